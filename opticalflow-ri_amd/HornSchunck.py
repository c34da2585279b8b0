"""Drop-in for the reference's HornSchunck.py: Horn-Schunck optical flow on
the GPU (libofri ofri_hs).

Reference: HornSchunck.py:28-126.  HS() keeps the reference's signature,
including its parameter naming HS(im2, im1, ...) where the first argument is
the frame at t and the second the frame at t+1; derivatives are the 2x2-cube
estimates with mirror padding, followed by Niter Jacobi sweeps of the weighted
3x3 average and the closed-form update.  The returned error is the change
from the initial flow, (||U-U0||_F + ||V-V0||_F) / (H*W), as the reference
returns it (there is no convergence stop).
"""
import numpy as np

from ofri import ops
from ofri.device import DeviceArray, as_device


class HSOpticalFlowAlgoAdapter(object):
    """Main-adapter for genericPyramidalOpticalFlow (HornSchunck.py:28-49).

    `alphas` is consumed from its end, one value per compute() call."""

    def __init__(self, alphas, Niter, provideGenericPyramidalDefaults=True):
        self.provideGenericPyramidalDefaults = provideGenericPyramidalDefaults
        self.alphas = alphas
        self.Niter = Niter

    def compute(self, im1, im2, U, V):
        alpha = self.alphas.pop()
        return HS(im1, im2, alpha, self.Niter, U, V)

    def getAlgoName(self):
        return 'Horn-Schunck'

    def hasGenericPyramidalDefaults(self):
        return self.provideGenericPyramidalDefaults

    def getGenericPyramidalDefaults(self):
        return {'warping': True, 'biLinear': True, 'scaling': True}


def HS(im2, im1, alpha, Niter, U, V):
    """Horn-Schunck from the initial flow (U, V).  The first image argument is
    the frame at t, the second the frame at t+1 (HornSchunck.py:72-104).
    Returns (U, V, total_error) as new float32 arrays."""
    d0, d1 = as_device(im2), as_device(im1)
    if d0.shape != d1.shape:
        raise ValueError("images differ in shape: %s vs %s" % (d0.shape, d1.shape))
    dU = DeviceArray.from_host(np.float32(U))
    dV = DeviceArray.from_host(np.float32(V))
    err = ops.hs(d0, d1, np.float32(alpha), int(Niter), dU, dV)
    return dU.download(), dV.download(), err
