"""Drop-in for the reference's PhysicsBasedOpticalFlowLiuShen.py: the Liu-Shen
physics-based optical flow refinement on the GPU (libofri ofri_ls).

Reference: PhysicsBasedOpticalFlowLiuShen.py:33-158.  Each frame is normalised
by its own maximum, the per-pixel 2x2 inverse B and the coefficient fields are
formed once, then Jacobi sweeps run `while total_error > tol and k < maxnum`
(tol = 1e-8, maxnum = 60) with total_error = (||du||_2 + ||dv||_2) / (r*c).
Inside the solver u is the row (y) component; the adapter swaps U and V.
"""
import numpy as np

from ofri import ops
from ofri.device import DeviceArray, as_device

MAXNUM = 60
TOL = 1e-8


class LiuShenOpticalFlowAlgoAdapter(object):
    """Optional (refinement) adapter for genericPyramidalOpticalFlow
    (PhysicsBasedOpticalFlowLiuShen.py:33-45)."""

    def __init__(self, alpha):
        self.alpha = alpha

    def compute(self, im1, im2, U, V):
        [resV, resU, error] = physicsBasedOpticalFlowLiuShen(im1, im2, self.alpha, V, U)
        return [resU, resV, error]

    def getAlgoName(self):
        return 'Liu-Shen Physics based OF'

    # Declared without `self`, exactly as the reference declares it
    # (PhysicsBasedOpticalFlowLiuShen.py:44): calling it on an instance raises
    # TypeError, so the adapter cannot serve as a main adapter.
    def hasGenericPyramidalDefaults():
        return False


def physicsBasedOpticalFlowLiuShen(im1, im2, h, U, V, maxnum=MAXNUM, tol=TOL):
    """Liu-Shen refinement from the initial flow; U is the row component.
    Returns (u, v, error) as new float32 arrays and a float32 error."""
    d1, d2 = as_device(im1), as_device(im2)
    if d1.shape != d2.shape:
        raise ValueError("images differ in shape: %s vs %s" % (d1.shape, d2.shape))
    du = DeviceArray.from_host(np.float32(U))
    dv = DeviceArray.from_host(np.float32(V))
    err, k = ops.ls(d1, d2, h, du, dv, maxnum=maxnum, tol=tol)
    if k:
        ops.log('Iteration: ' + str(k - 1) + ' - Total error: ' + str(err))
    return du.download(), dv.download(), err
