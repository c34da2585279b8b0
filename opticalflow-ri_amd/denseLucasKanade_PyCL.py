"""Drop-in for the reference's denseLucasKanade_PyCL.py: dense pyramidal
Lucas-Kanade on the GPU (libofri ofri_lk_dense) instead of PyOpenCL.

Reference: denseLucasKanade_PyCL.py:33-182 and pyrlkDenseLargeW.cl.  The
constructor keeps the reference's signature (platformID / deviceID accepted
and ignored).  compute() returns (U, V, True) like the reference; the
per-pixel tracking status is kept in self.status (1 tracked, 0 lost).  With
enableVorticityEnhancement the window flags come from the mean vorticity of
the incoming flow, computed on the host as the reference computes it.
"""
import numpy as np
from scipy.ndimage import convolve as filter2

from ofri import ops
from ofri.device import DeviceArray


class denseLucasKanade_PyCl(object):
    """Dense LK adapter for genericPyramidalOpticalFlow (denseLucasKanade_PyCL.py:33-182)."""

    def __init__(self, platformID=0, deviceID=0, Niter=5, halfWindow=13, provideGenericPyramidalDefaults=True,
                 enableVorticityEnhancement=False):
        self.provideGenericPyramidalDefaults = provideGenericPyramidalDefaults
        self.enableVorticityEnhancement = enableVorticityEnhancement
        self.windowHalfWidth = int(halfWindow)
        self.windowHalfHeight = int(halfWindow)
        self.windowWidth = 2 * int(halfWindow) + 1
        self.windowHeight = 2 * int(halfWindow) + 1
        self.Niter = Niter
        self.status = None
        ops.log('Runnning on device: libofri (gfx950)')

    def evaluateVorticityEnhancement(self, U, V):
        """Asymmetric window flags [left, right, top, bottom] from the mean of
        omega = convolve(V, D.T) - convolve(U, D) with the half central
        difference D, 'reflect' borders, beyond +-2e-3
        (denseLucasKanade_PyCL.py:75-92).  Convolution flips D, so omega is
        minus the vorticity dV/dx - dU/dy."""
        if not self.enableVorticityEnhancement:
            return [0, 0, 0, 0]
        D = np.array([[0, -1, 0], [0, 0, 0], [0, 1, 0]], dtype=np.float32) * np.float32(0.5)
        omega = filter2(V, D.T, mode='reflect') - filter2(U, D, mode='reflect')
        m = np.mean(omega)
        if m < -2e-3:
            return [0, 1, 0, 1]
        if m > 2e-3:
            return [1, 0, 0, 1]
        return [0, 0, 0, 0]

    def calcPatchSize(self, windowWidth, windowHeight):
        """The reference's patch-tiling helper (denseLucasKanade_PyCL.py:94-111)."""
        if windowWidth > 32 and windowWidth > 2 * windowHeight:
            blockX, blockY = 32, 8
        else:
            blockX, blockY = 16, 16
        return np.int32((windowWidth + blockX - 1) / blockX), np.int32((windowHeight + blockY - 1) / blockY)

    def compute(self, im1, im2, U, V):
        asym = self.evaluateVorticityEnhancement(U, V)
        I = np.ascontiguousarray(im1, dtype=np.float32)
        J = np.ascontiguousarray(im2, dtype=np.float32)
        u = DeviceArray.from_host(np.ascontiguousarray(U, dtype=np.float32))
        v = DeviceArray.from_host(np.ascontiguousarray(V, dtype=np.float32))
        st = ops.ByteArray(I.shape)
        ops.lk_dense(I, J, u, v, self.windowHalfWidth, self.Niter, st, asym)
        self.status = st.download()
        return u.download(), v.download(), True

    def getAlgoName(self):
        return 'OpenCL Dense LK'

    def hasGenericPyramidalDefaults(self):
        return self.provideGenericPyramidalDefaults

    def getGenericPyramidalDefaults(self):
        return {'warping': False, 'intermediateScaling': True, 'scaling': False}
