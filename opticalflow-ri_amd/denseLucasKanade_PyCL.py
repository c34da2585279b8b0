"""Drop-in for the reference's denseLucasKanade_PyCL.py: dense pyramidal
Lucas-Kanade on the GPU (libofri ofri_lk_dense) instead of PyOpenCL.

Reference: denseLucasKanade_PyCL.py:33-182 and pyrlkDenseLargeW.cl.  The
constructor keeps the reference's signature (platformID / deviceID accepted
and ignored).  compute() returns (U, V, True) like the reference; the
per-pixel tracking status is kept in self.status (1 tracked, 0 lost).
"""
import numpy as np

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
        """Asymmetric window selection (denseLucasKanade_PyCL.py:75-92); only
        the symmetric window is implemented."""
        if not self.enableVorticityEnhancement:
            return [0, 0, 0, 0]
        raise NotImplementedError("vorticity-driven asymmetric windows are not implemented")

    def calcPatchSize(self, windowWidth, windowHeight):
        """The reference's patch-tiling helper (denseLucasKanade_PyCL.py:94-111)."""
        if windowWidth > 32 and windowWidth > 2 * windowHeight:
            blockX, blockY = 32, 8
        else:
            blockX, blockY = 16, 16
        return np.int32((windowWidth + blockX - 1) / blockX), np.int32((windowHeight + blockY - 1) / blockY)

    def compute(self, im1, im2, U, V):
        self.evaluateVorticityEnhancement(U, V)
        I = np.ascontiguousarray(im1, dtype=np.float32)
        J = np.ascontiguousarray(im2, dtype=np.float32)
        u = DeviceArray.from_host(np.ascontiguousarray(U, dtype=np.float32))
        v = DeviceArray.from_host(np.ascontiguousarray(V, dtype=np.float32))
        st = ops.ByteArray(I.shape)
        ops.lk_dense(I, J, u, v, self.windowHalfWidth, self.Niter, st)
        self.status = st.download()
        return u.download(), v.download(), True

    def getAlgoName(self):
        return 'OpenCL Dense LK'

    def hasGenericPyramidalDefaults(self):
        return self.provideGenericPyramidalDefaults

    def getGenericPyramidalDefaults(self):
        return {'warping': False, 'intermediateScaling': True, 'scaling': False}
