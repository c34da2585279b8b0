"""Drop-in for the reference's Farneback_PyCL.py: Farneback dense optical flow
on the GPU (libofri ofri_farneback) instead of PyOpenCL.

Reference: Farneback_PyCL.py:63-615 and optical_flow_farneback.cl.  The
constructor keeps the reference's signature (platformID / deviceID are
accepted and ignored: the context's GPU is used).  The expansion constants
(FarnebackPrepareGaussian) and the blur kernels (getGaussianKernelBitExact)
are formed on the host with the reference's dtype sequence; the device runs
pre-blur, polynomial expansion and the flow iterations with all buffers
resident, without the reference's per-kernel host copies.  The internal
pyramid (pyramidalLevels > 1, Farneback_PyCL.py:475-596) keeps the
reference's level loop on the host; per level the frames are blurred at full
resolution and Pillow-BILINEAR resized on the device (ofri_fb_blur,
ofri_resize), and the flow is resized and scaled there too.
"""
import numpy as np

from GaussianKernelBitExact import getGaussianKernelBitExact
from ofri import ops
from ofri.device import DeviceArray

FLT_EPSILON = 1.19209289550781250000000000000000000e-7


class Farneback_PyCL(object):
    """Farneback adapter for genericPyramidalOpticalFlow (Farneback_PyCL.py:63-615)."""

    def __init__(self, windowSize=33, Niters=5, polyN=7, polySigma=1.5, useGaussian=True, pyrScale=0.5,
                 pyramidalLevels=1, platformID=0, deviceID=0, provideGenericPyramidalDefaults=True):
        self.useDouble = False
        self.windowSize = windowSize
        self.numIters = Niters
        self.polyN = int(polyN)
        self.polySigma = polySigma
        self.useGaussianFilter = useGaussian
        self.pyramidalLevels = pyramidalLevels - 1
        self.fastPyramids = False
        self.pyrScale = pyrScale
        self.provideGenericPyramidalDefaults = provideGenericPyramidalDefaults
        if windowSize & 1 == 0:
            raise Exception('windowSize must be an odd value')
        ops.log('Runnning on device: libofri (gfx950)')

    # ---- host-side constants (Farneback_PyCL.py:123-208) -----------------
    def FarnebackPrepareGaussian(self):
        """Expansion taps g, xg, xxg (float32, 2n+1) and the four inverse Gram
        entries (float64)."""
        n = self.polyN
        sigma = self.polySigma
        if sigma < FLT_EPSILON:
            sigma = n * 0.3
        g = np.zeros(2 * n + 1, dtype=np.float32)
        xg = np.zeros(2 * n + 1, dtype=np.float32)
        xxg = np.zeros(2 * n + 1, dtype=np.float32)
        total = np.float64(0.0)
        for x in range(-n, n + 1):
            g[x + n] = np.exp(-x * x / (2 * sigma * sigma))
            total += g[x + n]
        total = 1.0 / total
        for x in range(-n, n + 1):
            g[x + n] = np.float32(g[x + n] * total)
            xg[x + n] = np.float32(x * g[x + n])
            xxg[x + n] = np.float32(x * x * g[x + n])
        G = np.zeros((6, 6), np.float64)
        for y in range(-n, n + 1):
            gy = g[y + n]
            for x in range(-n, n + 1):
                gx = g[x + n]
                G[0, 0] += gy * gx
                G[1, 1] += gy * gx * x * x
                G[3, 3] += gy * gx * x * x * x * x
                G[5, 5] += gy * gx * x * x * y * y
        G[2, 2] = G[0, 3] = G[0, 4] = G[3, 0] = G[4, 0] = G[1, 1]
        G[4, 4] = G[3, 3]
        G[3, 4] = G[4, 3] = G[5, 5]
        invG = np.linalg.inv(G)
        return g, xg, xxg, invG[1, 1], invG[0, 3], invG[3, 3], invG[5, 5]

    def setPolynomialExpansionConsts(self):
        n = self.polyN
        g, xg, xxg, i11, i03, i33, i55 = self.FarnebackPrepareGaussian()
        self.matrixG = g[n:].reshape([1, n + 1]).copy()
        self.matrixXG = xg[n:].reshape([1, n + 1]).copy()
        self.matrixXXG = xxg[n:].reshape([1, n + 1]).copy()
        self.matrixIGD = np.array([i11, i03, i33, i55], dtype=np.float64)
        self.matrixIG = self.matrixIGD.astype(np.float32)

    def getGaussianKernel(self, n, sigma):
        _, kernel = getGaussianKernelBitExact(n, sigma)
        return np.float32(kernel).reshape([1, n])

    def setGaussianBlurKernel(self, smoothSize, sigma):
        g = self.getGaussianKernel(smoothSize, sigma)
        half = int(smoothSize / 2)
        self.matrixGKernel = g[0, half:].reshape([1, half + 1]).copy()

    def _consts(self):
        self.setPolynomialExpansionConsts()
        self.setGaussianBlurKernel(3, 0.0)                 # smoothSize 3, sigma 0 at scale 1
        pre = self.matrixGKernel[0].copy()
        c = {"g": self.matrixG[0], "xg": self.matrixXG[0], "xxg": self.matrixXXG[0], "ig": self.matrixIG,
             "pre_kernel": pre, "iters": int(self.numIters)}
        if self.useGaussianFilter:
            self.setGaussianBlurKernel(self.windowSize, self.windowSize / 2 * 0.3)
            c["win_kernel"] = self.matrixGKernel[0].copy()
        else:
            c["box_half"] = self.windowSize // 2            # boxFilter5(M, blockSize/2) truncated to int
        return c

    # ---- adapter protocol -------------------------------------------------
    def compute(self, im1, im2, U, V):
        assert self.polyN == 5 or self.polyN == 7
        assert np.shape(im1) == np.shape(im2) and self.pyrScale < 1
        assert np.shape(U) == np.shape(im1) and np.shape(V) == np.shape(im1)
        size = np.shape(im1)
        min_size = 32
        levels, scale = 0, 1.0
        while levels < self.pyramidalLevels:
            scale *= self.pyrScale
            if size[1] * scale < min_size or size[0] * scale < min_size:
                break
            levels += 1
        c = self._consts()
        a1 = np.ascontiguousarray(im1, dtype=np.float32)
        a2 = np.ascontiguousarray(im2, dtype=np.float32)
        fx = DeviceArray.from_host(np.ascontiguousarray(U, dtype=np.float32))
        fy = DeviceArray.from_host(np.ascontiguousarray(V, dtype=np.float32))
        if levels == 0:
            # one level: the resizes are identities, the pre-blur (3, 0) runs on the device
            ops.farneback(a1, a2, fx, fy, c)
            return fx.download(), fy.download(), 'Unknown'
        d1, d2 = DeviceArray.from_host(a1), DeviceArray.from_host(a2)
        blurred = DeviceArray(d1.shape)
        ident = dict(c, pre_kernel=np.ones(1, np.float32))
        first = True
        for k in range(levels, -1, -1):                    # Farneback_PyCL.py:507-596
            scale = 1.0
            for _ in range(k):
                scale *= self.pyrScale
            sigma = (1.0 / scale - 1.0) * 0.5
            smoothSize = max(int(round(sigma * 5)) | 1, 3)
            width, height = int(round(size[1] * scale)), int(round(size[0] * scale))
            fx = ops.resize(fx, height, width, ops.BILINEAR)
            fy = ops.resize(fy, height, width, ops.BILINEAR)
            self.setGaussianBlurKernel(smoothSize, sigma)
            pk = self.matrixGKernel[0]
            f1 = ops.resize(ops.fb_blur(d1, pk, out=blurred), height, width, ops.BILINEAR)
            f2 = ops.resize(ops.fb_blur(d2, pk, out=blurred), height, width, ops.BILINEAR)
            ops.farneback(f1, f2, fx, fy, ident, flow_scale=scale if first else 1.0 / self.pyrScale)
            first = False
        return fx.download(), fy.download(), 'Unknown'

    def getAlgoName(self):
        return 'Farneback CL'

    def hasGenericPyramidalDefaults(self):
        return self.provideGenericPyramidalDefaults

    def getGenericPyramidalDefaults(self):
        return {'warping': False, 'scaling': True}
