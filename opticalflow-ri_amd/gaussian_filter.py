"""Drop-in for the reference's gaussian_filter.py: the separable Gaussian
pre-filter, computed on the GPU (libofri ofri_gauss_sep).

Reference: gaussian_filter.py:47-94.  The weights are formed on the host with
the reference's numpy dtype sequence (float64 Gaussian values stored into a
float32 kernel, normalised by its float32 sum), so they are bit-identical; the
filtering itself pads each line as the reference does and accumulates the
taps in float32 in the reference's order.
"""
import numpy as np

from ofri import ops


def prepareGaussianKernel(sigma, kernelSizePx):
    """float32 Gaussian taps at integer offsets arange(-n/2, n/2) (gaussian_filter.py:47-52)."""
    offsets = np.arange(-kernelSizePx / 2, kernelSizePx / 2, 1, dtype=int)
    peak = 1.0 / np.sqrt(2.0 * np.pi * sigma ** 2)
    taps = np.zeros(kernelSizePx, dtype=np.float32)
    taps[:] = peak * np.exp(-offsets ** 2 / (2.0 * sigma ** 2))
    return taps / np.sum(taps)


def convolveSeparableFilter(kernel, image):
    """Filters `image` in place with the float32 `kernel` (rows, then columns)
    and returns it (gaussian_filter.py:54-85)."""
    k = np.ascontiguousarray(kernel, dtype=np.float32)
    src = np.ascontiguousarray(image, dtype=np.float32)
    out = ops.gauss_sep(src, k).download()
    image[...] = out
    return image


def gaussian_filter(image, sigma, truncate):
    """Kernel of 2*int(truncate*sigma + 0.5) + 1 taps (gaussian_filter.py:87-90)."""
    kernelSizePx = 2 * int(truncate * sigma + 0.5) + 1
    return convolveSeparableFilter(prepareGaussianKernel(sigma, kernelSizePx), image)


def gaussian_filterPx(image, sigma, kernelSizePx):
    """Kernel of kernelSizePx taps; filters in place (gaussian_filter.py:92-94)."""
    return convolveSeparableFilter(prepareGaussianKernel(sigma, kernelSizePx), image)
