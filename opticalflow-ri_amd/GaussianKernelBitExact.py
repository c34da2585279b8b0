"""Drop-in for the reference's GaussianKernelBitExact.py (OpenCV's
getGaussianKernelBitExact restated with the stdlib decimal module).

Reference: GaussianKernelBitExact.py:55-144.  Host-side constant generation:
the weights feed the Farneback blur kernels.  sigma <= 0 with n in
{1, 3, 5, 7, 9} gives the fixed binomial kernels; otherwise sigmaX = sigma only
when sigma < 0, else 0.15*n + 0.35 (a sigma >= 0 is ignored, as in the
reference), the taps exp(-(x/2)^2 / (2 sigmaX^2)) for x = 1-n, 3-n, ... are
summed and normalised in 28-digit decimal arithmetic and returned as float64.
"""
import decimal

import numpy as np

softOne = decimal.Decimal("1.0")
softZero = decimal.Decimal("0.0")

# binomial kernels for sigma <= 0, in units of 1/2^k
_FIXED = {
    3: ([1, 2, 1], 4),
    5: ([1, 4, 6, 4, 1], 16),
    7: ([2, 7, 14, 18, 14, 7, 2], 64),
    9: ([4, 13, 30, 51, 60, 51, 30, 13, 4], 256),
}


def getGaussianKernelBitExact(n, sigma):
    """Returns (sum, kernel[n] float64) like the reference."""
    assert n > 0
    if sigma <= 0:
        if n == 1:
            return softOne, np.array([softOne])
        if n in _FIXED:
            num, den = _FIXED[n]
            return softOne, np.array(num, dtype=np.float64) / den
    D = decimal.Decimal
    sigmaX = D(sigma) if sigma < 0 else D(n) * D("0.15") + D("0.35")
    scale2X = D("-0.125") / (sigmaX * sigmaX)
    half = int((n - 1) / 2)
    taps = []
    total = softZero
    x = 1 - n
    for _ in range(half):
        t = (D(x * x) * scale2X).exp()
        taps.append(t)
        total += t
        x += 2
    total *= D(2.0)
    total += softOne
    if (n & 1) == 0:
        total += softOne
    mul1 = softOne / total
    result = [softZero] * n
    sum2 = softZero
    for i in range(half):
        t = taps[i] * mul1
        result[i] = t
        result[n - 1 - i] = t
        sum2 += t
    sum2 *= D(2.0)
    result[half] = softOne * mul1
    sum2 += result[half]
    if (n & 1) == 0:
        result[half + 1] = result[half]
        sum2 += result[half]
    return np.float64(sum2), np.array([float(r) for r in result], dtype=np.float64)
