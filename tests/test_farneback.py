"""Farneback: oracle pins (scipy reflect-101 blurs, exact polynomial
expansion of quadratics, recovery of a known shift) and GPU parity."""
import numpy as np
import pytest
from scipy.ndimage import correlate1d

from oracle import oracle as orc


def consts(windowSize=33, Niters=5, polyN=7, polySigma=1.5):
    from Farneback_PyCL import Farneback_PyCL
    fb = Farneback_PyCL(windowSize, Niters, polyN, polySigma)
    return fb._consts()


def full_kernel(half):
    half = np.asarray(half, np.float64)
    return np.concatenate([half[:0:-1], half])


@pytest.mark.parametrize("key", ["pre_kernel", "win_kernel"])
def test_blur_is_reflect101_separable_gaussian(key):
    c = consts()
    rng = np.random.default_rng(1)
    img = rng.random((40, 57)) * 100
    out = orc.fb_blur(img, c[key], prec="f64")
    k = full_kernel(c[key])
    ref = correlate1d(correlate1d(img, k, axis=0, mode="mirror"), k, axis=1, mode="mirror")
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-10)


def test_window_kernel_is_bitexact_sigma_5_3(golden):
    c = consts()
    g = golden("bitexact")["gen_n33_s5.3"].astype(np.float32)
    assert np.array_equal(c["win_kernel"], g[16:])
    assert np.array_equal(c["pre_kernel"], np.float32([0.5, 0.25]))


@pytest.mark.parametrize("n,sigma", [(7, 1.5), (5, 1.1)])
def test_polynomial_expansion_recovers_quadratics(n, sigma):
    c = consts(polyN=n, polySigma=sigma)
    H, W = 48, 52
    y, x = np.mgrid[0:H, 0:W].astype(np.float64)
    # the expansion is local: planes = (df/dy, df/dx, f_yy/2, f_xx/2, f_xy) at each pixel
    a, b, cy2, cx2, cxy = 3.0, -2.0, 0.5, 0.25, 0.3
    img = 7.0 + a * y + b * x + cy2 * y * y + cx2 * x * x + cxy * x * y
    want = [a + 2 * cy2 * y + cxy * x, b + 2 * cx2 * x + cxy * y, cy2 + 0 * x, cx2 + 0 * x, cxy + 0 * x]
    R = orc.fb_polyexp(img, c["g"], c["xg"], c["xxg"], c["ig"], prec="f64").reshape(5, H, W)
    inner = (slice(n + 1, -n - 1), slice(n + 1, -n - 1))
    for plane in range(5):
        np.testing.assert_allclose(R[plane][inner], want[plane][inner], rtol=2e-5, atol=2e-5)
    # a constant image has no polynomial part
    R = orc.fb_polyexp(np.full((H, W), 9.0), c["g"], c["xg"], c["xxg"], c["ig"], prec="f64")
    assert np.abs(R).max() < 1e-6


def _smooth_pair(H, W, shift, seed=3):
    rng = np.random.default_rng(seed)
    base = rng.random((H + 40, W + 40))
    k = np.exp(-np.arange(-8, 9) ** 2 / (2 * 3.0 ** 2))
    k /= k.sum()
    base = correlate1d(correlate1d(base, k, axis=0, mode="wrap"), k, axis=1, mode="wrap")
    base = (base - base.mean()) / base.std() * 40 + 128
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64) + 20
    from scipy.ndimage import map_coordinates
    im1 = map_coordinates(base, [yy, xx], order=3)
    im2 = map_coordinates(base, [yy - shift[1], xx - shift[0]], order=3)
    return im1.astype(np.float32), im2.astype(np.float32)


def test_farneback_oracle_recovers_known_shift():
    c = consts()
    im1, im2 = _smooth_pair(96, 112, (1.3, -0.7))
    z = np.zeros_like(im1)
    fx, fy = orc.farneback(im1, im2, z, z, c)
    inner = (slice(24, -24), slice(24, -24))
    assert abs(np.median(fx[inner]) - 1.3) < 0.05 and abs(np.median(fy[inner]) + 0.7) < 0.05
    fx64, fy64 = orc.farneback(im1, im2, z, z, c, prec="f64")
    assert np.abs(fx - fx64).max() < 1e-3 and np.abs(fy - fy64).max() < 1e-3


# ------------------------------------------------------------------ GPU
@pytest.fixture(scope="module")
def ops():
    from ofri import ops as _ops
    from ofri.device import context
    context()
    return _ops


@pytest.mark.gpu
@pytest.mark.parametrize("H,W", [(128, 160), (512, 512), (97, 70)])
def test_farneback_gpu_matches_oracle(ops, H, W):
    from ofri.device import DeviceArray
    c = consts()
    im1, im2 = _smooth_pair(H, W, (1.1, 0.4), seed=H)
    rng = np.random.default_rng(W)
    U0 = (rng.standard_normal((H, W)) * 0.2).astype(np.float32)
    V0 = (rng.standard_normal((H, W)) * 0.2).astype(np.float32)
    fx, fy = DeviceArray.from_host(U0), DeviceArray.from_host(V0)
    ops.farneback(im1, im2, fx, fy, c)
    gx, gy = fx.download(), fy.download()
    ox, oy = orc.farneback(im1, im2, U0, V0, c)
    dx, dy = orc.farneback(im1, im2, U0, V0, c, prec="f64")
    d32 = max(np.abs(gx - ox).max(), np.abs(gy - oy).max())
    d64 = max(np.abs(gx - dx).max(), np.abs(gy - dy).max())
    print("[farneback %dx%d] max|d| vs oracle-f32 %.3g px, vs f64 %.3g px" % (H, W, d32, d64))
    assert d32 <= 1e-4
    assert d64 <= 1e-3


@pytest.mark.gpu
def test_farneback_dropin_adapter(ops):
    from Farneback_PyCL import Farneback_PyCL
    im1, im2 = _smooth_pair(80, 96, (0.8, 0.3), seed=9)
    z = np.zeros_like(im1)
    fb = Farneback_PyCL(platformID=0)
    U, V, err = fb.compute(im1, im2, z, z)
    assert err == 'Unknown'
    ox, oy = orc.farneback(im1, im2, z, z, consts())
    assert np.abs(U - ox).max() <= 1e-4 and np.abs(V - oy).max() <= 1e-4
    assert fb.getGenericPyramidalDefaults() == {'warping': False, 'scaling': True}
    with pytest.raises(Exception):
        Farneback_PyCL(windowSize=32)
