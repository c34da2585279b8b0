"""Farneback: oracle pins (scipy reflect-101 blurs, exact polynomial
expansion of quadratics, recovery of a known shift) and GPU parity."""
import numpy as np
import pytest
from scipy.ndimage import correlate1d

from oracle import oracle as orc


def consts(windowSize=33, Niters=5, polyN=7, polySigma=1.5, useGaussian=True):
    from Farneback_PyCL import Farneback_PyCL
    fb = Farneback_PyCL(windowSize, Niters, polyN, polySigma, useGaussian)
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


def level_kernel(n, sigma):
    """Half kernel of the level pre-blur: setGaussianBlurKernel(n, sigma)
    (Farneback_PyCL.py:200-208) on the bit-exact Gaussian."""
    from GaussianKernelBitExact import getGaussianKernelBitExact
    _, k = getGaussianKernelBitExact(n, sigma)
    return np.float32(k)[n // 2:]


def test_bilinear_resize_matches_pillow(golden):
    from PIL import Image
    g = golden("pil")
    for t, (H, W, H2, W2) in enumerate(g["cases"]):
        assert np.array_equal(orc.resize(g["in%d" % t], int(H2), int(W2), 2), g["bilinear%d" % t])
    rng = np.random.default_rng(5)
    for (H, W, H2, W2) in [(97, 70, 49, 35), (192, 224, 48, 56), (40, 52, 80, 104), (33, 65, 33, 65)]:
        img = (rng.random((H, W)) * 255).astype(np.float32)
        want = np.array(Image.fromarray(img).resize((W2, H2), Image.BILINEAR))
        assert np.array_equal(orc.resize(img, H2, W2, 2), want), (H, W, H2, W2)


def test_farneback_level_count_rule():
    # both sides must stay >= 32 after each pyr_scale step, at most levels-1 steps
    assert orc.farneback_levels(192, 224, 3, 0.5) == 2
    assert orc.farneback_levels(1024, 1024, 6, 0.5) == 5
    assert orc.farneback_levels(1024, 1024, 9, 0.5) == 5
    assert orc.farneback_levels(63, 100, 3, 0.5) == 0
    assert orc.farneback_levels(64, 64, 3, 0.5) == 1
    assert orc.farneback_levels(500, 700, 1, 0.5) == 0


def test_farneback_pyramid_without_coarser_levels_is_one_level():
    c = consts()
    im1, im2 = _smooth_pair(64, 80, (0.9, 0.4), seed=4)
    z = np.zeros_like(im1)
    a = orc.farneback_pyramid(im1, im2, z, z, c, 0, 0.5, level_kernel)
    b = orc.farneback(im1, im2, z, z, c)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_farneback_pyramid_recovers_a_large_shift():
    # 14 px is far outside one level's reach; two halvings bring it to 3.5 px
    c = consts()
    im1, im2 = _smooth_pair(192, 224, (14.0, -9.0))
    z = np.zeros_like(im1)
    inner = (slice(24, -24), slice(24, -24))
    fx, fy = orc.farneback_pyramid(im1, im2, z, z, c, 0, 0.5, level_kernel)
    assert abs(np.median(fx[inner]) - 14.0) > 5
    fx, fy = orc.farneback_pyramid(im1, im2, z, z, c, 2, 0.5, level_kernel)
    assert abs(np.median(fx[inner]) - 14.0) < 0.05 and abs(np.median(fy[inner]) + 9.0) < 0.05


@pytest.mark.parametrize("kh", [1, 16, 40])
def test_box_window_is_replicate_uniform_filter(kh):
    from scipy.ndimage import uniform_filter
    rng = np.random.default_rng(kh)
    H, W = 37, 61
    planes = rng.random((5 * H, W)) * 100
    out = orc.fb_box(planes, kh, planes=5, prec="f64").reshape(5, H, W)
    for k, p in enumerate(planes.reshape(5, H, W)):
        np.testing.assert_allclose(out[k], uniform_filter(p, size=2 * kh + 1, mode="nearest"), rtol=1e-12)


def test_farneback_box_window_recovers_known_shift():
    c = consts(useGaussian=False)
    assert c["box_half"] == 16 and "win_kernel" not in c
    im1, im2 = _smooth_pair(96, 112, (1.3, -0.7))
    z = np.zeros_like(im1)
    fx, fy = orc.farneback(im1, im2, z, z, c)
    inner = (slice(24, -24), slice(24, -24))
    assert abs(np.median(fx[inner]) - 1.3) < 0.05 and abs(np.median(fy[inner]) + 0.7) < 0.05
    fg, _ = orc.farneback(im1, im2, z, z, consts())
    assert np.abs(fx - fg).max() > 1e-3                    # a different window, not the Gaussian path


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


@pytest.mark.gpu
@pytest.mark.parametrize("H,W,n,sigma", [(192, 224, 3, 0.0), (192, 224, 39, 7.5), (97, 70, 19, 3.5),
                                         (40, 36, 159, 31.5), (515, 1030, 33, 4.95), (20, 300, 33, 4.95),
                                         (97, 70, 21, 3.15), (13, 1027, 3, 0.0)])
def test_fb_level_blur_matches_oracle(ops, H, W, n, sigma):
    # (40, 36, 159): a half kernel (79) wider than the image, reflect-101
    # indexes wrap; 3, 21 and 33 taps take the register-window passes (1030
    # and 1027 columns: a partial 1024-wide segment, rows not a multiple of
    # 4 floats; 20 rows against a 33-tap window: wrapped rows)
    rng = np.random.default_rng(n)
    img = (rng.random((H, W)) * 255).astype(np.float32)
    k = level_kernel(n, sigma)
    assert np.array_equal(ops.fb_blur(img, k).download(), orc.fb_blur(img, k))


@pytest.mark.gpu
@pytest.mark.parametrize("H,W,levels,shift", [(192, 224, 3, (14.0, -9.0)), (150, 134, 4, (2.5, 1.5)),
                                              (256, 256, 5, (6.0, -3.0))])
def test_farneback_pyramid_dropin_matches_oracle(ops, H, W, levels, shift):
    from Farneback_PyCL import Farneback_PyCL
    im1, im2 = _smooth_pair(H, W, shift, seed=H + W)
    rng = np.random.default_rng(W)
    U0 = (rng.standard_normal((H, W)) * 0.2).astype(np.float32)
    V0 = (rng.standard_normal((H, W)) * 0.2).astype(np.float32)
    fb = Farneback_PyCL(pyramidalLevels=levels)
    U, V, err = fb.compute(im1, im2, U0, V0)
    n = orc.farneback_levels(H, W, levels, 0.5)
    c = consts()
    ox, oy = orc.farneback_pyramid(im1, im2, U0, V0, c, n, 0.5, level_kernel)
    dx, dy = orc.farneback_pyramid(im1, im2, U0, V0, c, n, 0.5, level_kernel, prec="f64")
    d32 = max(np.abs(U - ox).max(), np.abs(V - oy).max())
    d64 = max(np.abs(U - dx).max(), np.abs(V - dy).max())
    inner = (slice(24, -24), slice(24, -24))
    print("[farneback pyramid %dx%d, %d coarser levels] max|d| vs oracle-f32 %.3g px, vs f64 %.3g px, "
          "median flow %.3f %.3f" % (H, W, n, d32, d64, np.median(U[inner]), np.median(V[inner])))
    assert U.shape == (H, W) and err == 'Unknown'
    assert d32 <= 1e-4
    assert d64 <= 1e-3
    assert abs(np.median(U[inner]) - shift[0]) < 0.05 and abs(np.median(V[inner]) - shift[1]) < 0.05


@pytest.mark.gpu
@pytest.mark.parametrize("H,W,levels,win", [(96, 112, 1, 33), (192, 224, 3, 33), (97, 70, 2, 21)])
def test_farneback_box_window_dropin_matches_oracle(ops, H, W, levels, win):
    from Farneback_PyCL import Farneback_PyCL
    shift = (1.3, -0.7) if levels == 1 else (6.0, 2.5)
    im1, im2 = _smooth_pair(H, W, shift, seed=H)
    z = np.zeros_like(im1)
    fb = Farneback_PyCL(windowSize=win, useGaussian=False, pyramidalLevels=levels)
    U, V, _ = fb.compute(im1, im2, z, z)
    n = orc.farneback_levels(H, W, levels, 0.5)
    c = consts(windowSize=win, useGaussian=False)
    ox, oy = orc.farneback_pyramid(im1, im2, z, z, c, n, 0.5, level_kernel)
    dx, dy = orc.farneback_pyramid(im1, im2, z, z, c, n, 0.5, level_kernel, prec="f64")
    d32 = max(np.abs(U - ox).max(), np.abs(V - oy).max())
    d64 = max(np.abs(U - dx).max(), np.abs(V - dy).max())
    print("[farneback box %dx%d, window %d, %d coarser levels] max|d| vs oracle-f32 %.3g px, vs f64 %.3g px"
          % (H, W, win, n, d32, d64))
    assert d32 <= 1e-4
    assert d64 <= 1e-3
