"""Learned form infrastructure: the device particle renderer and the
pyramid's exported per-level labels."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ops():
    from ofri import ops as _ops
    from ofri.device import context
    context()
    return _ops


def _lshs(ops, levels=2):
    from gaussian_filter import prepareGaussianKernel
    return dict(levels=levels, alphas=[45.0, 21.0][-levels:] if levels == 1 else [45.0, 21.0], niter=600,
                filt_w=prepareGaussianKernel(3.4, 3), opt_w=prepareGaussianKernel(0.48, 5), ls_alpha=5.0)


def test_render_is_a_function_of_the_seed(ops):
    a1, a2 = ops.render_pair(192, 256, 1, 0.05, 77, [128.0, 96.0, 20.0, 3.0])
    b1, b2 = ops.render_pair(192, 256, 1, 0.05, 77, [128.0, 96.0, 20.0, 3.0])
    c1, _ = ops.render_pair(192, 256, 1, 0.05, 78, [128.0, 96.0, 20.0, 3.0])
    x1, x2, y1 = a1.download(), a2.download(), c1.download()
    assert np.array_equal(x1, b1.download()) and np.array_equal(x2, b2.download())
    assert not np.array_equal(x1, y1)
    # particle images: bright spots on a dark background, noise sigma 2
    assert 15 < x1.mean() < 120 and x1.max() > 200 and np.percentile(x1, 10) < 8


@pytest.mark.parametrize("family", [0, 1, 2])
def test_rendered_pairs_carry_their_flow(ops, family):
    """The plain pyramid recovers the rendered flow field (interior, mean EPE)."""
    H = W = 256
    rng = np.random.default_rng(family)
    modes = None
    if family == 0:
        par = [3.0]
    elif family == 1:
        par = [128.0, 128.0, 25.0, 3.0]
    else:
        n = 64
        # large-scale modes only (wavelengths 64..128 px): the smoothness
        # terms of HS / Liu-Shen cannot follow finer turbulence
        k = np.exp(rng.uniform(np.log(2 * np.pi / 128), np.log(2 * np.pi / 64), n))
        ang, phi = rng.uniform(0, 2 * np.pi, n), rng.uniform(0, 2 * np.pi, n)
        modes = np.stack([k * np.cos(ang), k * np.sin(ang), phi, k ** (-5.0 / 6.0) * np.sqrt(k) / k])
        par = [2.0, 3.0]
    im1, im2, u, v = ops.render_pair(H, W, family, 0.06, 1000 + family, par, modes, truth=True)
    U, V, _ = ops.pyr_lshs(im1, im2, **_lshs(ops))
    ut, vt = u.download(), v.download()
    assert np.sqrt(ut ** 2 + vt ** 2).max() == pytest.approx(3.0, rel=0.05)
    e = np.sqrt((U.download() - ut) ** 2 + (V.download() - vt) ** 2)[24:-24, 24:-24]
    print("[render family %d] mean EPE to truth %.3f px" % (family, e.mean()))
    assert e.mean() < 0.25


def test_taps_run_equals_the_plain_pyramid(ops):
    im1, im2 = ops.render_pair(256, 320, 1, 0.05, 5, [160.0, 128.0, 20.0, 3.0])
    prm = _lshs(ops)
    U, V, st = ops.pyr_lshs(im1, im2, **prm)
    Ut, Vt, taps, stt = ops.pyr_lshs_taps(im1, im2, **prm)
    assert np.array_equal(U.download(), Ut.download()) and np.array_equal(V.download(), Vt.download())
    for a, b in zip(st, stt):
        assert a["ls_iters"] == b["ls_iters"] and a["ls_err"] == b["ls_err"] and a["hs_err"] == b["hs_err"]
        assert b["hs_iters"] == 600 and 0 < b["hs_res"] < 1e-6
    # one level: the accumulated flow is HS + (LS - HS)
    U1, V1, taps1, _ = ops.pyr_lshs_taps(im1, im2, **_lshs(ops, 1))
    t = taps1.download().reshape(4, 256, 320)
    assert np.abs(t[0] + t[2] - U1.download()).max() <= 1e-6 * max(1.0, np.abs(t[0]).max())
    assert np.abs(t[1] + t[3] - V1.download()).max() <= 1e-6 * max(1.0, np.abs(t[1]).max())
