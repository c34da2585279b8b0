"""GPU parity: libofri (through its C ABI) against the CPU oracle on the same
seeded inputs.  Bar: bit-exact for weights, sampling indices and the
deterministic stencil pipelines; flow max |du| <= 1e-3 px and mean-EPE
difference <= 1e-4 px against the oracle's reference dataflow, with the
FP32-vs-FP64 error printed for every solver result."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

FLOW_TOL = 1e-3      # max |delta u| px, against the oracle's reference dataflow
EPE_TOL = 1e-4       # mean end-point-error difference, px


@pytest.fixture(scope="module")
def ops():
    from ofri import ops as _ops
    from ofri.device import context
    context()  # fails loudly if libofri.so or the GPU is missing
    return _ops


def _img(rng, H, W, lo=0.0, hi=255.0):
    return (rng.random((H, W)) * (hi - lo) + lo).astype(np.float32)


def _epe(u1, v1, u2, v2):
    return np.sqrt((u1 - u2) ** 2 + (v1 - v2) ** 2)


def report(name, U, V, U32, V32, U64, V64):
    d32 = max(np.abs(U - U32).max(), np.abs(V - V32).max())
    d64 = max(np.abs(U - U64).max(), np.abs(V - V64).max())
    e64 = _epe(U, V, U64, V64).mean()
    print("[%s] max|d| vs oracle-f32 %.3g px, vs oracle-f64 %.3g px (mean EPE to f64 %.3g px)" % (name, d32, d64, e64))
    return d32, d64


# ------------------------------------------------------------- pre-filter
@pytest.mark.parametrize("shape", [(256, 256), (37, 53), (5, 9)])
@pytest.mark.parametrize("sigma,n", [(3.4, 3), (0.48, 5), (2.0, 3), (1.8, 7)])
def test_prefilter_bitexact(ops, golden, shape, sigma, n):
    from gaussian_filter import prepareGaussianKernel
    if min(shape) < n:
        pytest.skip("image smaller than kernel")
    w = prepareGaussianKernel(sigma, n)
    rng = np.random.default_rng(1)
    img = _img(rng, *shape)
    out = ops.gauss_sep(img, w).download()
    assert np.array_equal(out, orc.gauss_sep(img, w))


def test_prefilter_dropin_mutates_and_returns(ops):
    from gaussian_filter import gaussian_filterPx, prepareGaussianKernel
    rng = np.random.default_rng(2)
    img = _img(rng, 31, 40)
    ref = orc.gauss_sep(img, prepareGaussianKernel(3.4, 3))
    out = gaussian_filterPx(img, 3.4, 3)
    assert out is img and np.array_equal(img, ref)


# ------------------------------------------------------------- resize / spline / warp
def test_resize_bicubic_matches_pillow_fixture(ops, golden):
    g = golden("pil")
    for t, (H, W, H2, W2) in enumerate(g["cases"]):
        out = ops.resize(g["in%d" % t], int(H2), int(W2), ops.BICUBIC).download()
        assert np.array_equal(out, g["bicubic%d" % t]), (H, W, H2, W2)
        out = ops.resize(g["in%d" % t], int(H2), int(W2), ops.BILINEAR).download()
        assert np.array_equal(out, g["bilinear%d" % t]), (H, W, H2, W2)


@pytest.mark.parametrize("H,W", [(1024, 1024), (2048, 768), (385, 257)])
def test_resize_half_matches_oracle(ops, H, W):
    rng = np.random.default_rng(3)
    img = _img(rng, H, W)
    H2, W2 = int(np.round(H * 0.5)), int(np.round(W * 0.5))
    assert np.array_equal(ops.resize(img, H2, W2).download(), orc.resize_bicubic(img, H2, W2))


def test_spline_matches_scipy_fixture(ops, golden):
    g = golden("spline")
    for t, (hc, wc, h, w) in enumerate(g["cases"]):
        out = ops.spline_upsample(g["in%d" % t], int(h), int(w), 1.0).download()
        np.testing.assert_allclose(out, g["out%d" % t].astype(np.float32), rtol=2e-7, atol=1e-7)


@pytest.mark.parametrize("hc,wc,h,w", [(512, 512, 1024, 1024), (96, 160, 192, 320), (100, 70, 201, 141)])
def test_spline_matches_oracle(ops, hc, wc, h, w):
    rng = np.random.default_rng(4)
    c = (rng.standard_normal((hc, wc)) * 3).astype(np.float32)
    sx = np.float32(w) / np.float32(wc)
    out = ops.spline_upsample(c, h, w, sx).download()
    ref = orc.spline_upsample(c, h, w, sx)
    # double sums in a different order: at most one float32 rounding step apart
    assert np.abs(out - ref).max() <= 4e-7 * max(1.0, np.abs(ref).max())
    assert (out != ref).mean() < 1e-3


def test_warp_matches_fixture_and_oracle(ops, golden):
    g = golden("warp")
    o1, o2 = ops.warp_symmetric(g["im1"], g["im2"], g["U"], g["V"])
    assert np.array_equal(o1.download(), g["o1"]) and np.array_equal(o2.download(), g["o2"])
    rng = np.random.default_rng(5)
    H, W = 300, 411
    a, b = _img(rng, H, W), _img(rng, H, W)
    U = (rng.standard_normal((H, W)) * 4).astype(np.float32)
    V = (rng.standard_normal((H, W)) * 4).astype(np.float32)
    r1, r2 = orc.warp_symmetric(a, b, U, V)
    o1, o2 = ops.warp_symmetric(a, b, U, V)
    assert np.array_equal(o1.download(), r1) and np.array_equal(o2.download(), r2)


def test_warp_coords_dropin(ops, golden):
    from GenericPyramidalOpticalFlow import doBiLinearWarping
    g = golden("warp")
    H, W = g["im1"].shape
    ys, xs = np.meshgrid(np.arange(H, dtype=np.int32), np.arange(W, dtype=np.int32), indexing="ij")
    out = doBiLinearWarping(g["im1"], np.float32(ys - g["V"] / 2.0), np.float32(xs - g["U"] / 2.0), 1, "nearest")
    assert np.array_equal(out, g["o1"])


# ------------------------------------------------------------- Horn-Schunck
def _piv(H, W, seed=7, shift=(1.3, 0.6)):
    from ofri import synth
    rng = np.random.default_rng(seed)
    im1, im2, _, _ = synth.piv_pair(H, W, 0.05, lambda x, y: (shift[0] + 0 * x, shift[1] + 0 * y), rng)
    return im1, im2


@pytest.mark.parametrize("H,W,niter", [(256, 256, 600), (97, 131, 50), (12, 10, 7), (300, 70, 33)])
def test_hs_matches_oracle(ops, H, W, niter):
    from ofri.device import DeviceArray
    im1, im2 = _piv(H, W)
    U = DeviceArray.zeros((H, W)); V = DeviceArray.zeros((H, W))
    err = ops.hs(im1, im2, 21.0, niter, U, V)
    Ug, Vg = U.download(), V.download()
    U32, V32, e32 = orc.hs(im1, im2, 21.0, niter)
    U64, V64, _ = orc.hs(im1, im2, 21.0, niter, prec="f64")
    d32, d64 = report("hs %dx%d/%d" % (H, W, niter), Ug, Vg, U32, V32, U64, V64)
    assert d32 <= 1e-5
    assert abs(_epe(Ug, Vg, U64, V64).mean() - _epe(U32, V32, U64, V64).mean()) <= EPE_TOL
    assert d64 <= FLOW_TOL
    assert abs(float(err) - e32) <= 1e-5 * abs(e32) + 1e-12


def test_hs_nonzero_initial_flow_and_small_golden(ops, golden):
    from ofri.device import DeviceArray
    g = golden("hs_small")
    U = DeviceArray.from_host(g["U0"]); V = DeviceArray.from_host(g["V0"])
    ops.hs(g["I0"], g["I1"], float(g["alpha"]), int(g["niter"]), U, V)
    assert np.abs(U.download() - g["U"]).max() <= 2e-6 * np.abs(g["U"]).max()
    assert np.abs(V.download() - g["V"]).max() <= 2e-6 * np.abs(g["V"]).max()


@pytest.mark.parametrize("depth", [1, 2, 5, 8, 12])
def test_hs_temporal_depth_does_not_change_result(ops, depth):
    from ofri.device import DeviceArray, context
    im1, im2 = _piv(200, 333, seed=11)
    ctx = context()
    res = []
    for d in (depth, 6):
        ctx.set_tuning(hs_depth=d)
        U = DeviceArray.zeros(im1.shape); V = DeviceArray.zeros(im1.shape)
        ops.hs(im1, im2, 45.0, 61, U, V)
        res.append((U.download(), V.download()))
    ctx.set_tuning()
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


@pytest.mark.parametrize("variant", range(12))
def test_hs_tile_variants_identical(ops, variant):
    from ofri.device import DeviceArray, context
    im1, im2 = _piv(150, 211, seed=13)
    ctx = context()
    U32, V32, _ = orc.hs(im1, im2, 21.0, 40)
    for depth in (3, 7):
        ctx.set_tuning(hs_depth=depth, hs_variant=variant + 1)
        U = DeviceArray.zeros(im1.shape); V = DeviceArray.zeros(im1.shape)
        try:
            ops.hs(im1, im2, 21.0, 40, U, V)
        finally:
            ctx.set_tuning()
        assert np.array_equal(U.download(), U32) and np.array_equal(V.download(), V32), (variant, depth)


@pytest.mark.parametrize("H,W", [(170, 230), (5, 7), (2, 66), (730, 1000)])
@pytest.mark.parametrize("variant", range(7))
def test_ls_tile_variants_identical(ops, variant, H, W):
    # (5, 7) and (2, 66): every pixel on the image border, most of each tile
    # outside the image (the coefficients formed in each launch's load phase
    # take the clamped neighbours there); (730, 1000): border tiles cut at the
    # right and bottom edges around many interior tiles, the image's first
    # and last rows at different rows of a wavefront across the shapes
    from ofri.device import DeviceArray, context
    im1, im2 = _piv(H, W, seed=14)
    u32, v32, e32, it32 = orc.ls(im1, im2, 5.0, np.zeros((H, W)), np.zeros((H, W)), maxnum=17, tol=0.0)
    ctx = context()
    for depth in (2, 5):
        ctx.set_tuning(ls_depth=depth, ls_variant=variant + 1)
        du, dv = DeviceArray.zeros((H, W)), DeviceArray.zeros((H, W))
        try:
            err, it = ops.ls(im1, im2, 5.0, du, dv, maxnum=17, tol=0.0)
        finally:
            ctx.set_tuning()
        assert it == 17
        assert np.array_equal(du.download(), u32) and np.array_equal(dv.download(), v32), (variant, depth)


def test_hs_dropin_matches_oracle(ops):
    from HornSchunck import HS, HSOpticalFlowAlgoAdapter
    im1, im2 = _piv(64, 80, seed=12)
    z = np.zeros(im1.shape, np.float32)
    U, V, err = HS(im1, im2, 21, 100, z, z)
    U32, V32, e32 = orc.hs(im1, im2, 21.0, 100)
    assert np.abs(U - U32).max() <= 1e-5 and np.abs(V - V32).max() <= 1e-5
    ad = HSOpticalFlowAlgoAdapter([21, 45], 100)
    U2, V2, _ = ad.compute(im1, im2, z, z)        # pops 45
    assert ad.alphas == [21]
    U45, _, _ = orc.hs(im1, im2, 45.0, 100)
    assert np.abs(U2 - U45).max() <= 1e-5


# ------------------------------------------------------------- Liu-Shen
@pytest.mark.parametrize("H,W", [(256, 256), (131, 97), (9, 14)])
def test_ls_matches_oracle(ops, H, W):
    from ofri.device import DeviceArray
    im1, im2 = _piv(H, W, seed=21)
    rng = np.random.default_rng(22)
    u0 = (rng.standard_normal((H, W)) * 0.3 + 0.6).astype(np.float32)
    v0 = (rng.standard_normal((H, W)) * 0.3 + 1.3).astype(np.float32)
    du, dv = DeviceArray.from_host(u0), DeviceArray.from_host(v0)
    err, it = ops.ls(im1, im2, 5.0, du, dv)
    u32, v32, e32, it32 = orc.ls(im1, im2, 5.0, u0, v0)
    u64, v64, _, _ = orc.ls(im1, im2, 5.0, u0, v0, prec="f64")
    ug, vg = du.download(), dv.download()
    d32, d64 = report("ls %dx%d" % (H, W), ug, vg, u32, v32, u64, v64)
    assert it == it32
    assert d32 <= 1e-5
    assert d64 <= FLOW_TOL
    assert abs(float(err) - e32) <= 1e-5 * abs(e32)


def test_ls_stops_inside_a_fused_launch(ops):
    """A tolerance between the 8th and 7th sweep's norm stops after sweep 8,
    which sits inside a fused launch: the re-run must give exactly 8 sweeps."""
    from ofri.device import DeviceArray
    H, W = 96, 120
    im1, im2 = _piv(H, W, seed=23)
    z = np.zeros((H, W), np.float32)
    f = orc.ls_setup(im1, im2, np.float32(5))
    u, v, errs = z, z, []
    for _ in range(10):
        u, v, e = orc.ls_sweep(f, 5.0, u, v)
        errs.append(float(np.float32(e)))
    assert errs[7] < errs[6]
    tol = 0.5 * (errs[7] + errs[6])
    u8, v8, e8, it8 = orc.ls(im1, im2, 5.0, z, z, maxnum=60, tol=tol)
    assert it8 == 8
    du, dv = DeviceArray.zeros((H, W)), DeviceArray.zeros((H, W))
    err, it = ops.ls(im1, im2, 5.0, du, dv, maxnum=60, tol=tol)
    assert it == 8
    assert np.abs(du.download() - u8).max() <= 1e-5 and np.abs(dv.download() - v8).max() <= 1e-5
    # maxnum smaller than the launch depth and zero
    du, dv = DeviceArray.zeros((H, W)), DeviceArray.zeros((H, W))
    err, it = ops.ls(im1, im2, 5.0, du, dv, maxnum=3, tol=0.0)
    u3, v3, _, _ = orc.ls(im1, im2, 5.0, z, z, maxnum=3, tol=0.0)
    assert it == 3 and np.abs(du.download() - u3).max() <= 1e-5
    du = DeviceArray.from_host(u3)
    err, it = ops.ls(im1, im2, 5.0, du, dv, maxnum=0)
    assert it == 0 and np.array_equal(du.download(), u3)


def _hs_norm_history(im1, im2, alpha, n):
    """Per-sweep (||dU||_2 + ||dV||_2)/(H*W) of the oracle's float32 sweeps."""
    fx, fy, ft = orc.hs_derivs(im1, im2)
    U = np.zeros(im1.shape, np.float32); V = np.zeros(im1.shape, np.float32)
    hist = []
    for _ in range(n):
        Un, Vn = orc.hs_sweep(fx, fy, ft, alpha, U, V)
        du = (Un - U).astype(np.float64); dv = (Vn - V).astype(np.float64)
        hist.append(np.float32((np.sqrt((du * du).sum()) + np.sqrt((dv * dv).sum())) / U.size))
        U, V = Un, Vn
    return np.array(hist, np.float32), U, V


@pytest.mark.parametrize("H,W,depth", [(160, 230, 0), (97, 131, 3), (64, 64, 8)])
def test_hs_residual_history_matches_oracle(ops, H, W, depth):
    """ofri_hs_tol with no tolerance: the per-sweep update norms recorded on the
    device equal the oracle's, and the flow equals ofri_hs's bit for bit."""
    from ofri.device import DeviceArray, context
    im1, im2 = _piv(H, W, seed=H)
    n = 23
    ref_hist, U32, V32 = _hs_norm_history(im1, im2, 21.0, n)
    Uo, Vo, _ = orc.hs(im1, im2, 21.0, n)
    assert np.array_equal(U32, Uo) and np.array_equal(V32, Vo)
    ctx = context()
    ctx.set_tuning(hs_depth=depth)
    try:
        U = DeviceArray.zeros((H, W)); V = DeviceArray.zeros((H, W))
        hist = DeviceArray.zeros((n,))
        err, it = ops.hs_tol(im1, im2, 21.0, n, -1.0, U, V, hist)
    finally:
        ctx.set_tuning()
    assert it == n
    assert np.array_equal(U.download(), U32) and np.array_equal(V.download(), V32)
    h = hist.download()
    assert np.allclose(h, ref_hist, rtol=1e-6, atol=0), np.abs(h / ref_hist - 1).max()
    assert err == h[-1]


def test_hs_tol_stops_inside_a_fused_launch(ops):
    """A tolerance equal to the 9th sweep's norm stops after the first sweep
    whose norm is at or below it; that sweep sits inside a fused launch and the
    result must equal ofri_hs run for exactly that many sweeps."""
    from ofri.device import DeviceArray, context
    H, W = 120, 150
    im1, im2 = _piv(H, W, seed=31)
    ref_hist, _, _ = _hs_norm_history(im1, im2, 45.0, 40)
    ctx = context()
    ctx.set_tuning(hs_depth=5)
    try:
        hist = DeviceArray.zeros((40,))
        U = DeviceArray.zeros((H, W)); V = DeviceArray.zeros((H, W))
        ops.hs_tol(im1, im2, 45.0, 40, -1.0, U, V, hist)
        h = hist.download()
        tol = float(h[8])
        stop = int(np.argmax(h <= np.float32(tol))) + 1
        assert stop % 5 != 0            # not at a launch boundary
        U = DeviceArray.zeros((H, W)); V = DeviceArray.zeros((H, W))
        err, it = ops.hs_tol(im1, im2, 45.0, 40, tol, U, V)
        assert it == stop and err == h[stop - 1]
        Ur = DeviceArray.zeros((H, W)); Vr = DeviceArray.zeros((H, W))
        ops.hs(im1, im2, 45.0, stop, Ur, Vr)
        assert np.array_equal(U.download(), Ur.download()) and np.array_equal(V.download(), Vr.download())
        # zero sweeps leaves the flow untouched
        U0 = DeviceArray.from_host(np.full((H, W), 0.5, np.float32))
        err, it = ops.hs_tol(im1, im2, 45.0, 0, 0.0, U0, V)
        assert it == 0 and np.all(U0.download() == 0.5)
    finally:
        ctx.set_tuning()


_fma = orc.fma32


def _accel_step(fx, fy, ft, alpha, U, V):
    """One accelerated-phase sweep from (U, V) as the device forms it:
    t = T / w from oracle.hs_sweep_accel_scaled, the plain step's value
    T = RN(w t) and the update fma(w, t, -x).  Returns (Tu, Tv, du, dv)."""
    tu, tv = orc.hs_sweep_accel_scaled(fx, fy, ft, alpha, U, V)
    w = orc.ACCEL_W
    return w * tu, w * tv, _fma(w, tu, -U), _fma(w, tv, -V)


def _cheb_restated(im1, im2, alpha, bound, tol, maxnum, finish=0.0):
    """Chebyshev semi-iteration over the accelerated phase's float32 sweep T
    (oracle.hs_sweep_accel), with the device's float32 operation order: x_{k+1} = x_{k-1} + w_{k+1} ((x_k +
    g (T(x_k) - x_k)) - x_{k-1}), each multiply-add fused (fma(g, d, x_k),
    then fma(w, t - x_{k-1}, x_{k-1})); stop at the first x_k with plain residual
    <= tol (finish * tol when finish > 1) and take T(x_k); with finish > 1 the
    accelerated phase has at most maxnum - maxnum // 2 sweeps and plain
    sweeps (the oracle's reference-dataflow sweep) then continue to tol on
    the same count.  Returns (U, V,
    sweeps, residual history)."""
    t1 = tol * finish if finish > 1 else tol
    lo, hi = -1.0 / 3.0, float(np.float32(bound))
    sig = (hi - lo) / (2.0 - lo - hi)
    g = np.float32(2.0 / (2.0 - lo - hi))
    om, w = [], 1.0
    for k in range(maxnum):
        w = 1.0 if k == 0 else (2.0 / (2.0 - sig * sig) if k == 1 else 1.0 / (1.0 - 0.25 * sig * sig * w))
        om.append(np.float32(w))
    fx, fy, ft = orc.hs_derivs(im1, im2)
    U = np.zeros(im1.shape, np.float32); V = np.zeros(im1.shape, np.float32)
    UP, VP = U.copy(), V.copy()
    hist = []
    n_acc = maxnum - max(1, maxnum // 2) if finish > 1 else maxnum
    for k in range(n_acc):
        Tu, Tv, du, dv = _accel_step(fx, fy, ft, alpha, U, V)
        r = np.float32((np.sqrt((du.astype(np.float64) ** 2).sum()) + np.sqrt((dv.astype(np.float64) ** 2).sum()))
                       / U.size)
        hist.append(r)
        if not r > t1:
            U, V = Tu, Tv
            break
        nu = _fma(om[k], _fma(g, du, U) - UP, UP)
        nv = _fma(om[k], _fma(g, dv, V) - VP, VP)
        UP, VP, U, V = U, V, nu, nv
    else:
        if not finish > 1:
            return U, V, n_acc, np.array(hist)
    if finish > 1:
        # the finishing phase follows whether or not the accelerated one met
        # its tolerance
        while len(hist) < maxnum and hist[-1] > tol:
            Tu, Tv = orc.hs_sweep(fx, fy, ft, alpha, U, V)
            du, dv = Tu - U, Tv - V
            hist.append(np.float32((np.sqrt((du.astype(np.float64) ** 2).sum()) +
                                    np.sqrt((dv.astype(np.float64) ** 2).sum())) / U.size))
            U, V = Tu, Tv
    return U, V, len(hist), np.array(hist)


def _cheb_restated_launches(im1, im2, alpha, bound, tol, maxnum, finish, depth):
    """_cheb_restated with the stop test only at the last sweep of each
    launch of `depth` sweeps (launches counted from each phase's start): the
    accelerated phase (at most maxnum - maxnum // 2 sweeps when finishing)
    ends with the iterate its stopping launch produced, the plain phase (at
    most maxnum // 2 sweeps) with T(x_k).  Returns (U, V,
    sweeps, last tested residual)."""
    t1 = tol * finish if finish > 1 else tol
    lo, hi = -1.0 / 3.0, float(np.float32(bound))
    sig = (hi - lo) / (2.0 - lo - hi)
    g = np.float32(2.0 / (2.0 - lo - hi))
    om, w = [], 1.0
    for k in range(maxnum):
        w = 1.0 if k == 0 else (2.0 / (2.0 - sig * sig) if k == 1 else 1.0 / (1.0 - 0.25 * sig * sig * w))
        om.append(np.float32(w))
    fx, fy, ft = orc.hs_derivs(im1, im2)
    U = np.zeros(im1.shape, np.float32); V = np.zeros(im1.shape, np.float32)
    UP, VP = U.copy(), V.copy()

    def res(du, dv):
        return np.float32((np.sqrt((du.astype(np.float64) ** 2).sum()) +
                           np.sqrt((dv.astype(np.float64) ** 2).sum())) / U.size)
    count, r = 0, None
    n_acc = maxnum - max(1, maxnum // 2) if finish > 1 else maxnum
    for k in range(n_acc):
        Tu, Tv, du, dv = _accel_step(fx, fy, ft, alpha, U, V)
        tested = k % depth == depth - 1 or k == n_acc - 1
        if tested:
            r = res(du, dv)
        nu = _fma(om[k], _fma(g, du, U) - UP, UP)
        nv = _fma(om[k], _fma(g, dv, V) - VP, VP)
        UP, VP, U, V = U, V, nu, nv
        count = k + 1
        if tested and not r > t1:
            break
    if finish > 1:
        # the finishing phase follows whether or not the accelerated one met
        # its tolerance
        fin = max(1, maxnum // 2)
        for j in range(fin):
            if count >= maxnum:
                break
            Tu, Tv = orc.hs_sweep(fx, fy, ft, alpha, U, V)
            tested = j % depth == depth - 1 or j == fin - 1
            if tested:
                r = res(Tu - U, Tv - V)
            U, V = Tu, Tv
            count += 1
            if tested and not r > tol:
                break
    return U, V, count, r


@pytest.mark.parametrize("depth,finish", [(4, 0.0), (3, 4.0), (10, 4.0)])
def test_hs_accel_launch_test_matches_restated_iteration(ops, depth, finish):
    """ofri_hs_accel with the stop test once per launch (the learned form's
    mode) equals its restatement over the oracle's sweep bit for bit."""
    from ofri.device import DeviceArray, context
    H, W = 96, 120
    im1, im2 = _piv(H, W, seed=41)
    plain_hist, _, _ = _hs_norm_history(im1, im2, 45.0, 200)
    tol = float(plain_hist[199])
    Ur, Vr, kr, rr = _cheb_restated_launches(im1, im2, 45.0, 0.99, tol, 200, finish, depth)
    assert kr < 150, kr
    ctx = context()
    ctx.set_tuning(hs_depth=depth)
    try:
        U = DeviceArray.zeros((H, W)); V = DeviceArray.zeros((H, W))
        err, it = ops.hs_accel(im1, im2, 45.0, 200, tol, 0.99, U, V, finish=finish, launch_test=True)
    finally:
        ctx.set_tuning()
    print("[hs accel per-launch test, depth %d, finish %g] %d sweeps" % (depth, finish, it))
    assert it == kr
    assert abs(float(err) - float(rr)) <= 1e-6 * float(rr)
    assert np.array_equal(U.download(), Ur) and np.array_equal(V.download(), Vr)


@pytest.mark.parametrize("depth,finish", [(0, 0.0), (4, 0.0), (3, 4.0)])
def test_hs_accel_matches_restated_iteration(ops, depth, finish):
    """ofri_hs_accel equals the Chebyshev iteration restated over the
    accelerated float32 sweep bit for bit (sweep count, residual history and the returned T(x_k)),
    and stops in far fewer sweeps than plain sweeps need for the same norm."""
    from ofri.device import DeviceArray, context
    H, W = 96, 120
    im1, im2 = _piv(H, W, seed=41)
    plain_hist, _, _ = _hs_norm_history(im1, im2, 45.0, 200)
    tol = float(plain_hist[199])
    Ur, Vr, kr, hr = _cheb_restated(im1, im2, 45.0, 0.99, tol, 200, finish)
    assert kr < 150, kr
    ctx = context()
    ctx.set_tuning(hs_depth=depth)
    try:
        U = DeviceArray.zeros((H, W)); V = DeviceArray.zeros((H, W))
        hist = DeviceArray.zeros((200,))
        err, it = ops.hs_accel(im1, im2, 45.0, 200, tol, 0.99, U, V, hist, finish=finish)
    finally:
        ctx.set_tuning()
    print("[hs accel %dx%d finish %g] plain needs 200 sweeps for %.3g, accelerated %d" % (H, W, finish, tol, it))
    assert it == kr
    assert np.allclose(hist.download()[:it], hr, rtol=1e-6, atol=0)
    assert np.array_equal(U.download(), Ur) and np.array_equal(V.download(), Vr)


def _ls_learned_restated(f, h, u, v, tol, maxnum, depth):
    """ofri_ls_learned restated over the oracle: launches of `depth` sweeps,
    every sweep but a launch's last the float32 sweep (oracle.ls_sweep_fast),
    the last the reference sweep, whose update norm is tested; stop after the
    first launch whose norm is <= tol.  Returns (u, v, sweeps, norms)."""
    count, norms = 0, []
    while count < maxnum:
        ns = min(depth, maxnum - count)
        for _ in range(ns - 1):
            u, v = orc.ls_sweep_fast(f, h, u, v)
        un, vn, _ = orc.ls_sweep(f, h, u, v)
        du, dv = un - u, vn - v
        norms.append(np.float32((np.sqrt((du.astype(np.float64) ** 2).sum()) +
                                 np.sqrt((dv.astype(np.float64) ** 2).sum())) / u.size))
        u, v = un, vn
        count += ns
        if not norms[-1] > tol:
            break
    return u, v, count, norms


@pytest.mark.parametrize("shape,variant,depth", [((96, 120), 2, 4), ((70, 90), 1, 3), ((130, 64), 4, 5)])
def test_ls_learned_matches_restated_iteration(ops, shape, variant, depth):
    """ofri_ls_learned (the learned form's Liu-Shen: float32 sweeps between
    the per-launch tests, the tested sweep the reference update) equals its
    restatement over the oracle bit for bit: iterate, sweep count and tested
    norm, with tiles on and off the image border."""
    from ofri.device import DeviceArray, context
    H, W = shape
    im1, im2 = _piv(H, W, seed=53)
    rng = np.random.default_rng(54)
    u0 = (rng.standard_normal((H, W)) * 0.3).astype(np.float32)
    v0 = (rng.standard_normal((H, W)) * 0.3).astype(np.float32)
    f = orc.ls_setup(im1, im2, 5.0)
    _, _, _, norms = _ls_learned_restated(f, 5.0, u0, v0, -1.0, 6 * depth, depth)
    tol = float(norms[2])                     # met at the end of the third launch
    ur, vr, kr, nr = _ls_learned_restated(f, 5.0, u0, v0, tol, 6 * depth, depth)
    assert kr == 3 * depth
    ctx = context()
    ctx.set_tuning(ls_depth=depth, ls_variant=variant)
    try:
        du, dv = DeviceArray.from_host(u0), DeviceArray.from_host(v0)
        err, it = ops.ls_learned(im1, im2, 5.0, du, dv, maxnum=6 * depth, tol=tol)
    finally:
        ctx.set_tuning()
    print("[ls learned %dx%d depth %d] %d sweeps, norm %.4g" % (H, W, depth, it, err))
    assert it == kr
    assert abs(float(err) - float(nr[-1])) <= 1e-6 * float(nr[-1])
    assert np.array_equal(du.download(), ur) and np.array_equal(dv.download(), vr)


@pytest.mark.parametrize("H,W,n", [(130, 98, 3), (66, 161, 5), (256, 256, 9), (40, 34, 11)])
def test_pyramid_hs_prefilter_fused_equals_separate_passes(ops, H, W, n):
    """The pyramid runs its HS pre-filter inside the derivative pass (one
    kernel for up to 9 taps, separate passes above): the coarse level's HS
    flow equals HS on the separately pre-filtered resized pair bit for bit,
    with tiles cut by the image border on both sides."""
    from gaussian_filter import prepareGaussianKernel
    from ofri.device import DeviceArray
    rng = np.random.default_rng(n)
    im1 = (rng.random((H, W)) * 255).astype(np.float32)
    im2 = np.roll(im1, 2, axis=1) + rng.standard_normal((H, W)).astype(np.float32)
    fw, ow = prepareGaussianKernel(1.5, n), prepareGaussianKernel(0.48, 5)
    _, _, taps, st = ops.pyr_lshs_taps(im1, im2, 2, [45.0, 21.0], 50, fw, ow, 5.0)
    h, w = st[0]["H"], st[0]["W"]
    t = taps.download()
    f1 = ops.gauss_sep(ops.resize(im1, h, w), fw)
    f2 = ops.gauss_sep(ops.resize(im2, h, w), fw)
    U, V = DeviceArray.zeros((h, w)), DeviceArray.zeros((h, w))
    ops.hs(f1, f2, 45.0, 50, U, V)
    assert np.array_equal(t[:h * w].reshape(h, w), U.download())
    assert np.array_equal(t[h * w:2 * h * w].reshape(h, w), V.download())


@pytest.mark.parametrize("H,W,seed", [(64, 80, 1), (37, 53, 2), (130, 257, 3)])
def test_liu_shen_warp_matches_oracle(ops, H, W, seed):
    """ofri_warp_liushen equals the oracle's restatement of the reference's
    Liu-Shen warping bit for bit: integer scatter with wrapping negative
    targets and colliding sources, the 73-tap smoothing of the fractional
    flow, the intensity correction; a target past the end raises IndexError."""
    rng = np.random.default_rng(seed)
    im = (rng.random((H, W)) * 200).astype(np.float32)
    # flows up to +-3 px, kept inside the image at the far edges (the
    # reference raises there); negative targets near the near edges wrap
    ys, xs = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    u = (rng.random((H, W)) * 6 - 3).astype(np.float32)
    v = (rng.random((H, W)) * 6 - 3).astype(np.float32)
    u = np.where(xs + np.floor(u + 0.5) >= W, np.float32(-1.25), u).astype(np.float32)
    v = np.where(ys + np.floor(v + 0.5) >= H, np.float32(-2.75), v).astype(np.float32)
    ref = orc.ls_warp(im, u, v)
    out = ops.warp_liushen(im, u, v).download()
    assert np.array_equal(out, ref)
    bad = u.copy()
    bad[H // 2, W - 1] = 0.5
    with pytest.raises(IndexError):
        ops.warp_liushen(im, bad, v)


def test_pyramid_liu_shen_warping_dropin(ops):
    """The drop-in's updateNextPyramidalLevel with biLinear=False: the first
    frame is warped in place (spline-upsampled, scaled flow) as the oracle
    restates it, the second frame is left alone, the flow becomes the
    accumulated one and the solver restarts from zero."""
    import GenericPyramidalOpticalFlow as G
    rng = np.random.default_rng(9)
    im1 = (rng.random((64, 96)) * 200).astype(np.float32)
    im2 = (rng.random((64, 96)) * 200).astype(np.float32)
    # flows pointing up and left (negative targets wrap; none past the end)
    Uc = (-rng.random((32, 48))).astype(np.float32)
    Vc = (-rng.random((32, 48))).astype(np.float32)
    prev = np.empty((32, 48), np.float32)
    a = im1.copy()
    w1, w2, Ua, Va, U, V = G.updateNextPyramidalLevel(a, prev, im2, Uc, Vc, None, None, True, False, True)
    us = orc.spline_upsample(Uc, 64, 96, np.float32(2.0))
    vs = orc.spline_upsample(Vc, 64, 96, np.float32(2.0))
    assert np.array_equal(Ua, us) and np.array_equal(Va, vs)
    assert np.array_equal(w1, orc.ls_warp(im1, us, vs)) and w1 is a
    assert np.array_equal(w2, im2) and not U.any() and not V.any()


def test_ls_small_golden(ops, golden):
    from ofri.device import DeviceArray
    g = golden("ls_small")
    du, dv = DeviceArray.from_host(g["u0"]), DeviceArray.from_host(g["v0"])
    err, it = ops.ls(g["im1"], g["im2"], float(g["h"]), du, dv, maxnum=int(g["nsweep"]), tol=0.0)
    assert it == int(g["nsweep"])
    np.testing.assert_allclose(du.download(), g["u"], rtol=2e-6, atol=2e-6 * np.abs(g["u"]).max())
    np.testing.assert_allclose(float(err), g["errs"][-1], rtol=1e-5)


def test_ls_adapter_swaps_components(ops):
    from PhysicsBasedOpticalFlowLiuShen import LiuShenOpticalFlowAlgoAdapter
    im1, im2 = _piv(40, 56, seed=24)
    rng = np.random.default_rng(25)
    U = rng.standard_normal(im1.shape).astype(np.float32)
    V = rng.standard_normal(im1.shape).astype(np.float32)
    resU, resV, err = LiuShenOpticalFlowAlgoAdapter(5).compute(im1, im2, U, V)
    rv, ru, _, _ = orc.ls(im1, im2, 5.0, V, U)   # solver u = caller V
    assert np.abs(resU - ru).max() <= 1e-5 and np.abs(resV - rv).max() <= 1e-5
    with pytest.raises(TypeError):
        LiuShenOpticalFlowAlgoAdapter(5).hasGenericPyramidalDefaults()


# ------------------------------------------------------------- whole LS(HS) pyramid
def _lshs_params():
    from gaussian_filter import prepareGaussianKernel
    return prepareGaussianKernel(3.4, 3), prepareGaussianKernel(0.48, 5)


@pytest.mark.parametrize("size,levels", [(256, 2), (256, 1), (200, 3)])
def test_pyramid_lshs_matches_oracle(ops, size, levels):
    from ofri import synth
    fw, ow = _lshs_params()
    im1, im2, ut, vt = synth.config_pair(4, size=size)
    alphas = [45.0, 21.0, 21.0][:levels] if levels > 1 else [21.0]
    U, V, st = ops.pyr_lshs(im1, im2, levels, alphas, 600, fw, ow, 5.0)
    Ug, Vg = U.download(), V.download()
    U32, V32, s32 = orc.pyr_lshs(im1, im2, levels, alphas, 600, fw, ow, 5.0)
    U64, V64, _ = orc.pyr_lshs(im1, im2, levels, alphas, 600, fw, ow, 5.0, prec="f64")
    d32, d64 = report("lshs %d^2 L%d" % (size, levels), Ug, Vg, U32, V32, U64, V64)
    assert d32 <= FLOW_TOL
    epe_g, epe_o = _epe(Ug, Vg, ut, vt).mean(), _epe(U32, V32, ut, vt).mean()
    assert abs(epe_g - epe_o) <= EPE_TOL
    for l in range(levels):
        assert st[l]["ls_iters"] == int(s32[l, 2])
        assert abs(float(st[l]["hs_err"]) - s32[l, 0]) <= 1e-5 * abs(s32[l, 0]) + 1e-12
    print("  EPE to truth: gpu %.4f px, oracle %.4f px" % (epe_g, epe_o))


def test_dropin_driver_on_reference_piv_pair(ops, golden):
    """The reference's two-level LS(HS) example script
    (examples/LiuSE_PyHSchunck_Fs3_4_PyrLvls2.py) on the reference's own TIFF
    pair, through the drop-in API."""
    from GenericPyramidalOpticalFlow import genericPyramidalOpticalFlow
    from HornSchunck import HSOpticalFlowAlgoAdapter
    from PhysicsBasedOpticalFlowLiuShen import LiuShenOpticalFlowAlgoAdapter
    g = golden("piv_pair")
    Iold, Inew = g["im0"].astype(np.float32), g["im1"].astype(np.float32)
    hs = HSOpticalFlowAlgoAdapter([21, 45], 600)
    U, V = genericPyramidalOpticalFlow(Iold, Inew, 3.4, hs, 2, 1, 0.48, LiuShenOpticalFlowAlgoAdapter(5))
    assert hs.alphas == []
    fw, ow = _lshs_params()
    U32, V32, _ = orc.pyr_lshs(Iold, Inew, 2, [45.0, 21.0], 600, fw, ow, 5.0)
    U64, V64, _ = orc.pyr_lshs(Iold, Inew, 2, [45.0, 21.0], 600, fw, ow, 5.0, prec="f64")
    d32, _ = report("reference TIFF pair", U, V, U32, V32, U64, V64)
    assert d32 <= FLOW_TOL
    # Poiseuille profile: |u(y)| peaked mid-channel, small at the walls
    prof = np.abs(U.mean(axis=1))
    assert prof[200:312].mean() > 3 * prof[:40].mean()


def test_generic_path_equals_device_path(ops):
    """An adapter wrapper forces the level loop (numpy at the adapter
    boundary); it must reproduce the single device-resident call."""
    from GenericPyramidalOpticalFlow import genericPyramidalOpticalFlow
    from HornSchunck import HSOpticalFlowAlgoAdapter
    from PhysicsBasedOpticalFlowLiuShen import LiuShenOpticalFlowAlgoAdapter
    from ofri import synth
    im1, im2, _, _ = synth.config_pair(4, size=128)

    class Wrapped(HSOpticalFlowAlgoAdapter):
        pass

    Ua, Va = genericPyramidalOpticalFlow(im1, im2, 3.4, HSOpticalFlowAlgoAdapter([21, 45], 200), 2, 1, 0.48,
                                         LiuShenOpticalFlowAlgoAdapter(5))
    Ub, Vb = genericPyramidalOpticalFlow(im1, im2, 3.4, Wrapped([21, 45], 200), 2, 1, 0.48,
                                         LiuShenOpticalFlowAlgoAdapter(5))
    assert np.array_equal(Ua, Ub) and np.array_equal(Va, Vb)


def test_generic_pyramid_liu_shen_warping_runs_the_reference_loop(ops):
    """genericPyramidalOpticalFlow with Liu-Shen warping (biLinear=False),
    through the level loop: with one iteration per level its flow is close
    to (not the same as) bilinear warping's; with a re-warp iteration
    (kLevels=2) it runs to finite flows.  (At levels after the first the
    reference re-warps the frame its first warp moved in place, so kLevels=2
    drifts from the shift; the drop-in keeps that.)"""
    from GenericPyramidalOpticalFlow import genericPyramidalOpticalFlow
    from HornSchunck import HSOpticalFlowAlgoAdapter
    from PhysicsBasedOpticalFlowLiuShen import LiuShenOpticalFlowAlgoAdapter

    class NoDefaults(HSOpticalFlowAlgoAdapter):
        def hasGenericPyramidalDefaults(self):
            return False

    im1, im2 = _piv(96, 128, seed=61, shift=(-1.3, -0.6))
    out = {}
    for bil in (False, True):
        out[bil] = genericPyramidalOpticalFlow(im1, im2, 3.4, NoDefaults([21, 45], 60), 2, 1, 0.48,
                                               LiuShenOpticalFlowAlgoAdapter(5), True, bil, True, True)
    (U, V), (Ub, Vb) = out[False], out[True]
    assert U.shape == (96, 128) and np.isfinite(U).all() and np.isfinite(V).all()
    dU = np.median(np.abs(U - Ub)[16:-16, 16:-16])
    dV = np.median(np.abs(V - Vb)[16:-16, 16:-16])
    print("[liu-shen warping] median U %.3f V %.3f, median |difference to bilinear| %.4f %.4f px"
          % (np.median(U), np.median(V), dU, dV))
    assert dU < 0.3 and dV < 0.3
    assert not np.array_equal(U, Ub)
    # the adapter pops one alpha per compute call: two per level with kLevels=2
    U2, V2 = genericPyramidalOpticalFlow(im1, im2, 3.4, NoDefaults([21, 21, 45, 45], 60), 2, 2, 0.48,
                                         LiuShenOpticalFlowAlgoAdapter(5), True, False, True, True)
    assert np.isfinite(U2).all() and np.isfinite(V2).all()


def test_full_size_pipeline_properties(ops):
    """1024^2 LS(HS) 2-level, the headline configuration at its own size:
    deterministic across runs, finite, close to the synthetic ground truth
    away from the borders, and within FLOW_TOL of the float32 oracle's
    pyramid (same Liu-Shen sweep counts per level)."""
    from ofri import synth
    fw, ow = _lshs_params()
    im1, im2, ut, vt = synth.config_pair(4)
    U1, V1, st = ops.pyr_lshs(im1, im2, 2, [45.0, 21.0], 600, fw, ow, 5.0)
    U2, V2, _ = ops.pyr_lshs(im1, im2, 2, [45.0, 21.0], 600, fw, ow, 5.0)
    a, b = U1.download(), V1.download()
    assert np.array_equal(a, U2.download()) and np.array_equal(b, V2.download())
    assert np.isfinite(a).all() and np.isfinite(b).all()
    epe = _epe(a, b, ut, vt)[64:-64, 64:-64].mean()
    print("  1024^2 EPE to truth %.4f px, stats %s" % (epe, st))
    assert epe < 1.0
    U32, V32, s32 = orc.pyr_lshs(im1, im2, 2, [45.0, 21.0], 600, fw, ow, 5.0)
    d32 = max(np.abs(a - U32).max(), np.abs(b - V32).max())
    print("  1024^2 max|d| vs oracle-f32 %.3g px" % d32)
    assert d32 <= FLOW_TOL
    assert [s["ls_iters"] for s in st] == [int(s32[l, 2]) for l in range(2)]


def test_pyramid_graph_replays_match_eager(ops):
    """ofri_pyr_lshs runs eagerly, then captures, then replays a HIP graph:
    every replay (with unrelated eager work and syncs in between) must equal
    the eager result bit for bit."""
    from ofri import synth
    from ofri.device import DeviceArray, context
    fw, ow = _lshs_params()
    im1, im2, _, _ = synth.config_pair(4, size=192)
    ctx = context()
    d1, d2 = DeviceArray.from_host(im1), DeviceArray.from_host(im2)
    Ua, Va = DeviceArray((192, 192)), DeviceArray((192, 192))
    ctx.set_graphs(False)
    ops.pyr_lshs(d1, d2, 2, [45.0, 21.0], 120, fw, ow, 5.0, Uacc=Ua, Vacc=Va)
    ref = (Ua.download(), Va.download())
    ctx.set_graphs(True)
    for rep in range(5):
        ops.pyr_lshs(d1, d2, 2, [45.0, 21.0], 120, fw, ow, 5.0, Uacc=Ua, Vacc=Va)
        got = (Ua.download(), Va.download())
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), rep
        ops.gauss_sep(d1, fw).download()         # unrelated eager work between replays
        Ua.upload(np.full((192, 192), 7.0, np.float32))
