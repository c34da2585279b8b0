"""Pins the CPU oracle against the library-probed golden vectors, brute force on
tiny inputs and the mathematics (fixed points, closed forms).  CPU only."""
import numpy as np
import pytest

from oracle import oracle as orc


def ulp_diff(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, -(a & 0x7FFFFFFF), a)
    b = np.where(b < 0, -(b & 0x7FFFFFFF), b)
    return np.abs(a - b).max()


# --------------------------------------------------------------- stencils
def test_hs_derivative_stencils_match_scipy(golden):
    g = golden("stencils")
    X = g["X"]
    z = np.zeros_like(X)
    fx, fy, ft = orc.hs_derivs(X, z)           # I0 = X, I1 = 0
    assert np.array_equal(fx, g["hs_cx"])
    assert np.array_equal(fy, g["hs_cy"])
    assert np.array_equal(ft, g["hs_ct"])
    fx, fy, ft = orc.hs_derivs(z, X)           # I1 = X: ft = -CT[X]
    assert np.array_equal(fx, g["hs_cx"])
    assert np.array_equal(ft, -g["hs_ct"])


def test_hs_average_matches_scipy(golden):
    g = golden("stencils")
    Y = g["Y"]
    z = np.zeros_like(Y)
    un, vn = orc.hs_sweep(z, z, z, np.float32(21), Y, 2 * Y)   # der = 0 -> ubar
    assert np.array_equal(un, g["hs_avg"])
    assert np.array_equal(vn, (2 * Y).astype(np.float32) * 0 + orc.hs_sweep(z, z, z, 21, 2 * Y, Y)[0])


def _ls_probe(Y, **fields):
    H, W = Y.shape
    base = {k: np.zeros((H, W), np.float32) for k in ("IIx", "IIy", "II", "Ixt", "Iyt", "B11", "B12", "B22")}
    base["B11"][:] = -1.0                    # u' = -(B11 bu) = bu
    base.update({k: np.full((H, W), v, np.float32) for k, v in fields.items()})
    return base


def test_ls_sweep_stencils_match_scipy(golden):
    """Each LS sweep stencil isolated by choosing the coefficient fields so that
    u' equals one term of bu (PhysicsBasedOpticalFlowLiuShen.py:142-144)."""
    g = golden("stencils")
    Y = g["Y"]
    z = np.zeros_like(Y)
    un, _, _ = orc.ls_sweep(_ls_probe(Y, IIx=0.5), 0, Y, z)          # 2*.5*Dr[u]
    assert np.array_equal(un, g["ls_dr"])
    un, _, _ = orc.ls_sweep(_ls_probe(Y, IIx=1.0), 0, z, Y)          # Dc[v]
    assert np.array_equal(un, g["ls_dc"])
    un, _, _ = orc.ls_sweep(_ls_probe(Y, IIy=1.0), 0, z, Y)          # Dr[v]
    assert np.array_equal(un, g["ls_dr"])
    un, _, _ = orc.ls_sweep(_ls_probe(Y, II=1.0), 0, Y, z)           # Fr[u]
    assert np.array_equal(un, g["ls_fr"])
    un, _, _ = orc.ls_sweep(_ls_probe(Y, II=1.0), 0, z, Y)           # M[v]
    assert np.array_equal(un, g["ls_m"])
    un, _, _ = orc.ls_sweep(_ls_probe(Y), 1, Y, z)                   # H0[u], zero padded
    assert np.array_equal(un, g["ls_h0"])


# --------------------------------------------------------------- solvers vs scipy-driven runs
def test_hs_small_matches_scipy_driven_run(golden):
    g = golden("hs_small")
    fx, fy, ft = orc.hs_derivs(g["I0"], g["I1"])
    assert np.array_equal(fx, g["fx"]) and np.array_equal(fy, g["fy"]) and np.array_equal(ft, g["ft"])
    U, V, _ = orc.hs(g["I0"], g["I1"], g["alpha"], int(g["niter"]), g["U0"], g["V0"])
    assert ulp_diff(U, g["U"]) <= 1 and ulp_diff(V, g["V"]) <= 1


def test_ls_small_matches_scipy_driven_run(golden):
    g = golden("ls_small")
    f = orc.ls_setup(g["im1"], g["im2"], np.float32(g["h"]))
    for k in ("IIx", "IIy", "II", "Ixt", "Iyt", "B11", "B12", "B22"):
        assert ulp_diff(f[k], g[k]) <= 1, k
    u, v = g["u0"], g["v0"]
    errs = []
    for _ in range(int(g["nsweep"])):
        u, v, e = orc.ls_sweep(f, float(g["h"]), u, v)
        errs.append(e)
    assert ulp_diff(u, g["u"]) <= 2 and ulp_diff(v, g["v"]) <= 2
    np.testing.assert_allclose(errs, g["errs"], rtol=1e-6)
    # driver: fixed sweep count through maxnum, tolerance 0 never stops early
    u2, v2, e2, it = orc.ls(g["im1"], g["im2"], float(g["h"]), g["u0"], g["v0"], maxnum=int(g["nsweep"]), tol=0.0)
    assert it == int(g["nsweep"])
    assert np.array_equal(u2, u) and np.array_equal(v2, v)


def test_ls_stops_on_tolerance():
    rng = np.random.default_rng(3)
    a = (rng.random((20, 24)) * 200 + 10).astype(np.float32)
    b = np.roll(a, 1, axis=1)
    z = np.zeros_like(a)
    _, _, e_all, it_all = orc.ls(a, b, 5.0, z, z, maxnum=60, tol=0.0)
    assert it_all == 60
    # a tolerance just above the 10th sweep's norm stops there
    errs = []
    u, v = z, z
    f = orc.ls_setup(a, b, np.float32(5))
    for _ in range(12):
        u, v, e = orc.ls_sweep(f, 5.0, u, v)
        errs.append(np.float32(e))
    tol = float(errs[9]) * 1.0000001
    _, _, e10, it10 = orc.ls(a, b, 5.0, z, z, maxnum=60, tol=tol)
    assert it10 <= 10 and e10 <= tol


# --------------------------------------------------------------- fixed points (mathematics)
def test_hs_fixed_point_satisfies_discrete_equations():
    """After many FP64 sweeps, (U, V) satisfy the HS Euler-Lagrange equations
    (alpha^2 + fx^2 + fy^2) U = alpha^2 ... in Jacobi form: U = ubar - fx*der."""
    rng = np.random.default_rng(5)
    H, W = 12, 10
    I0 = rng.random((H, W)) * 50
    I1 = np.roll(I0, 1, axis=1) + rng.random((H, W))
    U, V, _ = orc.hs(I0, I1, 3.0, 4000, prec="f64")
    fx, fy, ft = orc.hs_derivs(I0, I1, prec="f64")
    ub, vb = orc.hs_sweep(np.zeros((H, W)), np.zeros((H, W)), np.zeros((H, W)), 3.0, U, V, prec="f64")
    # residual of  fx (fx U + fy V + ft) + alpha^2 (U - ubar) = 0
    r = fx * (fx * U + fy * V + ft) + 9.0 * (U - ub)
    assert np.abs(r).max() < 1e-8 * (1 + np.abs(fx).max() ** 2)
    r = fy * (fx * U + fy * V + ft) + 9.0 * (V - vb)
    assert np.abs(r).max() < 1e-8 * (1 + np.abs(fy).max() ** 2)


def test_accel_sweep_is_the_hs_sweep_to_float32_rounding():
    """The accelerated phase's float32 sweep is the same operator as the
    reference-dataflow sweep: on PIV-like fields both agree to a few float32
    ulps (a dropped or mis-weighted neighbour, a wrong sign or a transposed
    derivative moves them apart by orders of magnitude), and the float64
    instance of the exact sweep sits between them to the same tolerance."""
    rng = np.random.default_rng(7)
    H, W = 37, 53
    fx, fy, ft = (rng.standard_normal((H, W)).astype(np.float32) * s for s in (20.0, 20.0, 10.0))
    U, V = (rng.standard_normal((H, W)).astype(np.float32) for _ in range(2))
    for alpha in (45.0, 21.0, 3.0):
        tu, tv = orc.hs_sweep_accel(fx, fy, ft, alpha, U, V)
        eu, ev = orc.hs_sweep(fx, fy, ft, alpha, U, V)
        du, dv = orc.hs_sweep(fx, fy, ft, alpha, U, V, prec="f64")
        scale = max(np.abs(eu).max(), np.abs(ev).max())
        assert max(np.abs(tu - eu).max(), np.abs(tv - ev).max()) <= 8 * np.finfo(np.float32).eps * scale
        assert max(np.abs(tu - du).max(), np.abs(tv - dv).max()) <= 8 * np.finfo(np.float32).eps * scale
        assert not np.array_equal(tu, eu)  # it is a different rounding of the same operator
    # a mirror-symmetric field stays mirror-symmetric (the tile's ghost cells
    # evolve exactly like their mirror pixels)
    S = rng.standard_normal((H, W)).astype(np.float32)
    S = np.concatenate([S, S[:, -2::-1]], axis=1)     # mirror about column W-1
    F = [np.concatenate([f, f[:, -2::-1]], axis=1) for f in (fx, fy, ft)]
    tu, _ = orc.hs_sweep_accel(F[0], F[1], F[2], 21.0, S, S)
    assert np.array_equal(tu[:, W - 1:], tu[:, W - 1::-1])


def test_fast_ls_sweep_is_the_ls_sweep_to_float32_rounding():
    """The learned form's float32 Liu-Shen sweep is the same operator as the
    reference-dataflow sweep on a real setup (borders included): both agree
    to a few float32 ulps of the update's scale, while a dropped neighbour, a
    sign flip of the corner term or swapped derivative fields move them
    apart by orders of magnitude."""
    rng = np.random.default_rng(11)
    H, W = 29, 41
    im1 = (rng.random((H, W)) * 200 + 20).astype(np.float32)
    im2 = np.roll(im1, 1, axis=1) + rng.standard_normal((H, W)).astype(np.float32)
    f = orc.ls_setup(im1, im2, 5.0)
    u, v = (rng.standard_normal((H, W)).astype(np.float32) * 0.5 for _ in range(2))
    eu, ev, _ = orc.ls_sweep(f, 5.0, u, v)
    tu, tv = orc.ls_sweep_fast(f, 5.0, u, v)
    scale = max(np.abs(eu).max(), np.abs(ev).max())
    tol = 16 * np.finfo(np.float32).eps * scale
    assert max(np.abs(tu - eu).max(), np.abs(tv - ev).max()) <= tol
    assert not np.array_equal(tu, eu)
    g = dict(f)
    g["IIx"], g["IIy"] = f["IIy"], f["IIx"]
    su, sv = orc.ls_sweep_fast(g, 5.0, u, v)
    assert max(np.abs(su - eu).max(), np.abs(sv - ev).max()) > 1e3 * tol
    su, sv = orc.ls_sweep_fast(f, 5.0, u, -v)
    assert max(np.abs(su - eu).max(), np.abs(sv - ev).max()) > 1e3 * tol


def test_scipy_gaussian_restatement_is_bit_exact():
    """The Liu-Shen warp's 73-tap smoothing restated (float64 symmetric sums
    in scipy's order, float32 between the axes, 'reflect' borders, repeated
    on lines shorter than the filter) equals scipy.ndimage.gaussian_filter
    bit for bit."""
    import scipy.ndimage as nd
    rng = np.random.default_rng(3)
    for shape in [(50, 61), (37, 20), (12, 9), (1, 40)]:
        x = (rng.random(shape) - 0.5).astype(np.float32)
        assert np.array_equal(orc.gauss_filter_scipy(x, 1.8, 20.0), nd.gaussian_filter(x, 1.8, truncate=20.0))
    assert len(orc.gauss_kernel_scipy(1.8, 20.0)) == 73


def test_liu_shen_warp_semantics():
    """Liu-Shen warping: a zero flow leaves the image as it is; a whole-pixel
    flow moves pixels (rounded half up, negative targets wrapping, the last
    source in row-major order winning) with no intensity correction; a
    target past the end raises IndexError as the reference's assignment
    does; a fractional flow changes all but the last row and column."""
    rng = np.random.default_rng(4)
    H, W = 9, 11
    im = (rng.random((H, W)) * 100).astype(np.float32)
    z = np.zeros((H, W), np.float32)
    assert np.array_equal(orc.ls_warp(im, z, z), im)
    u = np.full((H, W), -1.0, np.float32)         # every pixel one column left, column 0 wrapping
    out = orc.ls_warp(im, u, z)
    assert np.array_equal(out[:, :-1], im[:, 1:]) and np.array_equal(out[:, -1], im[:, 0])
    u = np.zeros((H, W), np.float32)
    u[4, 2] = 1.0                                  # (4, 2) and (4, 3) both land on (4, 3): the later source wins
    u[4, 3] = 0.0
    out = orc.ls_warp(im, u, z)
    assert out[4, 3] == im[4, 3] and out[4, 2] == im[4, 2]
    u[4, 3] = -1.0                                 # now (4, 3) goes to (4, 2) and (4, 2) to (4, 3): a swap
    out = orc.ls_warp(im, u, z)
    assert out[4, 3] == im[4, 2] and out[4, 2] == im[4, 3]
    v = np.zeros((H, W), np.float32)
    v[0, 3] = -1.0                                 # row 0 -> row -1 = the last row,
    v[H - 1, 3] = -1.0                             # whose own pixel moves up
    w = orc.ls_warp(im, z, v)
    assert w[H - 1, 3] == im[0, 3] and w[H - 2, 3] == im[H - 1, 3]
    u2 = z.copy()
    u2[2, W - 1] = 0.5                             # rounds half up to column W: out of range
    try:
        orc.ls_warp(im, u2, z)
        raise AssertionError("expected IndexError")
    except IndexError:
        pass
    f = orc.ls_warp(im, np.full((H, W), 0.25, np.float32), z)
    assert not np.array_equal(f[:-1, :-1], im[:-1, :-1])
    assert np.array_equal(f[-1, :], im[-1, :]) and np.array_equal(f[:, -1], im[:, -1])


def test_ls_fixed_point_solves_linear_system():
    """LS Jacobi fixed point: A [u; v] = -b(u, v) per pixel, where the sweep is
    u' = -(B bu): at convergence u = u'."""
    rng = np.random.default_rng(6)
    H, W = 6, 7
    a = (rng.random((H, W)) * 200 + 10)
    b = np.roll(a, 1, axis=0) * 0.98 + 3
    z = np.zeros_like(a)
    f = orc.ls_setup(a, b, 5.0, prec="f64")
    # the sweep is affine, x' = G x + c: probe G column by column, solve directly
    c = np.concatenate([t.ravel() for t in orc.ls_sweep(f, 5.0, z, z, prec="f64")[:2]])
    n = H * W
    G = np.zeros((2 * n, 2 * n))
    for q in range(2 * n):
        e = np.zeros(2 * n); e[q] = 1.0
        un, vn, _ = orc.ls_sweep(f, 5.0, e[:n].reshape(H, W), e[n:].reshape(H, W), prec="f64")
        G[:, q] = np.concatenate([un.ravel(), vn.ravel()]) - c
    xs = np.linalg.solve(np.eye(2 * n) - G, c)
    assert np.abs(np.linalg.eigvals(G)).max() < 1.0          # Jacobi contraction
    u, v, e, it = orc.ls(a, b, 5.0, z, z, maxnum=5000, tol=1e-16, prec="f64")
    np.testing.assert_allclose(np.concatenate([u.ravel(), v.ravel()]), xs, atol=1e-9)


def test_f32_and_f64_instances_agree():
    rng = np.random.default_rng(7)
    I0 = (rng.random((32, 40)) * 255).astype(np.float32)
    I1 = np.roll(I0, 1, axis=1)
    U32, V32, e32 = orc.hs(I0, I1, 21.0, 200)
    U64, V64, e64 = orc.hs(I0, I1, 21.0, 200, prec="f64")
    assert np.abs(U32 - U64).max() < 1e-4 and np.abs(V32 - V64).max() < 1e-4
    assert abs(e32 - e64) < 1e-6 * max(1, abs(e64)) + 1e-9


# --------------------------------------------------------------- pre-filter brute force
def _prefilter_literal(img, k):
    """The reference's own construction (gaussian_filter.py:54-85): padded line
    buffer, full 1-D convolution accumulated in float32, then slicing."""
    img = img.astype(np.float32).copy()
    n = len(k); h = n // 2
    H, W = img.shape

    def line(v):
        L = len(v)
        tmp = np.zeros(L + n - 1, np.float32)
        tmp[h:h + L] = v
        for j in range(h):
            tmp[h - 1 - j] = v[j]
            tmp[L + n - 2 - j] = v[L - 1 - j]
        full = np.zeros(L + n - 1 + 2 * h, np.float32)
        for i in range(len(tmp) + n - 1):
            for j in range(n):
                if 0 <= i - j < len(tmp):
                    full[i] = np.float32(full[i] + np.float32(tmp[i - j] * k[j]))
        res = full[h:h + len(tmp)]
        return res[h:-h]
    for i in range(H):
        img[i, :] = line(img[i, :])
    for j in range(W):
        img[:, j] = line(img[:, j])
    return img


@pytest.mark.parametrize("key", ["s3.4_n3", "s0.48_n5", "s2_n3", "s2.5_n7"])
def test_prefilter_matches_literal_construction(golden, key):
    k = golden("gauss_w")[key]
    rng = np.random.default_rng(8)
    img = (rng.random((9, 13)) * 255).astype(np.float32)
    ref = _prefilter_literal(img, k)
    out = orc.gauss_sep(img, k)
    assert np.array_equal(out, ref)


def test_prefilter_constant_and_impulse():
    k = np.array([0.25, 0.5, 0.25], np.float32)
    out = orc.gauss_sep(np.full((6, 7), 8.0, np.float32), k)
    assert np.all(out == 8.0)
    imp = np.zeros((7, 7), np.float32); imp[3, 3] = 1
    out = orc.gauss_sep(imp, k)
    assert np.array_equal(out[2:5, 2:5], np.outer(k, k).astype(np.float32))


# --------------------------------------------------------------- resize / spline / warp
def test_resize_bicubic_matches_pillow(golden):
    g = golden("pil")
    for t, (H, W, H2, W2) in enumerate(g["cases"]):
        out = orc.resize_bicubic(g["in%d" % t], int(H2), int(W2))
        assert np.array_equal(out, g["bicubic%d" % t]), (H, W, H2, W2)


def test_pil_coefficients_impulse(golden):
    g = golden("pil")
    k, b = orc.pil_coeffs(21, 10, 3)
    # column response to an impulse at 10 equals the coefficient of input 10
    resp = np.array([k[xx, 10 - b[xx, 0]] if 0 <= 10 - b[xx, 0] < b[xx, 1] else 0.0 for xx in range(10)])
    assert np.array_equal(resp.astype(np.float32), g["impulse_bicubic"][0])


def test_spline_matches_rectbivariatespline(golden):
    g = golden("spline")
    for t, (hc, wc, h, w) in enumerate(g["cases"]):
        _, d64 = orc.spline_upsample(g["in%d" % t], int(h), int(w), 1.0, want_f64=True)
        np.testing.assert_allclose(d64, g["out%d" % t], rtol=0, atol=1e-11)


def test_spline_reproduces_cubics_and_interpolates():
    n, N = 9, 18
    Wm = orc.spline_matrix(n, N)
    x = np.arange(n, dtype=np.float64)
    u = np.minimum(np.arange(N) * n / N, n - 1)
    for p in (lambda t: 1 + 0 * t, lambda t: t, lambda t: t ** 2 - 3 * t, lambda t: 0.1 * t ** 3 - t):
        np.testing.assert_allclose(Wm @ p(x), p(u), atol=1e-10)
    assert np.allclose(Wm[::2], np.eye(n)[: N // 2 + 0][: len(Wm[::2])], atol=1e-13)


def test_warp_matches_numpy(golden):
    g = golden("warp")
    o1, o2 = orc.warp_symmetric(g["im1"], g["im2"], g["U"], g["V"])
    assert np.array_equal(o1, g["o1"]) and np.array_equal(o2, g["o2"])


def test_warp_zero_flow_is_identity():
    rng = np.random.default_rng(9)
    a = (rng.random((8, 9)) * 255).astype(np.float32)
    z = np.zeros_like(a)
    o1, o2 = orc.warp_symmetric(a, a * 2, z, z)
    assert np.array_equal(o1, a) and np.array_equal(o2, a * 2)


def test_oracle_is_clean_under_address_and_undefined_behaviour_sanitizers():
    """The oracle built with -fsanitize=address,undefined and driven over every
    entry point on small, odd and edge-case shapes in both instances: no
    out-of-bounds access, leak or undefined operation."""
    import os
    import subprocess
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    subprocess.run(["make", "-s", "-C", here, "sanitize"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([os.path.join(here, "san_check")], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "oracle sanitizer check: ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
