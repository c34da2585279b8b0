"""Dense Lucas-Kanade: oracle pins (the OpenCL sampler's own formulation on a
tiny image, recovery of a known shift, the singular-tensor rule) and GPU
parity."""
import math

import numpy as np
import pytest
from scipy.ndimage import correlate1d, map_coordinates

from oracle import oracle as orc


def _pair(H, W, shift, seed=3, sig=3.0):
    rng = np.random.default_rng(seed)
    base = rng.random((H + 60, W + 60))
    r = int(3 * sig)
    k = np.exp(-np.arange(-r, r + 1) ** 2 / (2 * sig * sig))
    k /= k.sum()
    base = correlate1d(correlate1d(base, k, axis=0, mode="wrap"), k, axis=1, mode="wrap")
    base = (base - base.mean()) / base.std() * 40 + 128
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64) + 30
    im1 = map_coordinates(base, [yy, xx], order=3)
    im2 = map_coordinates(base, [yy - shift[1], xx - shift[0]], order=3)
    return im1.astype(np.float32), im2.astype(np.float32)


def _sampler(T, x, y):
    """OpenCL 1.2 LINEAR sampler, unnormalised coordinates, clamp-to-edge:
    i0 = floor(x - .5), a = frac(x - .5) (OpenCL spec 8.2.1)."""
    H, W = T.shape
    xs, ys = np.float32(x) - np.float32(0.5), np.float32(y) - np.float32(0.5)
    i0, j0 = math.floor(xs), math.floor(ys)
    a, b = np.float32(xs - i0), np.float32(ys - j0)
    c = lambda v, n: min(max(v, 0), n - 1)
    t00, t10 = T[c(j0, H), c(i0, W)], T[c(j0, H), c(i0 + 1, W)]
    t01, t11 = T[c(j0 + 1, H), c(i0, W)], T[c(j0 + 1, H), c(i0 + 1, W)]
    one = np.float32(1)
    t = (one - a) * (one - b) * t00
    t = t + a * (one - b) * t10
    t = t + (one - a) * b * t01
    t = t + a * b * t11
    return np.float32(t)


def _lk_literal(I, J, hw, niter, wx, wy):
    """The lkDense kernel restated per output pixel with the sampler read at
    loc = prevPt + (c + .5, r + .5) (pyrlkDenseLargeW.cl:231,272-301,505-615);
    wx / wy: 0/1 weights of the window grid's columns / rows, applied to the
    derivatives and to the differences as the kernel applies them."""
    H, W = I.shape
    win = int(max(np.nonzero(wx)[0].max(), np.nonzero(wy)[0].max())) + 1
    wgt = np.outer(wy[:win], wx[:win]).astype(np.float32)
    U = np.zeros((H, W), np.float32)
    V = np.zeros((H, W), np.float32)
    for i in range(H):
        for j in range(W):
            Px, Py = j - hw, i - hw
            # I patch via the sampler at pixel centres (x - .5 in READI)
            P = np.zeros((win + 2, win + 2), np.float32)
            for r in range(win + 2):
                for c in range(win + 2):
                    P[r, c] = _sampler(I, Px + c - 1 + 0.5, Py + r - 1 + 0.5)
            Dx = 3 * (P[:-2, 2:] + P[2:, 2:] - P[:-2, :-2] - P[2:, :-2]) + 10 * (P[1:-1, 2:] - P[1:-1, :-2])
            Dy = 3 * (P[2:, :-2] + P[2:, 2:] - P[:-2, :-2] - P[:-2, 2:]) + 10 * (P[2:, 1:-1] - P[:-2, 1:-1])
            Dx, Dy = Dx * wgt, Dy * wgt
            A11 = np.float32((Dx.astype(np.float64) ** 2).sum())
            A12 = np.float32((Dx.astype(np.float64) * Dy).sum())
            A22 = np.float32((Dy.astype(np.float64) ** 2).sum())
            D = A11 * A22 - A12 * A12
            if D < np.float32(1.192092896e-07):
                continue
            A11, A12, A22 = A11 / D, A12 / D, A22 / D
            px, py = np.float32(j) - np.float32(hw), np.float32(i) - np.float32(hw)
            for _ in range(niter):
                if px < -hw or px >= W or py < -hw or py >= H:
                    break
                b1 = b2 = 0.0
                for r in range(win):
                    for c in range(win):
                        d = (_sampler(J, px + c + 0.5, py + r + 0.5) - P[r + 1, c + 1]) * wgt[r, c]
                        b1 += float(d) * float(Dx[r, c])
                        b2 += float(d) * float(Dy[r, c])
                B1, B2 = np.float32(b1), np.float32(b2)
                ddx = (A12 * B2 - A22 * B1) * np.float32(32)
                ddy = (A12 * B1 - A11 * B2) * np.float32(32)
                px, py = px + ddx, py + ddy
                if abs(ddx) < 0.01 and abs(ddy) < 0.01:
                    break
            U[i, j] = px + hw - j
            V[i, j] = py + hw - i
    return U, V


def _cols(w):
    return [int(c) for c in np.nonzero(w)[0]]


def test_window_grid_weights():
    """Window columns counted by lkDense (pyrlkDenseLargeW.cl:338-371, 4 tiles
    of 8 per axis, WSX = window >= 16): a window of 16 or more columns counts
    exactly its 2hw+1 columns up to the grid's 32; a smaller one counts tile 0
    whole plus column 8; the left flag drops column 8, the right flag the
    last counted column past tile 0."""
    w = orc.lk_window_weights
    assert _cols(w(13)) == list(range(27))
    assert _cols(w(8)) == list(range(17))
    assert _cols(w(16)) == list(range(32))                 # 33 columns asked, the grid holds 32
    assert _cols(w(2)) == list(range(9))
    assert _cols(w(5)) == list(range(11))
    assert _cols(w(13, 1, 0)) == [c for c in range(27) if c != 8]
    assert _cols(w(13, 0, 1)) == list(range(26))
    assert _cols(w(13, 1, 1)) == [c for c in range(26) if c != 8]
    assert _cols(w(5, 0, 1)) == list(range(10))
    assert _cols(w(2, 0, 1)) == list(range(9))             # column 8 follows the left flag alone
    assert _cols(w(2, 1, 0)) == list(range(8))


@pytest.mark.parametrize("hw,asym", [(2, (0, 0, 0, 0)), (3, (1, 0, 0, 1)), (4, (0, 1, 1, 0))])
def test_oracle_matches_sampler_formulation(hw, asym):
    I, J = _pair(10, 11, (0.6, -0.3), seed=5)
    u, v, st = orc.lk_dense(I, J, np.zeros_like(I), np.zeros_like(I), half_win=hw, niter=3, asym=asym)
    wx = orc.lk_window_weights(hw, asym[0], asym[1])
    wy = orc.lk_window_weights(hw, asym[2], asym[3])
    U, V = _lk_literal(I, J, hw, 3, wx, wy)
    assert np.abs(u - U).max() < 2e-5 and np.abs(v - V).max() < 2e-5


def test_oracle_recovers_known_shift():
    I, J = _pair(80, 96, (1.3, -0.7))
    z = np.zeros_like(I)
    u, v, st = orc.lk_dense(I, J, z, z, half_win=13, niter=5)
    inner = (slice(20, -20), slice(20, -20))
    assert abs(np.median(u[inner]) - 1.3) < 0.03 and abs(np.median(v[inner]) + 0.7) < 0.03
    assert st[inner].all()
    u64, v64, _ = orc.lk_dense(I, J, z, z, half_win=13, niter=5, prec="f64")
    assert np.abs(u - u64).max() < 1e-3


def test_singular_tensor_leaves_flow_and_flags_status():
    """A ramp along x only has a rank-1 structure tensor: D < FLT_EPSILON, the
    pixel keeps its incoming flow and status 0 (pyrlkDenseLargeW.cl:492-500)."""
    x = np.tile(np.arange(40, dtype=np.float32), (30, 1)) * 2
    u0 = np.full(x.shape, 0.25, np.float32)
    u, v, st = orc.lk_dense(x, x + 1, u0, u0, half_win=3, niter=3)
    assert np.array_equal(u, u0) and np.array_equal(v, u0) and not st.any()


# ------------------------------------------------------------------ GPU
@pytest.fixture(scope="module")
def ops():
    from ofri import ops as _ops
    from ofri.device import context
    context()
    return _ops


# sd: spread of the initial flow.  0.3 keeps every block's displaced windows
# inside the kernel's LDS stage of J; 12 spreads them beyond it (the step
# then samples J from global memory) and sends pixels off the image; 3 mixes
# both kinds of block in one launch.
@pytest.mark.gpu
@pytest.mark.parametrize("H,W,hw,asym,sd", [(64, 80, 13, (0, 0, 0, 0), 0.3), (131, 97, 13, (0, 0, 0, 0), 0.3),
                                            (40, 50, 5, (0, 0, 0, 0), 0.3), (33, 29, 16, (0, 0, 0, 0), 0.3),
                                            (64, 80, 13, (1, 0, 0, 1), 0.3), (64, 80, 13, (0, 1, 0, 1), 0.3),
                                            (40, 50, 3, (1, 0, 1, 0), 0.3), (37, 45, 16, (0, 1, 1, 1), 0.3),
                                            (64, 80, 13, (0, 0, 0, 0), 12.0), (96, 96, 8, (0, 1, 1, 0), 3.0)])
def test_lk_gpu_matches_oracle(ops, H, W, hw, asym, sd):
    from ofri.device import DeviceArray
    I, J = _pair(H, W, (1.1, 0.5), seed=H + W)
    rng = np.random.default_rng(hw)
    u0 = (rng.standard_normal((H, W)) * sd).astype(np.float32)
    v0 = (rng.standard_normal((H, W)) * sd).astype(np.float32)
    u, v = DeviceArray.from_host(u0), DeviceArray.from_host(v0)
    st = ops.ByteArray((H, W))
    ops.lk_dense(I, J, u, v, hw, 5, st, asym)
    gu, gv, gs = u.download(), v.download(), st.download()
    ou, ov, os_ = orc.lk_dense(I, J, u0, v0, hw, 5, asym=asym)
    du, dv, _ = orc.lk_dense(I, J, u0, v0, hw, 5, prec="f64", asym=asym)
    d32 = max(np.abs(gu - ou).max(), np.abs(gv - ov).max())
    d64 = max(np.abs(gu - du).max(), np.abs(gv - dv).max())
    print("[lk %dx%d hw%d asym %s sd %g] max|d| vs oracle-f32 %.3g px, vs f64 %.3g px" % (H, W, hw, asym, sd, d32, d64))
    assert d32 <= 1e-4                 # parity bar: against the reference dataflow
    assert d64 <= 1e-2                 # FP32-vs-FP64 error: reported, sanity-bounded only
    assert np.array_equal(gs, os_)


@pytest.mark.gpu
def test_lk_dropin_adapter(ops):
    from denseLucasKanade_PyCL import denseLucasKanade_PyCl
    I, J = _pair(48, 64, (0.7, 0.2), seed=8)
    z = np.zeros_like(I)
    lk = denseLucasKanade_PyCl(Niter=5, halfWindow=13)
    U, V, flag = lk.compute(I, J, z, z)
    assert flag is True
    ou, ov, os_ = orc.lk_dense(I, J, z, z, 13, 5)
    assert np.abs(U - ou).max() <= 1e-4 and np.abs(V - ov).max() <= 1e-4
    assert np.array_equal(lk.status, os_)
    assert lk.getGenericPyramidalDefaults() == {'warping': False, 'intermediateScaling': True, 'scaling': False}


def _rotation_flow(H, W, omega):
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float32)
    cy, cx = (H - 1) / 2, (W - 1) / 2
    # solid rotation u = -w/2 (y - cy), v = w/2 (x - cx): vorticity dv/dx - du/dy = w
    return (np.float32(-omega / 2) * (yy - cy)).astype(np.float32), (np.float32(omega / 2) * (xx - cx)).astype(np.float32)


def test_vorticity_window_selection():
    """scipy's convolve flips the difference kernel, so the reference's omega
    is minus the vorticity: a counter-clockwise (positive) rotation selects
    the right/bottom window, a clockwise one the left/bottom window."""
    from denseLucasKanade_PyCL import denseLucasKanade_PyCl
    lk = denseLucasKanade_PyCl(enableVorticityEnhancement=True)
    assert lk.evaluateVorticityEnhancement(*_rotation_flow(40, 50, 0.01)) == [0, 1, 0, 1]
    assert lk.evaluateVorticityEnhancement(*_rotation_flow(40, 50, -0.01)) == [1, 0, 0, 1]
    assert lk.evaluateVorticityEnhancement(*_rotation_flow(40, 50, 0.001)) == [0, 0, 0, 0]
    assert denseLucasKanade_PyCl().evaluateVorticityEnhancement(*_rotation_flow(40, 50, 0.01)) == [0, 0, 0, 0]


@pytest.mark.gpu
@pytest.mark.parametrize("omega", [0.02, -0.02])
def test_lk_dropin_vorticity_enhancement(ops, omega):
    from denseLucasKanade_PyCL import denseLucasKanade_PyCl
    I, J = _pair(56, 72, (0.7, 0.2), seed=11)
    U0, V0 = _rotation_flow(56, 72, omega)
    lk = denseLucasKanade_PyCl(Niter=5, halfWindow=13, enableVorticityEnhancement=True)
    asym = lk.evaluateVorticityEnhancement(U0, V0)
    assert asym == ([0, 1, 0, 1] if omega > 0 else [1, 0, 0, 1])
    U, V, flag = lk.compute(I, J, U0, V0)
    ou, ov, os_ = orc.lk_dense(I, J, U0, V0, 13, 5, asym=asym)
    assert np.abs(U - ou).max() <= 1e-4 and np.abs(V - ov).max() <= 1e-4
    assert np.array_equal(lk.status, os_)
    su, sv, _ = orc.lk_dense(I, J, U0, V0, 13, 5)
    assert np.abs(U - su).max() > 1e-4                     # the asymmetric window changed the result
