"""ctypes wrapper of the CPU oracle (liboracle.so).  TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the benchmark's CPU-baseline leg import this
module, and only as the checker.  The shipped library never calls it.

Every function takes numpy arrays and returns new arrays; `prec` selects the
instance: "f32" is the reference dataflow (float32 storage, float64 stencil
sums), "f64" the same method in float64.  See ofri_oracle.c for the reference
file:line each entry follows.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
        _declare(_LIB)
    return _LIB


_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _declare(L):
    for suf, P, R in (("f32", _fp, C.c_float), ("f64", _dp, C.c_double)):
        getattr(L, "orc_gauss_sep_" + suf).argtypes = [P, C.c_int, C.c_int, P, C.c_int]
        getattr(L, "orc_hs_derivs_" + suf).argtypes = [P, P, C.c_int, C.c_int, P, P, P]
        getattr(L, "orc_hs_sweep_" + suf).argtypes = [P, P, P, R, C.c_int, C.c_int, P, P, P, P]
        f = getattr(L, "orc_hs_" + suf)
        f.argtypes = [P, P, C.c_int, C.c_int, R, C.c_int, P, P]
        f.restype = C.c_double
        getattr(L, "orc_ls_setup_" + suf).argtypes = [P, P, C.c_int, C.c_int, R] + [P] * 8
        f = getattr(L, "orc_ls_sweep_" + suf)
        f.argtypes = [P] * 8 + [R, C.c_int, C.c_int, P, P, P, P]
        f.restype = C.c_double
        f = getattr(L, "orc_ls_" + suf)
        f.argtypes = [P, P, C.c_int, C.c_int, R, C.c_int, C.c_double, P, P, _ip]
        f.restype = C.c_double
        getattr(L, "orc_resize_bicubic_" + suf).argtypes = [P, C.c_int, C.c_int, P, C.c_int, C.c_int]
        getattr(L, "orc_resize_" + suf).argtypes = [P, C.c_int, C.c_int, P, C.c_int, C.c_int, C.c_int]
        getattr(L, "orc_spline_upsample_" + suf).argtypes = [P, C.c_int, C.c_int, P, C.c_int, C.c_int, R, _dp]
        getattr(L, "orc_warp_symmetric_" + suf).argtypes = [P, P, P, P, C.c_int, C.c_int, P, P]
        getattr(L, "orc_pyr_lshs_" + suf).argtypes = [
            P, P, C.c_int, C.c_int, C.c_int, P, C.c_int, P, C.c_int, P, C.c_int, R, C.c_int,
            C.c_double, P, P, _dp]
    L.orc_spline_matrix.argtypes = [C.c_int, C.c_int, _dp]
    L.orc_pil_coeffs.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip]
    L.orc_pil_coeffs.restype = _dp
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_get_threads.restype = C.c_int


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def _dt(prec):
    return np.float32 if prec == "f32" else np.float64


def _arr(a, prec):
    return np.ascontiguousarray(a, dtype=_dt(prec))


def _p(a):
    return a.ctypes.data_as(_fp if a.dtype == np.float32 else _dp)


def gauss_sep(img, weights, prec="f32"):
    """gaussian_filterPx with the given weights; returns a new array."""
    out = _arr(img, prec).copy()
    w = _arr(weights, prec)
    H, W = out.shape
    getattr(lib(), "orc_gauss_sep_" + prec)(_p(out), H, W, _p(w), len(w))
    return out


def hs_derivs(I0, I1, prec="f32"):
    I0 = _arr(I0, prec); I1 = _arr(I1, prec)
    H, W = I0.shape
    fx, fy, ft = (np.empty_like(I0) for _ in range(3))
    getattr(lib(), "orc_hs_derivs_" + prec)(_p(I0), _p(I1), H, W, _p(fx), _p(fy), _p(ft))
    return fx, fy, ft


def hs_sweep(fx, fy, ft, alpha, U, V, prec="f32"):
    fx, fy, ft, U, V = (_arr(a, prec) for a in (fx, fy, ft, U, V))
    H, W = U.shape
    Un, Vn = np.empty_like(U), np.empty_like(V)
    getattr(lib(), "orc_hs_sweep_" + prec)(_p(fx), _p(fy), _p(ft), alpha, H, W, _p(U), _p(V), _p(Un), _p(Vn))
    return Un, Vn


def fma32(a, b, c):
    """float32 fused multiply-add a*b + c, rounded once: the 64-bit-mantissa
    intermediate holds the float32 product exactly, so only the sum rounds
    (to float32 as a single rounding would, but for a one-in-2^40 tie)."""
    return (np.asarray(a, np.longdouble) * np.asarray(b, np.longdouble)
            + np.asarray(c, np.longdouble)).astype(np.float32)


ACCEL_W = np.float32(1.0 / 12.0)   # the corner weight; the edge weight is 2 ACCEL_W exactly


def hs_sweep_accel_scaled(fx, fy, ft, alpha, U, V):
    """The Horn-Schunck sweep of the accelerated (Chebyshev) phase, all
    float32, scaled by 1/w (w = float32(1/12)): the same operator as hs_sweep
    (HornSchunck.py:51-66, the 3x3 average with scipy 'mirror' borders, then
    ub - fx (fx ub + fy vb + ft) / den) with the device's operation order.
    The average's kernel [[1,2,1],[2,0,2],[1,2,1]] w is the separable
    [1,2,1] x [1,2,1] kernel less its centre 4, so S = 12 x average is
    formed from the vertical pair sum p = x_up + x_down as c = fma(2, x, p)
    (the column's [1,2,1] sum), S = fma(2, p, c_left + c_right) (no
    cancellation: the centre never enters).  With ub = w S and the time
    derivative taken as ft' = RN(ft / w):
    der' = fma(fx, S_u, fma(fy, S_v, ft')) * RN(1 / fma(fy, fy, fma(fx, fx, alpha^2))),
    t = fma(-fx, der', S_u) = T / w (up to rounding).  Returns (t_u, t_v)."""
    f32 = np.float32
    fx, fy, ft, U, V = (np.asarray(x, f32) for x in (fx, fy, ft, U, V))
    a2 = f32(alpha) * f32(alpha)
    two = f32(2.0)

    def S(X):
        P = np.pad(X, 1, mode="reflect")
        p = P[:-2, :] + P[2:, :]
        c = fma32(two, P[1:-1, :], p)
        return fma32(two, p[:, 1:-1], c[:, :-2] + c[:, 2:])
    su, sv = S(U), S(V)
    fts = ft / ACCEL_W
    rcp = f32(1.0) / fma32(fy, fy, fma32(fx, fx, a2))
    der = fma32(fx, su, fma32(fy, sv, fts)) * rcp
    return fma32(-fx, der, su), fma32(-fy, der, sv)


def hs_sweep_accel(fx, fy, ft, alpha, U, V):
    """T(x) of the accelerated phase's sweep, RN(w t) with (t_u, t_v) from
    hs_sweep_accel_scaled: the value a plain step of that phase stores (its
    update is fma(w, t, -x)).  Returns (Tu, Tv)."""
    tu, tv = hs_sweep_accel_scaled(fx, fy, ft, alpha, U, V)
    return ACCEL_W * tu, ACCEL_W * tv


def ls_sweep_fast(fields, h, u, v):
    """The Liu-Shen sweep the learned form runs between its tests, all
    float32: the same operator as ls_sweep (PhysicsBasedOpticalFlowLiuShen.py
    :142-150; derivatives with 'nearest' borders, the 8-neighbour sum H0 with
    'constant' zero borders) with the device's operation order: full
    differences dr = u_down - u_up, dc = u_right - u_left, M = 0.25 ((ul - ur)
    - (dl - dr)), H0 in pairs, then
    bu = fma(h, H0u, fma(II, Fr(u) + M(v), fma(Iy/2, dr(v), fma(Ix/2, dc(v), fma(Ix, dr(u), Ixt)))))
    bv = fma(h, H0v, fma(II, M(u) + Fc(v), fma(Iy, dc(v), fma(Ix/2, dc(u), fma(Iy/2, dr(u), Iyt)))))
    u' = -fma(B11, bu, B12 bv), v' = -fma(B12, bu, B22 bv) (Ix = IIx, Iy = IIy).
    Returns (u', v')."""
    f32 = np.float32
    u, v = np.asarray(u, f32), np.asarray(v, f32)
    iix, iiy, ii2, ixt, iyt, b11, b12, b22 = (np.asarray(fields[k], f32) for k in
                                              ("IIx", "IIy", "II", "Ixt", "Iyt", "B11", "B12", "B22"))
    h = f32(h)

    def terms(X):
        E = np.pad(X, 1, mode="edge")
        Z = np.pad(X, 1, mode="constant")
        up, dn, lf, rt = E[:-2, 1:-1], E[2:, 1:-1], E[1:-1, :-2], E[1:-1, 2:]
        m = f32(0.25) * ((E[:-2, :-2] - E[:-2, 2:]) - (E[2:, :-2] - E[2:, 2:]))
        h0 = ((Z[:-2, :-2] + Z[:-2, 2:]) + (Z[2:, :-2] + Z[2:, 2:])) + \
             ((Z[:-2, 1:-1] + Z[2:, 1:-1]) + (Z[1:-1, :-2] + Z[1:-1, 2:]))
        return dn - up, rt - lf, up + dn, lf + rt, m, h0
    dru, dcu, fru, _, mu, h0u = terms(u)
    drv, dcv, _, fcv, mv, h0v = terms(v)
    hix, hiy = f32(0.5) * iix, f32(0.5) * iiy
    bu = fma32(iix, dru, ixt)
    bu = fma32(hix, dcv, bu)
    bu = fma32(hiy, drv, bu)
    bu = fma32(ii2, fru + mv, bu)
    bu = fma32(h, h0u, bu)
    bv = fma32(hiy, dru, iyt)
    bv = fma32(hix, dcu, bv)
    bv = fma32(iiy, dcv, bv)
    bv = fma32(ii2, mu + fcv, bv)
    bv = fma32(h, h0v, bv)
    return -fma32(b11, bu, b12 * bv), -fma32(b12, bu, b22 * bv)


def gauss_kernel_scipy(sigma, truncate):
    """scipy.ndimage's Gaussian weights (float64, normalised; radius
    int(truncate * sigma + 0.5)), as gaussian_filter1d forms them."""
    r = int(truncate * float(sigma) + 0.5)
    x = np.arange(-r, r + 1)
    phi = np.exp(-0.5 / (float(sigma) * float(sigma)) * x ** 2)
    return phi / phi.sum()


def _reflect(p, L):
    """scipy 'reflect' (half-sample symmetric) index of any position."""
    p = np.mod(p, 2 * L)
    return np.where(p >= L, 2 * L - 1 - p, p)


def gauss_filter_scipy(x, sigma, truncate):
    """scipy.ndimage.gaussian_filter of a float32 image, restated: axis 0
    then axis 1, each a symmetric float64 correlation summed as scipy's
    NI_Correlate1D sums a symmetric filter (centre tap first, then the pairs
    x[i-j] + x[i+j] from the outermost inwards), rounded to float32 between
    the axes, 'reflect' borders.  Used by the Liu-Shen warping mode
    (GenericPyramidalOpticalFlow.py:203-204)."""
    w = gauss_kernel_scipy(sigma, truncate)
    r = (len(w) - 1) // 2
    out = np.asarray(x, np.float32)
    for axis in (0, 1):
        a = np.moveaxis(out, axis, 0).astype(np.float64)
        L = a.shape[0]
        idx = np.arange(L)
        acc = a * w[r]
        for j in range(r, 0, -1):
            acc = acc + (a[_reflect(idx - j, L)] + a[_reflect(idx + j, L)]) * w[r - j]
        out = np.moveaxis(acc.astype(np.float32), 0, axis)
    return np.ascontiguousarray(out)


def ls_warp(im1, us, vs):
    """The Liu-Shen warping mode of updateNextPyramidalLevel
    (GenericPyramidalOpticalFlow.py:188-220) on float32 arrays: every pixel
    moved by its flow rounded half up (numpy fancy assignment: negative
    targets wrap, the last source in row-major order wins, targets past the
    end raise IndexError), the fractional flow smoothed by
    gaussian_filter(sigma 1.8, truncate 20: 73 taps), then the
    optical-flow-equation intensity correction of all but the last row and
    column.  Returns the warped first image."""
    im = np.array(im1, dtype=np.float32)
    us = np.asarray(us, np.float32)
    vs = np.asarray(vs, np.float32)
    H, W = im.shape
    ys, xs = np.meshgrid(np.arange(H, dtype=np.int32), np.arange(W, dtype=np.int32), indexing="ij")
    fu, fv = np.floor(us + np.float32(0.5)), np.floor(vs + np.float32(0.5))
    uswap = np.int32(xs + fu)
    vswap = np.int32(ys + fv)
    du = us - fu
    dv = vs - fv
    im[vswap, uswap] = im[ys, xs]
    mask = 3
    du = gauss_filter_scipy(du, 0.6 * mask, 4.0 / 0.6 * mask)
    dv = gauss_filter_scipy(dv, 0.6 * mask, 4.0 / 0.6 * mask)
    tdx = im[0:-1, 1:] * du[0:-1, 1:] - im[0:-1, 0:-1] * du[0:-1, 0:-1]
    tdy = im[1:, 0:-1] * dv[1:, 0:-1] - im[0:-1, 0:-1] * dv[0:-1, 0:-1]
    im[0:-1, 0:-1] = im[0:-1, 0:-1] - (tdx + tdy)
    return im


def hs(I0, I1, alpha, niter, U=None, V=None, prec="f32"):
    """Returns (U, V, total_error)."""
    I0 = _arr(I0, prec); I1 = _arr(I1, prec)
    H, W = I0.shape
    U = np.zeros((H, W), _dt(prec)) if U is None else _arr(U, prec).copy()
    V = np.zeros((H, W), _dt(prec)) if V is None else _arr(V, prec).copy()
    e = getattr(lib(), "orc_hs_" + prec)(_p(I0), _p(I1), H, W, alpha, int(niter), _p(U), _p(V))
    return U, V, e


def ls_setup(im1, im2, h, prec="f32"):
    """Returns dict of IIx, IIy, II, Ixt, Iyt, B11, B12, B22."""
    im1 = _arr(im1, prec); im2 = _arr(im2, prec)
    H, W = im1.shape
    names = ("IIx", "IIy", "II", "Ixt", "Iyt", "B11", "B12", "B22")
    f = {n: np.empty_like(im1) for n in names}
    getattr(lib(), "orc_ls_setup_" + prec)(_p(im1), _p(im2), H, W, h, *[_p(f[n]) for n in names])
    return f


def ls_sweep(fields, h, u, v, prec="f32"):
    names = ("IIx", "IIy", "II", "Ixt", "Iyt", "B11", "B12", "B22")
    fs = [_arr(fields[n], prec) for n in names]
    u = _arr(u, prec); v = _arr(v, prec)
    H, W = u.shape
    un, vn = np.empty_like(u), np.empty_like(v)
    e = getattr(lib(), "orc_ls_sweep_" + prec)(*[_p(a) for a in fs], h, H, W, _p(u), _p(v), _p(un), _p(vn))
    return un, vn, e


def ls(im1, im2, h, u, v, maxnum=60, tol=1e-8, prec="f32"):
    """Solver-internal orientation: u = row component, v = column component.
    Returns (u, v, err, iters)."""
    im1 = _arr(im1, prec); im2 = _arr(im2, prec)
    H, W = im1.shape
    u = _arr(u, prec).copy(); v = _arr(v, prec).copy()
    it = C.c_int(0)
    e = getattr(lib(), "orc_ls_" + prec)(_p(im1), _p(im2), H, W, h, int(maxnum), float(tol), _p(u), _p(v), C.byref(it))
    return u, v, e, it.value


def resize(src, H2, W2, filt, prec="f32"):
    """Pillow resize, filt 2 = BILINEAR, 3 = BICUBIC (Farneback_PyCL.py:61-62,
    GenericPyramidalOpticalFlow.py:66-67)."""
    src = _arr(src, prec)
    H, W = src.shape
    dst = np.empty((H2, W2), _dt(prec))
    getattr(lib(), "orc_resize_" + prec)(_p(src), H, W, _p(dst), H2, W2, int(filt))
    return dst


def resize_bicubic(src, H2, W2, prec="f32"):
    src = _arr(src, prec)
    H, W = src.shape
    dst = np.empty((H2, W2), _dt(prec))
    getattr(lib(), "orc_resize_bicubic_" + prec)(_p(src), H, W, _p(dst), H2, W2)
    return dst


def spline_upsample(src, H, W, scale=1.0, prec="f32", want_f64=False):
    src = _arr(src, prec)
    Hc, Wc = src.shape
    dst = np.empty((H, W), _dt(prec))
    d64 = np.empty((H, W), np.float64) if want_f64 else None
    getattr(lib(), "orc_spline_upsample_" + prec)(
        _p(src), Hc, Wc, _p(dst), H, W, scale, d64.ctypes.data_as(_dp) if want_f64 else None)
    return (dst, d64) if want_f64 else dst


def spline_matrix(n, N):
    out = np.empty((N, n), np.float64)
    lib().orc_spline_matrix(n, N, out.ctypes.data_as(_dp))
    return out


def pil_coeffs(in_size, out_size, filt=3):
    ks = C.c_int(0)
    bounds = np.zeros(2 * out_size, np.int32)
    p = lib().orc_pil_coeffs(in_size, out_size, filt, C.byref(ks), bounds.ctypes.data_as(_ip))
    k = np.ctypeslib.as_array(p, shape=(out_size * ks.value,)).copy().reshape(out_size, ks.value)
    C.CDLL(None).free(C.cast(p, C.c_void_p))
    return k, bounds.reshape(out_size, 2)


def warp_symmetric(im1, im2, U, V, prec="f32"):
    im1, im2, U, V = (_arr(a, prec) for a in (im1, im2, U, V))
    H, W = im1.shape
    o1, o2 = np.empty_like(im1), np.empty_like(im2)
    getattr(lib(), "orc_warp_symmetric_" + prec)(_p(im1), _p(im2), _p(U), _p(V), H, W, _p(o1), _p(o2))
    return o1, o2


def pyr_lshs(im1, im2, levels, alphas, niter, filt_w, opt_w, ls_alpha, ls_maxnum=60, ls_tol=1e-8,
             prec="f32"):
    """LS(HS) pyramid.  alphas: per level in the order the HS adapter pops them
    (coarse level first).  filt_w / opt_w: pre-filter weights or None.
    Returns (Uaccum, Vaccum, stats[levels, 3] = hs_err, ls_err, ls_iters)."""
    im1 = _arr(im1, prec); im2 = _arr(im2, prec)
    H, W = im1.shape
    al = _arr(np.asarray(alphas), prec)
    fw = _arr(filt_w if filt_w is not None else np.zeros(1), prec)
    ow = _arr(opt_w if opt_w is not None else np.zeros(1), prec)
    U = np.empty((H, W), _dt(prec)); V = np.empty((H, W), _dt(prec))
    st = np.zeros((levels, 3), np.float64)
    getattr(lib(), "orc_pyr_lshs_" + prec)(
        _p(im1), _p(im2), H, W, levels, _p(al), int(niter), _p(fw), 0 if filt_w is None else len(fw),
        _p(ow), 0 if opt_w is None else len(ow), ls_alpha if ls_alpha else 0.0, int(ls_maxnum),
        float(ls_tol), _p(U), _p(V), st.ctypes.data_as(_dp))
    return U, V, st


# ----------------------------------------------------------------- Farneback
def _decl_fb(L):
    for suf, P, R in (("f32", _fp, C.c_float), ("f64", _dp, C.c_double)):
        getattr(L, "orc_fb_blur_" + suf).argtypes = [P, P, C.c_int, C.c_int, C.c_int, P, C.c_int]
        getattr(L, "orc_fb_polyexp_" + suf).argtypes = [P, C.c_int, C.c_int, P, P, P, C.c_int, P, P]
        getattr(L, "orc_fb_update_matrices_" + suf).argtypes = [P, P, P, P, C.c_int, C.c_int, P]
        getattr(L, "orc_fb_update_flow_" + suf).argtypes = [P, C.c_int, C.c_int, P, P]
        getattr(L, "orc_farneback_" + suf).argtypes = [P, P, C.c_int, C.c_int, P, C.c_int, P, P, P, C.c_int, P, P,
                                                       C.c_int, C.c_int, P, P]
        getattr(L, "orc_fb_box_" + suf).argtypes = [P, P, C.c_int, C.c_int, C.c_int, C.c_int]


def _fblib():
    L = lib()
    if not getattr(L, "_fb_declared", False):
        _decl_fb(L)
        L._fb_declared = True
    return L


def fb_blur(src, half_kernel, planes=1, prec="f32"):
    src = _arr(src, prec)
    H = src.shape[0] // planes
    W = src.shape[1]
    gk = _arr(half_kernel, prec)
    dst = np.empty_like(src)
    getattr(_fblib(), "orc_fb_blur_" + prec)(_p(src), _p(dst), H, W, planes, _p(gk), len(gk) - 1)
    return dst


def fb_polyexp(src, g, xg, xxg, ig, prec="f32"):
    src = _arr(src, prec)
    H, W = src.shape
    g, xg, xxg, ig = (_arr(a, prec) for a in (g, xg, xxg, ig))
    dst = np.empty((5 * H, W), _dt(prec))
    getattr(_fblib(), "orc_fb_polyexp_" + prec)(_p(src), H, W, _p(g), _p(xg), _p(xxg), len(g) - 1, _p(ig), _p(dst))
    return dst


def fb_update_matrices(fx, fy, R0, R1, prec="f32"):
    fx, fy, R0, R1 = (_arr(a, prec) for a in (fx, fy, R0, R1))
    H, W = fx.shape
    M = np.empty((5 * H, W), _dt(prec))
    getattr(_fblib(), "orc_fb_update_matrices_" + prec)(_p(fx), _p(fy), _p(R0), _p(R1), H, W, _p(M))
    return M


def fb_update_flow(M, prec="f32"):
    M = _arr(M, prec)
    H, W = M.shape[0] // 5, M.shape[1]
    fx, fy = np.empty((H, W), _dt(prec)), np.empty((H, W), _dt(prec))
    getattr(_fblib(), "orc_fb_update_flow_" + prec)(_p(M), H, W, _p(fx), _p(fy))
    return fx, fy


def fb_box(src, kh, planes=1, prec="f32"):
    src = _arr(src, prec)
    dst = np.empty_like(src)
    getattr(_fblib(), "orc_fb_box_" + prec)(_p(src), _p(dst), src.shape[0] // planes, src.shape[1], planes, int(kh))
    return dst


def farneback(im1, im2, U, V, consts, prec="f32"):
    """One Farneback compute() at a single internal level.  consts: dict with
    pre_kernel, g, xg, xxg, ig, win_kernel (half kernels) and iters; an int
    box_half in place of win_kernel selects the box window."""
    im1, im2 = _arr(im1, prec), _arr(im2, prec)
    H, W = im1.shape
    fx, fy = _arr(U, prec).copy(), _arr(V, prec).copy()
    pk, g, xg, xxg, ig = (_arr(consts[k], prec) for k in ("pre_kernel", "g", "xg", "xxg", "ig"))
    if consts.get("box_half") is None:
        wk = _arr(consts["win_kernel"], prec)
        wkp, wkh = _p(wk), len(wk) - 1
    else:
        wkp, wkh = None, int(consts["box_half"])
    getattr(_fblib(), "orc_farneback_" + prec)(_p(im1), _p(im2), H, W, _p(pk), len(pk) - 1, _p(g), _p(xg), _p(xxg),
                                                len(g) - 1, _p(ig), wkp, wkh, int(consts["iters"]),
                                                _p(fx), _p(fy))
    return fx, fy


# ----------------------------------------------------------------- dense LK
def farneback_levels(H, W, pyramidal_levels, pyr_scale, min_size=32):
    """Coarser levels kept by Farneback_PyCL.compute (Farneback_PyCL.py:475-482):
    halve by pyr_scale while both sides stay >= min_size, at most
    pyramidal_levels - 1 times."""
    n, scale = 0, 1.0
    while n < pyramidal_levels - 1:
        scale *= pyr_scale
        if W * scale < min_size or H * scale < min_size:
            break
        n += 1
    return n


def farneback_pyramid(im1, im2, U, V, consts, n_levels, pyr_scale, level_kernel, prec="f32"):
    """Farneback compute() over n_levels coarser levels plus the full one
    (Farneback_PyCL.py:507-596, fastPyramids False).  Level k, coarsest
    first: scale = pyr_scale**k (a running product), sigma = (1/scale - 1)/2,
    smoothSize = max(round(5 sigma) | 1, 3); both full-resolution frames are
    blurred with level_kernel(smoothSize, sigma) (a half kernel, centre first)
    and Pillow-BILINEAR resized to (round(H scale), round(W scale)); the
    initial flow is the resized U, V times scale at the first level, else the
    resized previous level times 1/pyr_scale (float32 products); then one
    level runs on the resized frames with no further pre-blur."""
    im1, im2 = _arr(im1, prec), _arr(im2, prec)
    H, W = im1.shape
    dt = _dt(prec)
    flow, first = (_arr(U, prec), _arr(V, prec)), True
    c = dict(consts, pre_kernel=np.ones(1))
    for k in range(n_levels, -1, -1):
        scale = 1.0
        for _ in range(k):
            scale *= pyr_scale
        sigma = (1.0 / scale - 1.0) * 0.5
        smooth = max(int(round(sigma * 5)) | 1, 3)
        h, w = int(round(H * scale)), int(round(W * scale))
        mult = dt(scale if first else 1.0 / pyr_scale)
        fx, fy = (resize(f, h, w, 2, prec) * mult for f in flow)
        pk = level_kernel(smooth, sigma)
        f1 = resize(fb_blur(im1, pk, prec=prec), h, w, 2, prec)
        f2 = resize(fb_blur(im2, pk, prec=prec), h, w, 2, prec)
        flow, first = farneback(f1, f2, fx, fy, c, prec), False
    return flow


def lk_window_weights(hw, lo=0, hi=0):
    """0/1 weight of each of the 32 window-grid columns (or rows), taken from
    the per-work-item weights of pyrlkDenseLargeW.cl:338-371: work-item xid
    holds columns xid + 8 * tile with tile weights (1, wx0, wx1, wx2); the
    program is built with WSX = 1 when the window has 16 columns or more.
    lo / hi: the asymmetric left (top) and right (bottom) flags."""
    win = 2 * hw + 1
    w = np.zeros(32, np.uint8)
    for xid in range(8):
        if win >= 16:
            wx0 = 1
            wx1 = 1 if 2 * 8 + xid < win - hi else 0
            wx2 = 1 if 3 * 8 + xid < win - hi else 0
        else:
            wx0 = 1 if 1 * 8 + xid < win - hi else 0
            wx1 = wx2 = 0
        if xid == 0:
            wx0 = 1 - lo
        for tile, wt in enumerate((1, wx0, wx1, wx2)):
            w[tile * 8 + xid] = wt
    return w


def lk_dense(I, J, u, v, half_win=13, niter=5, prec="f32", asym=(0, 0, 0, 0)):
    """One dense-LK level.  asym: the (left, right, top, bottom) window flags.
    Returns (u, v, status uint8)."""
    L = lib()
    f = getattr(L, "orc_lk_dense_" + prec)
    P = _fp if prec == "f32" else _dp
    B = C.POINTER(C.c_ubyte)
    f.argtypes = [P, P, C.c_int, C.c_int, C.c_int, C.c_int, P, P, B, B, B]
    I, J = _arr(I, prec), _arr(J, prec)
    H, W = I.shape
    u, v = _arr(u, prec).copy(), _arr(v, prec).copy()
    st = np.zeros((H, W), np.uint8)
    wx = lk_window_weights(half_win, asym[0], asym[1])
    wy = lk_window_weights(half_win, asym[2], asym[3])
    f(_p(I), _p(J), H, W, int(half_win), int(niter), _p(u), _p(v), st.ctypes.data_as(B), wx.ctypes.data_as(B),
      wy.ctypes.data_as(B))
    return u, v, st
