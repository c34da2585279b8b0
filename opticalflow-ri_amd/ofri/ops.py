"""Device-array operations: one call per libofri entry point.

Every function takes and returns DeviceArray objects (host numpy arrays are
uploaded first) and runs on the process-wide context's stream.
"""
import ctypes as C
import os

import numpy as np

from . import _capi
from .device import DeviceArray, as_device, context

# Reference-like console output (the reference prints per level and per
# Liu-Shen sweep); OFRI_VERBOSE=0 silences it.
VERBOSE = os.environ.get("OFRI_VERBOSE", "1") != "0"


def log(*args):
    if VERBOSE:
        print(*args)


def _lib():
    return context().lib


def _h(ctx=None):
    """The handle calls go to: ctx (a Context or a lane) or the process-wide context."""
    return (ctx or context()).handle


def gauss_sep(img, weights, out=None):
    """Separable pre-filter with float32 weights (gaussian_filter.py:54-94)."""
    img = as_device(img)
    w = np.ascontiguousarray(weights, dtype=np.float32)
    out = out if out is not None else DeviceArray(img.shape)
    _capi.check(_lib().ofri_gauss_sep(_h(), img.ptr, out.ptr, img.H, img.W, _capi.fptr(w), len(w)))
    return out


BICUBIC = 3
BILINEAR = 2


def resize(img, H2, W2, filt=BICUBIC):
    """Pillow-compatible float-image resize (GenericPyramidalOpticalFlow.py:66-67)."""
    img = as_device(img)
    out = DeviceArray((H2, W2))
    _capi.check(_lib().ofri_resize(_h(), img.ptr, img.H, img.W, out.ptr, int(H2), int(W2), int(filt)))
    return out


def spline_upsample(flow, H, W, scale=1.0):
    """RectBivariateSpline upsample + scaling (GenericPyramidalOpticalFlow.py:151-171)."""
    flow = as_device(flow)
    out = DeviceArray((H, W))
    _capi.check(_lib().ofri_spline_upsample(_h(), flow.ptr, flow.H, flow.W, out.ptr, int(H), int(W),
                                            float(np.float32(scale))))
    return out


def warp_symmetric(im1, im2, U, V):
    """Symmetric bilinear warp (GenericPyramidalOpticalFlow.py:69-115,197-200)."""
    im1, im2, U, V = (as_device(a) for a in (im1, im2, U, V))
    o1, o2 = DeviceArray(im1.shape), DeviceArray(im1.shape)
    _capi.check(_lib().ofri_warp_symmetric(_h(), im1.ptr, im2.ptr, U.ptr, V.ptr, im1.H, im1.W, o1.ptr, o2.ptr))
    return o1, o2


def warp_coords(img, cy, cx):
    """Bilinear sample at float32 coordinates (GenericPyramidalOpticalFlow.py:69-115)."""
    img, cy, cx = (as_device(a) for a in (img, cy, cx))
    out = DeviceArray(img.shape)
    _capi.check(_lib().ofri_warp_coords(_h(), img.ptr, cy.ptr, cx.ptr, img.H, img.W, out.ptr))
    return out


def accumulate(acc, x):
    _capi.check(_lib().ofri_accumulate(_h(), acc.ptr, x.ptr, int(np.prod(acc.shape))))
    return acc


def hs(im1, im2, alpha, niter, U, V):
    """Horn-Schunck; U, V are updated in place.  Returns the reference's error."""
    im1, im2 = as_device(im1), as_device(im2)
    err = C.c_float()
    _capi.check(_lib().ofri_hs(_h(), im1.ptr, im2.ptr, im1.H, im1.W, float(np.float32(alpha)), int(niter),
                               U.ptr, V.ptr, C.byref(err)))
    return np.float32(err.value)


def hs_tol(im1, im2, alpha, maxnum, tol, U, V, hist=None):
    """Horn-Schunck until the per-sweep update norm <= tol or maxnum sweeps
    (tol < 0: exactly maxnum); U, V updated in place.  hist: optional device
    array of >= maxnum floats receiving each sweep's norm.  Returns (err, sweeps)."""
    im1, im2 = as_device(im1), as_device(im2)
    if hist is not None and hist.nbytes // 4 < maxnum:
        raise ValueError("hs_tol: history buffer holds %d < %d sweeps" % (hist.nbytes // 4, maxnum))
    err = C.c_float()
    it = C.c_int()
    _capi.check(_lib().ofri_hs_tol(_h(), im1.ptr, im2.ptr, im1.H, im1.W, float(np.float32(alpha)), int(maxnum),
                                   float(tol), U.ptr, V.ptr, hist.ptr if hist is not None else None,
                                   C.byref(err), C.byref(it)))
    return np.float32(err.value), it.value


def render_pair(H, W, family, density, seed, par, modes=None, noise=2.0, im1=None, im2=None, truth=False):
    """Synthetic PIV pair rendered on the device (ofri_render_pair).  family 0
    Poiseuille (par: umax), 1 Lamb-Oseen (cx, cy, rc, vmax), 2 shear +
    turbulence (shear, dmax; modes: (4, n) array kx, ky, phase, amp).
    Returns (im1, im2) or (im1, im2, u, v) DeviceArrays."""
    par = np.ascontiguousarray(par, dtype=np.float32)
    md = None if modes is None else np.ascontiguousarray(modes, dtype=np.float32)
    im1 = im1 if im1 is not None else DeviceArray((H, W))
    im2 = im2 if im2 is not None else DeviceArray((H, W))
    u = DeviceArray((H, W)) if truth else None
    v = DeviceArray((H, W)) if truth else None
    _capi.check(_lib().ofri_render_pair(_h(), H, W, int(family), float(density), int(seed) & ((1 << 64) - 1),
                                        _capi.fptr(par), par.size, _capi.fptr(md) if md is not None else None,
                                        0 if md is None else md.shape[-1], float(noise), im1.ptr, im2.ptr,
                                        u.ptr if truth else None, v.ptr if truth else None))
    return (im1, im2, u, v) if truth else (im1, im2)


def hs_accel(im1, im2, alpha, maxnum, tol, bound, U, V, hist=None, finish=0.0, launch_test=False, relative=False):
    """Chebyshev-accelerated Horn-Schunck to the plain residual tol
    (ofri_hs_accel; finish > 1 ends with plain sweeps from finish * tol;
    launch_test: stop test at each launch's last sweep only; relative (with
    launch_test): tol scales with the tested result's norm (||U|| + ||V||)/N);
    U, V in place.  Returns (err, sweeps)."""
    im1, im2 = as_device(im1), as_device(im2)
    if hist is not None and hist.nbytes // 4 < maxnum:
        raise ValueError("hs_accel: history buffer holds %d < %d sweeps" % (hist.nbytes // 4, maxnum))
    err = C.c_float()
    it = C.c_int()
    _capi.check(_lib().ofri_hs_accel(_h(), im1.ptr, im2.ptr, im1.H, im1.W, float(np.float32(alpha)), int(maxnum),
                                     float(tol), float(bound), float(finish), U.ptr, V.ptr,
                                     hist.ptr if hist is not None else None,
                                     C.byref(err), C.byref(it),
                                     int(bool(launch_test)) | (2 if relative else 0)))
    return np.float32(err.value), it.value


def ls(im1, im2, h, u, v, maxnum=60, tol=1e-8):
    """Liu-Shen in the solver's orientation (u = row component); u, v in place.
    Returns (err, sweeps)."""
    im1, im2 = as_device(im1), as_device(im2)
    err = C.c_float()
    it = C.c_int()
    _capi.check(_lib().ofri_ls(_h(), im1.ptr, im2.ptr, im1.H, im1.W, float(np.float32(h)), int(maxnum),
                               float(tol), u.ptr, v.ptr, C.byref(err), C.byref(it)))
    return np.float32(err.value), it.value


def scipy_gaussian_weights(sigma, truncate):
    """scipy.ndimage's float64 Gaussian weights (radius int(truncate * sigma +
    0.5), normalised), formed with numpy as scipy forms them."""
    r = int(truncate * float(sigma) + 0.5)
    x = np.arange(-r, r + 1)
    phi = np.exp(-0.5 / (float(sigma) * float(sigma)) * x ** 2)
    return phi / phi.sum()


def warp_liushen(im1, u, v, sigma=1.8, truncate=20.0):
    """ofri_warp_liushen: the Liu-Shen warping mode of the first frame by
    (u, v) (GenericPyramidalOpticalFlow.py:188-220).  Raises IndexError where
    the reference's pixel scatter would."""
    im1, u, v = as_device(im1), as_device(u), as_device(v)
    H, W = im1.shape[-2:]
    w = scipy_gaussian_weights(sigma, truncate)
    r = (len(w) - 1) // 2
    half = np.ascontiguousarray(w[r:], dtype=np.float64)
    out = DeviceArray((H, W))
    st = C.c_int(0)
    _capi.check(_lib().ofri_warp_liushen(_h(), im1.ptr, H, W, u.ptr, v.ptr,
                                         half.ctypes.data_as(C.POINTER(C.c_double)), r, out.ptr, C.byref(st)))
    if st.value:
        raise IndexError("Liu-Shen warping: a pixel's displaced position lies outside the %dx%d image" % (H, W))
    return out


def ls_learned(im1, im2, h, u, v, maxnum=24, tol=1e-8):
    """ofri_ls_learned: Liu-Shen tested once per fused launch (its last sweep,
    the reference update) with float32 sweeps in between; u, v in place.
    Returns (err, sweeps)."""
    im1, im2 = as_device(im1), as_device(im2)
    err = C.c_float()
    it = C.c_int()
    _capi.check(_lib().ofri_ls_learned(_h(), im1.ptr, im2.ptr, im1.H, im1.W, float(np.float32(h)), int(maxnum),
                                       float(tol), u.ptr, v.ptr, C.byref(err), C.byref(it)))
    return np.float32(err.value), it.value


def fb_blur(img, half_kernel, out=None):
    """Reflect-101 separable Gaussian blur (Farneback_PyCL.py:210-240) with a
    half kernel (centre first); returns a new device array unless out given."""
    img = as_device(img)
    k = np.ascontiguousarray(half_kernel, dtype=np.float32)
    out = out if out is not None else DeviceArray((img.H, img.W))
    _capi.check(_lib().ofri_fb_blur(_h(), img.ptr, img.H, img.W, out.ptr, _capi.fptr(k), len(k) - 1))
    return out


def _fb_params(consts, flow_scale=1.0):
    """FbParams plus the arrays it points into (kept alive by the caller)."""
    box = consts.get("box_half")
    arrs = {k: np.ascontiguousarray(consts[k], dtype=np.float32)
            for k in ("g", "xg", "xxg", "ig", "pre_kernel", "win_kernel") if k in consts}
    p = _capi.FbParams()
    p.poly_n = len(arrs["g"]) - 1
    p.g, p.xg, p.xxg, p.ig = (_capi.fptr(arrs[k]) for k in ("g", "xg", "xxg", "ig"))
    p.pre_kernel = _capi.fptr(arrs["pre_kernel"])
    p.pre_kh = len(arrs["pre_kernel"]) - 1
    if box is None:
        p.win_kernel = _capi.fptr(arrs["win_kernel"])
        p.win_kh = len(arrs["win_kernel"]) - 1
    else:
        p.box_filter, p.win_kh = 1, int(box)
    p.iters = int(consts["iters"])
    p.flow_scale = float(np.float32(flow_scale))
    return p, arrs


def pyr_generic(im1, im2, levels, filt_w, opt_w, warping, inter_scaling, final_scaling, ls_alpha=0.0,
                ls_maxnum=60, ls_tol=1e-8, fb_consts=None, lk=None, Uacc=None, Vacc=None, status=None, stats=True):
    """The generic pyramid on the device (ofri_pyr_generic) with a Farneback
    (fb_consts: the adapter's constants, one internal level) or dense-LK (lk:
    (halfWindow, Niter)) main solver and an optional Liu-Shen stage
    (ls_alpha > 0).  status: optional ByteArray receiving the last level's LK
    status.  Returns (Uacc, Vacc, stats)."""
    im1, im2 = as_device(im1), as_device(im2)
    H, W = im1.shape[-2:]
    if (fb_consts is None) == (lk is None):
        raise ValueError("pyr_generic: give exactly one of fb_consts and lk")
    fw = None if filt_w is None else np.ascontiguousarray(filt_w, dtype=np.float32)
    ow = None if opt_w is None else np.ascontiguousarray(opt_w, dtype=np.float32)
    p = _capi.PyrParams()
    p.levels = int(levels)
    p.filt_w = _capi.fptr(fw) if fw is not None else None
    p.filt_n = 0 if fw is None else len(fw)
    p.opt_w = _capi.fptr(ow) if ow is not None else None
    p.opt_n = 0 if ow is None else len(ow)
    p.warping, p.inter_scaling, p.final_scaling = int(bool(warping)), int(bool(inter_scaling)), int(bool(final_scaling))
    p.ls_alpha = float(ls_alpha) if ls_alpha else 0.0
    p.ls_maxnum = int(ls_maxnum)
    p.ls_tol = float(ls_tol)
    m = _capi.PyrMain()
    keep = None
    if fb_consts is not None:
        fbp, keep = _fb_params(fb_consts)
        m.kind = 0
        m.fb = C.pointer(fbp)
    else:
        m.kind = 1
        m.lk_half_win, m.lk_niter = int(lk[0]), int(lk[1])
        m.d_lk_status = status.ptr if status is not None else None
    Uacc = Uacc if Uacc is not None else DeviceArray((H, W))
    Vacc = Vacc if Vacc is not None else DeviceArray((H, W))
    st = (_capi.LevelStats * levels)() if stats else None
    _capi.check(_lib().ofri_pyr_generic(_h(), im1.ptr, im2.ptr, H, W, C.byref(p), C.byref(m), Uacc.ptr, Vacc.ptr,
                                        st))
    del keep
    return Uacc, Vacc, (_stats_dicts(st) if stats else [])


def farneback(im1, im2, flowx, flowy, consts, flow_scale=1.0):
    """One Farneback level (Farneback_PyCL.py:515-596); flowx / flowy device
    arrays updated in place, multiplied by float32(flow_scale) first.
    consts: dict of float32 arrays g, xg, xxg (polyN+1 taps), ig (4),
    pre_kernel, win_kernel (half kernels, centre first), and the int iters;
    an int box_half instead of win_kernel selects the box window
    (Farneback_PyCL.py:424-439)."""
    im1, im2 = as_device(im1), as_device(im2)
    p, arrs = _fb_params(consts, flow_scale)
    _capi.check(_lib().ofri_farneback(_h(), im1.ptr, im2.ptr, im1.H, im1.W, flowx.ptr, flowy.ptr, C.byref(p)))
    return flowx, flowy


def lk_dense(I, J, u, v, half_win=13, niter=5, status=None, asym=None):
    """One dense-LK level (denseLucasKanade_PyCL.py:113-169); u, v device
    arrays updated in place; status: optional device buffer of H*W bytes;
    asym: optional [left, right, top, bottom] window flags (0/1)."""
    I, J = as_device(I), as_device(J)
    a = (C.c_int * 4)(*[int(x) for x in asym]) if asym is not None else None
    _capi.check(_lib().ofri_lk_dense(_h(), I.ptr, J.ptr, I.H, I.W, u.ptr, v.ptr, int(half_win), int(niter),
                                     status.ptr if status is not None else None, a))
    return u, v


class ByteArray:
    """An H x W uint8 device buffer (status masks)."""

    def __init__(self, shape):
        self.ctx = context()
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape))
        p = C.c_void_p()
        _capi.check(self.ctx.lib.ofri_malloc(self.ctx.handle, self.nbytes, C.byref(p)))
        self.ptr = p

    def download(self):
        out = np.empty(self.shape, np.uint8)
        _capi.check(self.ctx.lib.ofri_download(self.ctx.handle, out.ctypes.data_as(C.c_void_p), self.ptr,
                                               self.nbytes))
        return out

    def __del__(self):
        try:
            if self.ptr:
                self.ctx.lib.ofri_free(self.ctx.handle, self.ptr)
                self.ptr = None
        except Exception:
            pass


def _lshs_params(levels, alphas, niter, filt_w, opt_w, ls_alpha, ls_maxnum, ls_tol):
    """LshsParams plus the arrays it points into (kept alive by the caller)."""
    al = np.ascontiguousarray(alphas, dtype=np.float32)
    if len(al) < levels:
        raise ValueError("need one alpha per level")
    fw = None if filt_w is None else np.ascontiguousarray(filt_w, dtype=np.float32)
    ow = None if opt_w is None else np.ascontiguousarray(opt_w, dtype=np.float32)
    p = _capi.LshsParams()
    p.levels = int(levels)
    p.alphas = _capi.fptr(al)
    p.niter = int(niter)
    p.filt_w = _capi.fptr(fw) if fw is not None else None
    p.filt_n = 0 if fw is None else len(fw)
    p.opt_w = _capi.fptr(ow) if ow is not None else None
    p.opt_n = 0 if ow is None else len(ow)
    p.ls_alpha = float(ls_alpha) if ls_alpha else 0.0
    p.ls_maxnum = int(ls_maxnum)
    p.ls_tol = float(ls_tol)
    return p, (al, fw, ow)


def batch_lshs(pairs, levels, alphas, niter, filt_w, opt_w, ls_alpha, ls_maxnum=60, ls_tol=1e-8, out=None,
               stats=True):
    """LS(HS) pyramids of many pairs of any sizes in one call (ofri_batch_lshs).
    pairs: sequence of (im1, im2) (numpy or DeviceArray).  out: optional list of
    (U, V) DeviceArrays.  Returns (list of (U, V), per-pair list of per-level stats)."""
    dev = [(as_device(a), as_device(b)) for a, b in pairs]
    n = len(dev)
    if n == 0:
        raise ValueError("batch_lshs: no pairs")
    for a, b in dev:
        if a.shape != b.shape:
            raise ValueError("images differ in shape: %s vs %s" % (a.shape, b.shape))
    if out is None:
        out = [(DeviceArray(a.shape), DeviceArray(a.shape)) for a, _ in dev]
    p, keep = _lshs_params(levels, alphas, niter, filt_w, opt_w, ls_alpha, ls_maxnum, ls_tol)
    tab = (_capi.Pair * n)()
    for i, ((a, b), (U, V)) in enumerate(zip(dev, out)):
        tab[i].im1, tab[i].im2 = a.ptr, b.ptr
        tab[i].H, tab[i].W = a.shape
        tab[i].U, tab[i].V = U.ptr, V.ptr
    st = (_capi.LevelStats * (n * levels))() if stats else None
    _capi.check(_lib().ofri_batch_lshs(_h(), tab, n, C.byref(p), st))
    res = []
    if stats:
        for i in range(n):
            res.append(_stats_dicts(st[i * levels:(i + 1) * levels]))
    return out, res


def _stats_dicts(st):
    return [dict(H=s.H, W=s.W, hs_err=np.float32(s.hs_err), ls_err=np.float32(s.ls_err), ls_iters=s.ls_iters,
                 hs_res=np.float32(s.hs_res), hs_iters=s.hs_iters, hs_nrm=np.float32(s.hs_nrm),
                 ls_nrm=np.float32(s.ls_nrm)) for s in st]


def flow_epe(U1, V1, U2, V2, out=None, ctx=None):
    """Mean endpoint error of two device flows, written to a one-word device
    array (ofri_flow_epe) with no synchronisation.  Returns that array.
    ctx: the context or lane queued on (default: the process-wide one)."""
    out = out if out is not None else DeviceArray((1,))
    n = int(np.prod(U1.shape))
    for a in (V1, U2, V2):
        if int(np.prod(a.shape)) != n:
            raise ValueError("flow_epe: flows differ in size")
    _capi.check(_lib().ofri_flow_epe(_h(ctx), U1.ptr, V1.ptr, U2.ptr, V2.ptr, n, out.ptr))
    return out


def pyr_lshs_taps(im1, im2, levels, alphas, niter, filt_w, opt_w, ls_alpha, ls_maxnum=60, ls_tol=1e-8,
                  Uacc=None, Vacc=None, taps=None, ls_tap_back=None):
    """ofri_pyr_lshs_taps: the pyramid with every level's HS result and Liu-Shen
    correction exported.  Returns (Uacc, Vacc, taps, stats); taps is a flat
    DeviceArray holding, per level, 4 planes of the level's size.
    ls_tap_back: optional per-level k, the Liu-Shen planes taken k sweeps
    before Liu-Shen's stop."""
    im1, im2 = as_device(im1), as_device(im2)
    H, W = im1.shape[-2:]
    p, keep = _lshs_params(levels, alphas, niter, filt_w, opt_w, ls_alpha, ls_maxnum, ls_tol)
    sizes = level_sizes(H, W, levels)
    n = 4 * sum(h * w for h, w in sizes)
    taps = taps if taps is not None else DeviceArray((n,))
    if taps.nbytes < 4 * n:
        raise ValueError("taps buffer holds %d floats, %d needed" % (taps.nbytes // 4, n))
    Uacc = Uacc if Uacc is not None else DeviceArray((H, W))
    Vacc = Vacc if Vacc is not None else DeviceArray((H, W))
    st = (_capi.LevelStats * levels)()
    back = None
    if ls_tap_back is not None:
        if len(ls_tap_back) != levels:
            raise ValueError("ls_tap_back: one offset per level")
        back = (C.c_int * levels)(*[int(k) for k in ls_tap_back])
    _capi.check(_lib().ofri_pyr_lshs_taps(_h(), im1.ptr, im2.ptr, H, W, C.byref(p), Uacc.ptr, Vacc.ptr,
                                          taps.ptr, back, st))
    return Uacc, Vacc, taps, _stats_dicts(st)


def pyr_lshs_learned(im1, im2, levels, alphas, niter, filt_w, opt_w, ls_alpha, net, hs_tol, ls_tol,
                     hs_budget=600, ls_budget=60, cheb_bound=0.99, cheb_finish=4.0, ls_cheb=(0.0, 0.0), ls_maxnum=60,
                     ls_tol_plain=1e-8, Uacc=None, Vacc=None, stats=True, flag=None, hs_finish_budget=0,
                     relative=False, ls_sweeps=0, hs_finish_first=0, ctx=None):
    """The learned form (ofri_pyr_lshs_learned): accelerated HS, the
    correction network and Liu-Shen run to the per-level residual targets
    hs_tol / ls_tol (relative: as multiples of the tested iterate's norm
    (||u|| + ||v||)/N).  ls_sweeps > 0: exactly that many Liu-Shen sweeps,
    ls_tol then accepts or rejects the level.  net: an ofri.net.Net or None
    (the method without it).  hs_budget, cheb_bound, cheb_finish,
    hs_finish_budget, hs_finish_first, ls_sweeps: one value or (coarse,
    fine).  flag: optional
    DeviceArray view of one 4-byte word receiving the convergence flag (int
    1 / 0) with no synchronisation.  ctx: the context or lane queued on
    (default: the process-wide one).  Returns (Uacc, Vacc, stats,
    converged); converged is None when stats is False."""
    im1, im2 = as_device(im1), as_device(im2)
    H, W = im1.shape[-2:]
    p, keep = _lshs_params(levels, alphas, niter, filt_w, opt_w, ls_alpha, ls_maxnum, ls_tol_plain)
    ht = np.ascontiguousarray(hs_tol, dtype=np.float32)
    lt = np.ascontiguousarray(ls_tol, dtype=np.float32)
    def pair(x):
        return (x, x) if np.isscalar(x) else tuple(x)
    hb, cb, cf, fb, lsw, ff = (pair(hs_budget), pair(cheb_bound), pair(cheb_finish), pair(hs_finish_budget),
                               pair(ls_sweeps), pair(hs_finish_first))
    lp = _capi.LearnedParams(_capi.fptr(ht), _capi.fptr(lt), (C.c_int * 2)(int(hb[0]), int(hb[1])), int(ls_budget),
                             (C.c_float * 2)(float(cb[0]), float(cb[1])), (C.c_float * 2)(float(cf[0]), float(cf[1])),
                             float(ls_cheb[0]), float(ls_cheb[1]), (C.c_int * 2)(int(fb[0]), int(fb[1])),
                             int(bool(relative)), (C.c_int * 2)(int(lsw[0]), int(lsw[1])),
                             (C.c_int * 2)(int(ff[0]), int(ff[1])))
    Uacc = Uacc if Uacc is not None else DeviceArray((H, W))
    Vacc = Vacc if Vacc is not None else DeviceArray((H, W))
    st = (_capi.LevelStats * levels)() if stats else None
    conv = C.c_int(0)
    _capi.check(_lib().ofri_pyr_lshs_learned(_h(ctx), im1.ptr, im2.ptr, H, W, C.byref(p), C.byref(lp),
                                             net.handle if net is not None else None, Uacc.ptr, Vacc.ptr, st,
                                             C.byref(conv) if stats else None,
                                             flag.ptr if flag is not None else None))
    return Uacc, Vacc, (_stats_dicts(st) if stats else []), (bool(conv.value) if stats else None)


def level_sizes(H, W, levels):
    """Level sizes of the pyramid, coarse to fine (GenericPyramidalOpticalFlow.py:291,335-342)."""
    out = []
    scale = 1.0 / (2.0 ** (levels - 1))
    for level in range(1, levels + 1):
        if scale < 1.0 and level != levels:
            out.append((int(np.int32(np.round(H * scale))), int(np.int32(np.round(W * scale)))))
        else:
            out.append((H, W))
        scale *= 2
    return out


def pyr_lshs(im1, im2, levels, alphas, niter, filt_w, opt_w, ls_alpha, ls_maxnum=60, ls_tol=1e-8,
             Uacc=None, Vacc=None, stats=True, ctx=None):
    """Whole LS(HS) pyramid on the device.  ctx: the context or lane queued
    on (default: the process-wide one).  Returns (Uacc, Vacc, stats list)."""
    im1, im2 = as_device(im1), as_device(im2)
    H, W = im1.shape
    al = np.ascontiguousarray(alphas, dtype=np.float32)
    if len(al) < levels:
        raise ValueError("need one alpha per level")
    fw = None if filt_w is None else np.ascontiguousarray(filt_w, dtype=np.float32)
    ow = None if opt_w is None else np.ascontiguousarray(opt_w, dtype=np.float32)
    p = _capi.LshsParams()
    p.levels = int(levels)
    p.alphas = _capi.fptr(al)
    p.niter = int(niter)
    p.filt_w = _capi.fptr(fw) if fw is not None else None
    p.filt_n = 0 if fw is None else len(fw)
    p.opt_w = _capi.fptr(ow) if ow is not None else None
    p.opt_n = 0 if ow is None else len(ow)
    p.ls_alpha = float(ls_alpha) if ls_alpha else 0.0
    p.ls_maxnum = int(ls_maxnum)
    p.ls_tol = float(ls_tol)
    Uacc = Uacc if Uacc is not None else DeviceArray((H, W))
    Vacc = Vacc if Vacc is not None else DeviceArray((H, W))
    st = (_capi.LevelStats * levels)() if stats else None
    _capi.check(_lib().ofri_pyr_lshs(_h(ctx), im1.ptr, im2.ptr, H, W, C.byref(p), Uacc.ptr, Vacc.ptr,
                                     st if stats else None))
    out = _stats_dicts(st) if stats else []
    return Uacc, Vacc, out
