"""ctypes binding of libofri.so (the C ABI declared in include/ofri.h).

The library is the only compute path: if it is missing or no GPU is present,
loading raises OfriError.  There is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_DIR, "libofri.so")

OFRI_OK = 0
OFRI_E_ARG = -1
OFRI_E_HIP = -2
OFRI_E_NOMEM = -3
OFRI_E_STATE = -4
OFRI_E_NUMERIC = -5


class OfriError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("libofri error %d: %s" % (code, msg))
        self.code = code


class LshsParams(C.Structure):
    _fields_ = [
        ("levels", C.c_int),
        ("alphas", C.POINTER(C.c_float)),
        ("niter", C.c_int),
        ("filt_w", C.POINTER(C.c_float)),
        ("filt_n", C.c_int),
        ("opt_w", C.POINTER(C.c_float)),
        ("opt_n", C.c_int),
        ("ls_alpha", C.c_float),
        ("ls_maxnum", C.c_int),
        ("ls_tol", C.c_double),
    ]


class LevelStats(C.Structure):
    _fields_ = [("H", C.c_int), ("W", C.c_int), ("hs_err", C.c_float), ("ls_err", C.c_float),
                ("ls_iters", C.c_int), ("hs_res", C.c_float), ("hs_iters", C.c_int), ("hs_nrm", C.c_float),
                ("ls_nrm", C.c_float)]


class Pair(C.Structure):
    _fields_ = [("im1", C.c_void_p), ("im2", C.c_void_p), ("H", C.c_int), ("W", C.c_int),
                ("U", C.c_void_p), ("V", C.c_void_p)]


class LearnedParams(C.Structure):
    _fields_ = [("hs_tol", C.POINTER(C.c_float)), ("ls_tol", C.POINTER(C.c_float)), ("hs_budget", C.c_int * 2),
                ("ls_budget", C.c_int), ("cheb_bound", C.c_float * 2), ("cheb_finish", C.c_float * 2),
                ("ls_cheb_lo", C.c_float), ("ls_cheb_hi", C.c_float), ("hs_finish_budget", C.c_int * 2),
                ("relative", C.c_int), ("ls_sweeps", C.c_int * 2), ("hs_finish_first", C.c_int * 2)]


class FbParams(C.Structure):
    _fields_ = [
        ("poly_n", C.c_int),
        ("g", C.POINTER(C.c_float)),
        ("xg", C.POINTER(C.c_float)),
        ("xxg", C.POINTER(C.c_float)),
        ("ig", C.POINTER(C.c_float)),
        ("pre_kernel", C.POINTER(C.c_float)),
        ("pre_kh", C.c_int),
        ("win_kernel", C.POINTER(C.c_float)),
        ("win_kh", C.c_int),
        ("iters", C.c_int),
        ("flow_scale", C.c_float),
        ("box_filter", C.c_int),
    ]


class PyrParams(C.Structure):
    _fields_ = [
        ("levels", C.c_int),
        ("filt_w", C.POINTER(C.c_float)),
        ("filt_n", C.c_int),
        ("opt_w", C.POINTER(C.c_float)),
        ("opt_n", C.c_int),
        ("warping", C.c_int),
        ("inter_scaling", C.c_int),
        ("final_scaling", C.c_int),
        ("ls_alpha", C.c_float),
        ("ls_maxnum", C.c_int),
        ("ls_tol", C.c_double),
    ]


class PyrMain(C.Structure):
    _fields_ = [("kind", C.c_int), ("fb", C.POINTER(FbParams)), ("lk_half_win", C.c_int), ("lk_niter", C.c_int),
                ("d_lk_status", C.c_void_p)]


class Tuning(C.Structure):
    _fields_ = [("hs_depth", C.c_int), ("ls_depth", C.c_int), ("hs_variant", C.c_int), ("ls_variant", C.c_int)]


class TuningClass(C.Structure):
    _fields_ = [("min_px", C.c_longlong), ("t", Tuning)]


_vp = C.c_void_p
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_sz = C.c_size_t

# name -> argtypes; every function returns int status
SIGNATURES = {
    "ofri_version": [],
    "ofri_create": [C.c_int, C.POINTER(_vp)],
    "ofri_destroy": [_vp],
    "ofri_sync": [_vp],
    "ofri_lane": [_vp, C.c_int, C.POINTER(_vp)],
    "ofri_fork": [_vp, C.c_int],
    "ofri_join": [_vp, C.c_int],
    "ofri_malloc": [_vp, _sz, C.POINTER(_vp)],
    "ofri_free": [_vp, _vp],
    "ofri_upload": [_vp, _vp, _vp, _sz],
    "ofri_download": [_vp, _vp, _vp, _sz],
    "ofri_copy": [_vp, _vp, _vp, _sz],
    "ofri_memset": [_vp, _vp, C.c_int, _sz],
    "ofri_timer_start": [_vp],
    "ofri_timer_stop": [_vp, _fp],
    "ofri_gauss_sep": [_vp, _vp, _vp, C.c_int, C.c_int, _fp, C.c_int],
    "ofri_resize": [_vp, _vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int, C.c_int],
    "ofri_spline_upsample": [_vp, _vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int, C.c_float],
    "ofri_warp_symmetric": [_vp, _vp, _vp, _vp, _vp, C.c_int, C.c_int, _vp, _vp],
    "ofri_warp_coords": [_vp, _vp, _vp, _vp, C.c_int, C.c_int, _vp],
    "ofri_accumulate": [_vp, _vp, _vp, _sz],
    "ofri_flow_epe": [_vp, _vp, _vp, _vp, _vp, _sz, _vp],
    "ofri_hs": [_vp, _vp, _vp, C.c_int, C.c_int, C.c_float, C.c_int, _vp, _vp, _fp],
    "ofri_hs_tol": [_vp, _vp, _vp, C.c_int, C.c_int, C.c_float, C.c_int, C.c_double, _vp, _vp, _vp, _fp,
                    _ip],
    "ofri_hs_accel": [_vp, _vp, _vp, C.c_int, C.c_int, C.c_float, C.c_int, C.c_double, C.c_float, C.c_float, _vp,
                      _vp, _vp, _fp, _ip, C.c_int],
    "ofri_ls": [_vp, _vp, _vp, C.c_int, C.c_int, C.c_float, C.c_int, C.c_double, _vp, _vp, _fp, _ip],
    "ofri_ls_noalloc": [_vp, _vp, _vp, C.c_int, C.c_int, C.c_float, C.c_int, C.c_double, _vp, _vp, _fp, _ip],
    "ofri_warp_liushen": [_vp, _vp, C.c_int, C.c_int, _vp, _vp, C.POINTER(C.c_double), C.c_int, _vp, _ip],
    "ofri_ls_learned": [_vp, _vp, _vp, C.c_int, C.c_int, C.c_float, C.c_int, C.c_double, _vp, _vp, _fp, _ip],
    "ofri_lk_dense": [_vp, _vp, _vp, C.c_int, C.c_int, _vp, _vp, C.c_int, C.c_int, _vp, C.POINTER(C.c_int)],
    "ofri_farneback": [_vp, _vp, _vp, C.c_int, C.c_int, _vp, _vp, C.POINTER(FbParams)],
    "ofri_fb_blur": [_vp, _vp, C.c_int, C.c_int, _vp, _fp, C.c_int],
    "ofri_batch_lshs": [_vp, C.POINTER(Pair), C.c_int, C.POINTER(LshsParams), C.POINTER(LevelStats)],
    "ofri_render_pair": [_vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_ulonglong, _fp, C.c_int, _fp, C.c_int,
                         C.c_float, _vp, _vp, _vp, _vp],
    "ofri_net_create": [_vp, C.c_int, C.POINTER(_vp)],
    "ofri_net_destroy": [_vp],
    "ofri_net_num_params": [_vp, C.POINTER(C.c_size_t)],
    "ofri_net_flops": [_vp, C.c_int, C.c_int, C.POINTER(C.c_double)],
    "ofri_net_correct": [_vp, _vp, C.POINTER(_vp), C.c_int, C.c_int, _vp, _vp],
    "ofri_net_set_params": [_vp, _vp, _fp, C.c_size_t],
    "ofri_net_init": [_vp, _vp, C.c_ulonglong],
    "ofri_net_get_params": [_vp, _vp, _fp, C.c_size_t],
    "ofri_net_get_grads": [_vp, _vp, _fp, C.c_size_t],
    "ofri_net_predict": [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp, _vp],
    "ofri_net_train": [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _fp, C.c_float, _fp],
    "ofri_pyr_lshs_taps": [_vp, _vp, _vp, C.c_int, C.c_int, C.POINTER(LshsParams), _vp, _vp, _vp, _ip,
                           C.POINTER(LevelStats)],
    "ofri_pyr_lshs_learned": [_vp, _vp, _vp, C.c_int, C.c_int, C.POINTER(LshsParams), C.POINTER(LearnedParams),
                              _vp, _vp, _vp, C.POINTER(LevelStats), _ip, _vp],
    "ofri_pyr_lshs": [_vp, _vp, _vp, C.c_int, C.c_int, C.POINTER(LshsParams), _vp, _vp,
                      C.POINTER(LevelStats)],
    "ofri_pyr_generic": [_vp, _vp, _vp, C.c_int, C.c_int, C.POINTER(PyrParams), C.POINTER(PyrMain), _vp, _vp,
                         C.POINTER(LevelStats)],
    "ofri_set_tuning": [_vp, C.POINTER(Tuning)],
    "ofri_get_tuning": [_vp, C.POINTER(Tuning)],
    "ofri_query_tuning": [_vp, C.c_int, C.c_int, C.POINTER(Tuning)],
    "ofri_set_tuning_classes": [_vp, C.c_void_p, C.c_int],
    "ofri_arch": [_vp, C.c_char_p, C.c_int],
    "ofri_set_graphs": [_vp, C.c_int],
}

_lib = None


def load(path=None):
    """Load libofri.so and declare every entry point; raises OfriError if absent."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise OfriError(OFRI_E_STATE, "libofri.so not built (%s); run __graft_entry__.build()" % path)
    lib = C.CDLL(path)
    lib.ofri_last_error.restype = C.c_char_p
    lib.ofri_last_error.argtypes = []
    for name, args in SIGNATURES.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    _lib = lib
    return lib


def check(status):
    if status != OFRI_OK:
        msg = _lib.ofri_last_error().decode(errors="replace") if _lib is not None else ""
        raise OfriError(status, msg)
    return status


def fptr(a):
    """float* of a C-contiguous float32 numpy array."""
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_fp)
