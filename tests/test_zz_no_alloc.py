"""The allocation-free Liu-Shen entry point (ofri_ls_noalloc) against ofri_ls,
both through the C ABI: the same flow bits, norm and sweep count on two
calls in a row (the second on a smaller image, with the first call's data
still in the shared device array), and a clear error above the array's
maximum size."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(H, W, seed):
    from ofri import ops
    a, b = ops.render_pair(H, W, 1, 0.05, seed, [W / 2.0, H / 2.0, 0.08 * min(H, W), 3.0])
    return a.download(), b.download()


def _call(fn, a, b, h, maxnum, tol, u0, v0):
    from ofri import _capi
    from ofri.device import DeviceArray, context
    ctx = context()
    da, db = DeviceArray.from_host(a), DeviceArray.from_host(b)
    du, dv = DeviceArray.from_host(u0), DeviceArray.from_host(v0)
    err, it = C.c_float(), C.c_int()
    st = getattr(ctx.lib, fn)(ctx.handle, da.ptr, db.ptr, a.shape[0], a.shape[1], float(h), int(maxnum), float(tol),
                              du.ptr, dv.ptr, C.byref(err), C.byref(it))
    return st, du.download(), dv.download(), err.value, it.value


def test_noalloc_ls_matches_ofri_ls_on_two_calls():
    from ofri import _capi
    rng = np.random.default_rng(7)
    for k, (H, W, maxnum) in enumerate(((512, 384, 60), (200, 136, 23))):
        a, b = _pair(H, W, 100 + k)
        u0 = (rng.standard_normal((H, W)) * 0.3).astype(np.float32)
        v0 = (rng.standard_normal((H, W)) * 0.3).astype(np.float32)
        st_r, ur, vr, er, ir = _call("ofri_ls", a, b, 5.0, maxnum, 1e-8, u0, v0)
        st_n, un, vn, en, i_n = _call("ofri_ls_noalloc", a, b, 5.0, maxnum, 1e-8, u0, v0)
        assert st_r == _capi.OFRI_OK and st_n == _capi.OFRI_OK
        print("[ls_noalloc %dx%d] sweeps %d / %d, norm %.6g / %.6g" % (H, W, i_n, ir, en, er))
        assert i_n == ir == maxnum and en == er
        assert np.array_equal(un, ur) and np.array_equal(vn, vr)
        assert np.isfinite(un).all() and not np.array_equal(un, u0)


def test_noalloc_ls_rejects_images_above_its_maximum():
    from ofri import _capi
    from ofri.device import context
    H = W = 1700
    a = np.zeros((H, W), np.float32)
    st, _, _, _, _ = _call("ofri_ls_noalloc", a, a, 5.0, 5, 1e-8, a, a)
    assert st == _capi.OFRI_E_ARG
    msg = context().lib.ofri_last_error().decode()
    print("[ls_noalloc 1700x1700] %s" % msg)
    assert "device scratch" in msg and "1700x1700" in msg
    # 1600 x 1600 (2.56 Mpx) still fits the array
    b = np.zeros((1600, 1600), np.float32)
    st, _, _, _, _ = _call("ofri_ls_noalloc", b, b, 5.0, 5, 1e-8, b, b)
    assert st == _capi.OFRI_OK
