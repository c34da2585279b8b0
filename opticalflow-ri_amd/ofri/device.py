"""Device context and device-resident float32 images.

One process-wide context (one HIP stream) drives all drop-in classes; the
pyramid driver keeps every intermediate image and flow in device memory and
only crosses to the host at the API boundary, where the reference's callers
hand over and receive numpy arrays.
"""
import ctypes as C
import os
import threading

import numpy as np

from . import _capi

_ctx = None
_lock = threading.Lock()


class Context:
    def __init__(self, device=0):
        self.lib = _capi.load()
        h = C.c_void_p()
        _capi.check(self.lib.ofri_create(int(device), C.byref(h)))
        self.handle = h
        self.device = device
        from . import autotune
        autotune.apply(self)       # the autotuner's cached size classes for this GPU, if any

    def sync(self):
        _capi.check(self.lib.ofri_sync(self.handle))

    def close(self):
        if self.handle:
            self.lib.ofri_destroy(self.handle)
            self.handle = None
        self._lanes = {}

    def lane(self, k):
        """Lane k of this context (ofri_lane): its own HIP stream on the same
        device, owned by this context."""
        lanes = self.__dict__.setdefault("_lanes", {})
        if k not in lanes:
            h = C.c_void_p()
            _capi.check(self.lib.ofri_lane(self.handle, int(k), C.byref(h)))
            lanes[k] = Lane(self, h, k)
        else:
            h = C.c_void_p()   # refreshes the lane's tuning and graph setting
            _capi.check(self.lib.ofri_lane(self.handle, int(k), C.byref(h)))
        return lanes[k]

    def fork(self, n):
        """Work queued on this context so far precedes lanes 0..n-1's later work."""
        for k in range(n):
            self.lane(k)
        _capi.check(self.lib.ofri_fork(self.handle, int(n)))

    def join(self, n):
        """Lanes 0..n-1's work so far precedes this context's later work."""
        _capi.check(self.lib.ofri_join(self.handle, int(n)))

    def set_tuning(self, hs_depth=0, ls_depth=0, hs_variant=0, ls_variant=0):
        t = _capi.Tuning(int(hs_depth), int(ls_depth), int(hs_variant), int(ls_variant))
        _capi.check(self.lib.ofri_set_tuning(self.handle, C.byref(t)))

    def get_tuning(self):
        t = _capi.Tuning()
        _capi.check(self.lib.ofri_get_tuning(self.handle, C.byref(t)))
        return t.hs_depth, t.ls_depth, t.hs_variant, t.ls_variant

    def set_tuning_classes(self, classes):
        """classes: iterable of (min_px, hs_variant, hs_depth, ls_variant,
        ls_depth) -- 1-based variants, 0 = the built-in default
        (ofri_set_tuning_classes); empty clears."""
        cl = list(classes)
        arr = (_capi.TuningClass * max(1, len(cl)))()
        for i, (mp, hv, hd, lv, ld) in enumerate(cl):
            arr[i].min_px = int(mp)
            arr[i].t = _capi.Tuning(int(hd), int(ld), int(hv), int(lv))
        _capi.check(self.lib.ofri_set_tuning_classes(self.handle, C.cast(arr, C.c_void_p), len(cl)))

    def arch(self):
        buf = C.create_string_buffer(64)
        _capi.check(self.lib.ofri_arch(self.handle, buf, 64))
        return buf.value.decode()

    def query_tuning(self, H, W):
        """(hs_depth, ls_depth, hs_variant, ls_variant) the sweeps use for H x W."""
        t = _capi.Tuning()
        _capi.check(self.lib.ofri_query_tuning(self.handle, int(H), int(W), C.byref(t)))
        return t.hs_depth, t.ls_depth, t.hs_variant, t.ls_variant

    def set_graphs(self, on):
        _capi.check(self.lib.ofri_set_graphs(self.handle, 1 if on else 0))

    def timer_start(self):
        _capi.check(self.lib.ofri_timer_start(self.handle))

    def timer_stop(self):
        ms = C.c_float()
        _capi.check(self.lib.ofri_timer_stop(self.handle, C.byref(ms)))
        return ms.value


class Lane(Context):
    """A lane of a Context: usable wherever a context is taken; destroyed
    with its parent (which this object keeps alive)."""

    def __init__(self, parent, handle, k):
        self.lib = parent.lib
        self.handle = handle
        self.device = parent.device
        self.parent = parent
        self.index = k

    def close(self):
        pass

    def lane(self, k):
        raise ValueError("a lane has no lanes of its own")


def context():
    """The process-wide context on device OFRI_DEVICE (default 0)."""
    global _ctx
    with _lock:
        if _ctx is None:
            _ctx = Context(int(os.environ.get("OFRI_DEVICE", "0")))
        return _ctx


class DeviceArray:
    """A float32 H x W array in device memory owned by this object."""

    __slots__ = ("ctx", "shape", "ptr", "nbytes")

    def __init__(self, shape, ctx=None):
        self.ctx = ctx or context()
        self.shape = tuple(int(s) for s in shape)
        self.nbytes = int(np.prod(self.shape)) * 4
        p = C.c_void_p()
        _capi.check(self.ctx.lib.ofri_malloc(self.ctx.handle, self.nbytes, C.byref(p)))
        self.ptr = p

    @classmethod
    def from_host(cls, a, ctx=None):
        a = np.ascontiguousarray(a, dtype=np.float32)
        d = cls(a.shape, ctx)
        d.upload(a)
        return d

    @classmethod
    def zeros(cls, shape, ctx=None):
        d = cls(shape, ctx)
        _capi.check(d.ctx.lib.ofri_memset(d.ctx.handle, d.ptr, 0, d.nbytes))
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        if a.shape != self.shape:
            raise ValueError("shape %s != %s" % (a.shape, self.shape))
        _capi.check(self.ctx.lib.ofri_upload(self.ctx.handle, self.ptr, a.ctypes.data_as(C.c_void_p),
                                             self.nbytes))

    def download(self):
        out = np.empty(self.shape, np.float32)
        _capi.check(self.ctx.lib.ofri_download(self.ctx.handle, out.ctypes.data_as(C.c_void_p), self.ptr,
                                               self.nbytes))
        return out

    def copy(self):
        d = DeviceArray(self.shape, self.ctx)
        _capi.check(self.ctx.lib.ofri_copy(self.ctx.handle, d.ptr, self.ptr, self.nbytes))
        return d

    @property
    def H(self):
        return self.shape[-2]

    @property
    def W(self):
        return self.shape[-1]

    def view(self, offset, shape):
        """A float32 window of this allocation starting `offset` elements in
        (no copy; valid while this array lives)."""
        return DeviceView(self, offset, shape)

    def __del__(self):
        try:
            if self.ptr:
                self.ctx.lib.ofri_free(self.ctx.handle, self.ptr)
                self.ptr = None
        except Exception:
            pass


class DeviceView(DeviceArray):
    """A window of another DeviceArray's memory; frees nothing."""

    __slots__ = ("base",)

    def __init__(self, base, offset, shape):
        self.ctx = base.ctx
        self.shape = tuple(int(s) for s in shape)
        self.nbytes = int(np.prod(self.shape)) * 4
        if offset < 0 or offset * 4 + self.nbytes > base.nbytes:
            raise ValueError("view [%d, +%d) outside an allocation of %d bytes" % (offset * 4, self.nbytes, base.nbytes))
        self.base = base
        self.ptr = C.c_void_p(base.ptr.value + 4 * int(offset))

    def __del__(self):
        pass


def as_device(a, ctx=None):
    return a if isinstance(a, DeviceArray) else DeviceArray.from_host(a, ctx)
