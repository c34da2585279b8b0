"""Warm-start network handle (libofri ofri_net_*): the learned form's
encoder-decoder, its parameters, prediction and training steps.

The network maps a batch of pre-filtered frame pairs to the plain LS(HS)
2-level solver's per-level results (coarse HS flow and Liu-Shen correction
at H/2 x W/2, fine HS increment and Liu-Shen correction at H x W).  All
compute runs in libofri's kernels; this module only moves parameters.
"""
import ctypes as C
import os

import numpy as np

from . import _capi
from .device import DeviceArray, as_device, context

# Loss weights per output field: coarse HS u, v, LS-HS u, v; fine likewise.
FIELD_WEIGHTS = np.array([1.0, 1.0, 4.0, 4.0, 1.0, 1.0, 4.0, 4.0], np.float32)
WEIGHTS_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "warmstart_weights.npz")


class Net(object):
    def __init__(self, params=None, ctx=None):
        self.ctx = ctx or context()
        h = C.c_void_p()
        _capi.check(self.ctx.lib.ofri_net_create(self.ctx.handle, C.byref(h)))
        self.handle = h
        n = C.c_size_t()
        _capi.check(self.ctx.lib.ofri_net_num_params(h, C.byref(n)))
        self.num_params = n.value
        if params is not None:
            self.set_params(params)

    @classmethod
    def load(cls, path=WEIGHTS_FILE, ctx=None):
        """The committed trained parameters (numpy .npz, no pickles)."""
        with np.load(path, allow_pickle=False) as z:
            return cls(z["params"].astype(np.float32), ctx)

    def save(self, path, **meta):
        np.savez_compressed(path, params=self.get_params(), **meta)

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.num_params:
            raise ValueError("network has %d parameters, got %d" % (self.num_params, p.size))
        _capi.check(self.ctx.lib.ofri_net_set_params(self.ctx.handle, self.handle, _capi.fptr(p), p.size))

    def init(self, seed=0):
        """Random initial parameters (He-normal, heads scaled down)."""
        _capi.check(self.ctx.lib.ofri_net_init(self.ctx.handle, self.handle, int(seed)))

    def get_params(self):
        out = np.empty(self.num_params, np.float32)
        _capi.check(self.ctx.lib.ofri_net_get_params(self.ctx.handle, self.handle, _capi.fptr(out), out.size))
        return out

    def get_grads(self):
        out = np.empty(self.num_params, np.float32)
        _capi.check(self.ctx.lib.ofri_net_get_grads(self.ctx.handle, self.handle, _capi.fptr(out), out.size))
        return out

    def predict(self, im1, im2, coarse=None, fine=None):
        """im1, im2: DeviceArray or numpy of shape (B, H, W) or (H, W).
        Returns (coarse [B, 4, H/2, W/2], fine [B, 4, H, W]) DeviceArrays."""
        d1, d2 = as_device(im1), as_device(im2)
        shp = d1.shape if len(d1.shape) == 3 else (1,) + d1.shape
        B, H, W = shp
        coarse = coarse if coarse is not None else DeviceArray((B, 4, H // 2, W // 2))
        fine = fine if fine is not None else DeviceArray((B, 4, H, W))
        _capi.check(self.ctx.lib.ofri_net_predict(self.ctx.handle, self.handle, d1.ptr, d2.ptr, B, H, W,
                                                  coarse.ptr, fine.ptr))
        return coarse, fine

    def train_step(self, im1, im2, lab_c, lab_f, lr, field_w=FIELD_WEIGHTS):
        """One step on a batch (lr = 0: gradients only).  Returns the loss."""
        d1, d2 = as_device(im1), as_device(im2)
        lc, lf = as_device(lab_c), as_device(lab_f)
        shp = d1.shape if len(d1.shape) == 3 else (1,) + d1.shape
        B, H, W = shp
        fw = np.ascontiguousarray(field_w, dtype=np.float32)
        loss = C.c_float()
        _capi.check(self.ctx.lib.ofri_net_train(self.ctx.handle, self.handle, d1.ptr, d2.ptr, B, H, W, lc.ptr,
                                                lf.ptr, _capi.fptr(fw), float(lr), C.byref(loss)))
        return float(loss.value)

    def __del__(self):
        try:
            if self.handle:
                self.ctx.lib.ofri_net_destroy(self.handle)
                self.handle = None
        except Exception:
            pass
