"""Correction network handle (libofri ofri_net_*): the learned form's
encoder-decoder, its parameters, prediction and training steps.

In the learned LS(HS) form the network (4 input planes) maps a level's
Liu-Shen input pair and the HS flow to the correction Liu-Shen's sweeps add
on top of HS (fine head, fields 0, 1).  All compute runs in libofri's
kernels; this module only moves parameters and buffers.
"""
import ctypes as C
import os

import numpy as np

from . import _capi
from .device import DeviceArray, as_device, context

# Loss weights per output field (coarse head fields 0..3, fine head fields
# 0..3): the correction network learns the fine head's first two fields.
FIELD_WEIGHTS = np.array([0.0, 0.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0], np.float32)
WEIGHTS_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "correction_weights.npz")


class Net(object):
    def __init__(self, params=None, ctx=None, planes=4):
        self.ctx = ctx or context()
        self.planes = int(planes)
        h = C.c_void_p()
        _capi.check(self.ctx.lib.ofri_net_create(self.ctx.handle, self.planes, C.byref(h)))
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
            planes = int(z["planes"]) if "planes" in z.files else 4
            return cls(z["params"].astype(np.float32), ctx, planes)

    def save(self, path, **meta):
        np.savez_compressed(path, params=self.get_params(), planes=self.planes, **meta)

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

    def flops(self, H, W):
        """Floating-point operations of one forward pass over an H x W input
        (2 per multiply-add on the real channels)."""
        f = C.c_double()
        _capi.check(self.ctx.lib.ofri_net_flops(self.handle, int(H), int(W), C.byref(f)))
        return f.value

    def get_grads(self):
        out = np.empty(self.num_params, np.float32)
        _capi.check(self.ctx.lib.ofri_net_get_grads(self.ctx.handle, self.handle, _capi.fptr(out), out.size))
        return out

    def predict(self, inp, coarse=None, fine=None):
        """inp: DeviceArray or numpy, (B, planes, H, W) or (planes, H, W).
        Returns (coarse [B, 4, H/2, W/2], fine [B, 4, H, W]) DeviceArrays."""
        d = as_device(inp)
        B, H, W = self._dims(d)
        coarse = coarse if coarse is not None else DeviceArray((B, 4, H // 2, W // 2))
        fine = fine if fine is not None else DeviceArray((B, 4, H, W))
        _capi.check(self.ctx.lib.ofri_net_predict(self.ctx.handle, self.handle, d.ptr, B, H, W, coarse.ptr,
                                                  fine.ptr))
        return coarse, fine

    def correct(self, planes, u, v):
        """Adds the fine head's fields 0, 1 for one input (a list of `planes`
        H x W DeviceArrays) into the H x W DeviceArrays u, v."""
        planes = [as_device(p) for p in planes]
        if len(planes) != self.planes:
            raise ValueError("network takes %d input planes, got %d" % (self.planes, len(planes)))
        H, W = planes[0].shape[-2:]
        arr = (C.c_void_p * self.planes)(*[p.ptr.value for p in planes])
        _capi.check(self.ctx.lib.ofri_net_correct(self.ctx.handle, self.handle, arr, H, W, u.ptr, v.ptr))
        return u, v

    def train_step(self, inp, lab_c, lab_f, lr, field_w=FIELD_WEIGHTS):
        """One step on a batch (lr = 0: gradients only).  Returns the loss."""
        d = as_device(inp)
        lc, lf = as_device(lab_c), as_device(lab_f)
        B, H, W = self._dims(d)
        fw = np.ascontiguousarray(field_w, dtype=np.float32)
        loss = C.c_float()
        _capi.check(self.ctx.lib.ofri_net_train(self.ctx.handle, self.handle, d.ptr, B, H, W, lc.ptr, lf.ptr,
                                                _capi.fptr(fw), float(lr), C.byref(loss)))
        return float(loss.value)

    def _dims(self, d):
        shp = d.shape if len(d.shape) == 4 else (1,) + tuple(d.shape)
        if len(shp) != 4 or shp[1] != self.planes:
            raise ValueError("input of shape %s: expected (B, %d, H, W)" % (d.shape, self.planes))
        return shp[0], shp[2], shp[3]

    def __del__(self):
        try:
            if self.handle:
                self.ctx.lib.ofri_net_destroy(self.handle)
                self.handle = None
        except Exception:
            pass
