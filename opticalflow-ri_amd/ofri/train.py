"""Training of the correction network on the GPU.

Every training pair is rendered on the device (ofri_render_pair) from its
own seed: H, W from {512, 576, ..., 1024}, density U[0.02, 0.08], flow
family uniform over Poiseuille, Lamb-Oseen and shear + turbulence (max 4 px).
The plain LS(HS) 2-level solver runs on it with the workload's parameters
(FILTER 3.4 / 3 taps, HS alpha 45 then 21 with 600 sweeps, FILTER_OPT 0.48 /
5 taps, Liu-Shen alpha 5, tol 1e-8, 60 sweeps) and exports every level's HS
flow and Liu-Shen correction (ofri_pyr_lshs_taps).  Each level gives one
sample: input = the level's Liu-Shen pair (FILTER_OPT-filtered, resized to
the level) and HS's flow; label = the Liu-Shen correction.  Forward, loss,
backward and Adam all run in libofri's kernels.

    python -m ofri.train --pairs 20000 --batch 8 --out ofri/correction_weights.npz

Labels are the Liu-Shen correction a few sweeps before the plain stop (the
learned form's own Liu-Shen sweeps, --tap-back), so that those sweeps land
on the plain iterate.

The pair set is a pure function of --seed; an epoch re-renders and
re-solves it (the labels of 20k pairs do not fit next to the images in HBM).
"""
import argparse
import math
import os
import sys
import time

import numpy as np

from . import ops
from .device import DeviceArray, context
from .net import FIELD_WEIGHTS, Net

SIZES = tuple(range(512, 1025, 64))
LSHS = dict(levels=2, alphas=[45.0, 21.0], niter=600, ls_alpha=5.0, ls_maxnum=60, ls_tol=1e-8)


def prefilter_weights():
    from gaussian_filter import prepareGaussianKernel
    return prepareGaussianKernel(3.4, 3), prepareGaussianKernel(0.48, 5)


def pair_spec(seed, H=None, W=None):
    """Size, flow family and flow parameters of training pair `seed`."""
    rng = np.random.default_rng(np.random.SeedSequence([20261203, int(seed)]))
    if H is None:
        H, W = int(rng.choice(SIZES)), int(rng.choice(SIZES))
    density = float(rng.uniform(0.02, 0.08))
    family = int(rng.integers(0, 3))
    modes = None
    if family == 0:
        par = [4.0]
    elif family == 1:
        m = min(H, W)
        par = [float(rng.uniform(0.3, 0.7) * W), float(rng.uniform(0.3, 0.7) * H),
               float(rng.uniform(0.05, 0.1) * m), 4.0]
    else:
        n = 256
        kmag = np.exp(rng.uniform(np.log(4 * np.pi / max(H, W)), np.log(2 * np.pi / 8.0), n))
        ang = rng.uniform(0, 2 * np.pi, n)
        phi = rng.uniform(0, 2 * np.pi, n)
        amp = kmag ** (-5.0 / 6.0) * np.sqrt(kmag) / kmag
        modes = np.stack([kmag * np.cos(ang), kmag * np.sin(ang), phi, amp]).astype(np.float32)
        par = [2.0, 4.0]
    return dict(H=H, W=W, density=density, family=family, par=par, modes=modes, seed=int(seed))


class Batch(object):
    """Device buffers for minibatches of up to Bmax pairs of up to Hmax x Wmax;
    set(B, H, W) lays the current batch out at the front of each buffer."""

    def __init__(self, Bmax, Hmax, Wmax):
        n = Bmax * Hmax * Wmax
        self.buf = {"im1": DeviceArray((n,)), "im2": DeviceArray((n,)),
                    "in_f": DeviceArray((4 * n,)), "in_c": DeviceArray((n,)),      # 4 planes, full / half size
                    # fine-head labels: fields 0, 1 written per sample, 2, 3 stay finite (weight 0)
                    "lab_f": DeviceArray.zeros((4 * n,)), "lab_c": DeviceArray.zeros((n,)),
                    "zero": DeviceArray.zeros((n,)),                              # coarse-head labels (unused)
                    "taps": DeviceArray((5 * Hmax * Wmax,)),
                    "U": DeviceArray((Hmax * Wmax,)), "V": DeviceArray((Hmax * Wmax,)),
                    "t1": DeviceArray((Hmax * Wmax,)), "t2": DeviceArray((Hmax * Wmax,))}
        self.cap = (Bmax, Hmax, Wmax)

    def set(self, B, H, W):
        Bm, Hm, Wm = self.cap
        if B > Bm or H * W > Hm * Wm:
            raise ValueError("batch %dx%dx%d exceeds %dx%dx%d" % (B, H, W, Bm, Hm, Wm))
        self.B, self.H, self.W = B, H, W
        b = self.buf
        h, w = H // 2, W // 2
        self.im1 = b["im1"].view(0, (B, H, W))
        self.im2 = b["im2"].view(0, (B, H, W))
        self.in_f = b["in_f"].view(0, (B, 4, H, W))
        self.in_c = b["in_c"].view(0, (B, 4, h, w))
        self.lab_f = b["lab_f"].view(0, (B, 4, H, W))
        self.lab_c = b["lab_c"].view(0, (B, 4, h, w))
        self.zero_f = b["zero"].view(0, (B, 4, h, w))      # coarse-head labels of the fine sample
        self.zero_c = b["zero"].view(0, (B, 4, h // 2, w // 2))
        self.taps = b["taps"].view(0, (4 * h * w + 4 * H * W,))
        self.U = b["U"].view(0, (H, W))
        self.V = b["V"].view(0, (H, W))
        return self


def _copy(dst, src, n):
    ctx = context()
    _ = ctx.lib.ofri_copy(ctx.handle, dst.ptr, src.ptr, 4 * n)


def fill(batch, specs, fw, ow, tap_back=None):
    """Render the batch's pairs, solve them with the plain pyramid and lay out
    the two samples per pair: fine level (in_f, lab_f) and coarse level
    (in_c, lab_c).  tap_back: per level (coarse, fine), the labels are the
    Liu-Shen correction that many sweeps before the plain stop (the learned
    form runs those last sweeps itself).  Returns the per-pair level
    statistics."""
    B, H, W = batch.B, batch.H, batch.W
    h, w = H // 2, W // 2
    hw, HW = h * w, H * W
    t1, t2 = batch.buf["t1"], batch.buf["t2"]
    stats = []
    for i, sp in enumerate(specs):
        a = batch.im1.view(i * HW, (H, W))
        b = batch.im2.view(i * HW, (H, W))
        ops.render_pair(H, W, sp["family"], sp["density"], sp["seed"], sp["par"], sp["modes"], 2.0, im1=a, im2=b)
        _, _, taps, st = ops.pyr_lshs_taps(a, b, filt_w=fw, opt_w=ow, Uacc=batch.U, Vacc=batch.V, taps=batch.taps,
                                           ls_tap_back=tap_back, **LSHS)
        stats.append(st)
        # fine level: Liu-Shen pair = FILTER_OPT of the frames
        base = i * 4 * HW
        ops.gauss_sep(a, ow, out=batch.in_f.view(base, (H, W)))
        ops.gauss_sep(b, ow, out=batch.in_f.view(base + HW, (H, W)))
        _copy(batch.in_f.view(base + 2 * HW, (2 * HW,)), taps.view(4 * hw, (2 * HW,)), 2 * HW)
        _copy(batch.lab_f.view(base, (2 * HW,)), taps.view(4 * hw + 2 * HW, (2 * HW,)), 2 * HW)
        # coarse level: FILTER_OPT of the bicubic half-size frames
        basec = i * 4 * hw
        r1 = ops.resize(a, h, w)
        r2 = ops.resize(b, h, w)
        ops.gauss_sep(r1, ow, out=batch.in_c.view(basec, (h, w)))
        ops.gauss_sep(r2, ow, out=batch.in_c.view(basec + hw, (h, w)))
        _copy(batch.in_c.view(basec + 2 * hw, (2 * hw,)), taps.view(0, (2 * hw,)), 2 * hw)
        _copy(batch.lab_c.view(basec, (2 * hw,)), taps.view(2 * hw, (2 * hw,)), 2 * hw)
    return stats


def lr_at(step, total, base, warm=100, floor=0.02):
    if step < warm:
        return base * (step + 1) / warm
    t = (step - warm) / max(1, total - warm)
    return base * (floor + (1 - floor) * 0.5 * (1 + math.cos(math.pi * min(1.0, t))))


def train(pairs=20000, batch=8, epochs=1.0, lr=1e-3, seed=0, out=None, resume=None, time_limit=None,
          log_every=25, tap_back=None):
    """Trains on `pairs` seeded pairs (seed .. seed + pairs - 1) for `epochs`
    passes in minibatches of `batch` same-size pairs; labels tapped
    `tap_back` Liu-Shen sweeps before the plain stop (default: the learned
    form's own Liu-Shen sweeps).  Returns the Net."""
    if tap_back is None:
        from .learned import LS_SWEEPS
        tap_back = LS_SWEEPS
    ctx = context()
    fw, ow = prefilter_weights()
    net = Net(planes=4)
    if resume:
        with np.load(resume, allow_pickle=False) as z:
            net.set_params(z["params"])
    else:
        net.init(seed)
    # minibatches: pairs grouped by size, order shuffled per epoch
    specs = [pair_spec(seed + i) for i in range(pairs)]
    groups = {}
    for sp in specs:
        groups.setdefault((sp["H"], sp["W"]), []).append(sp)
    mbs = []
    for key, lst in groups.items():
        for j in range(0, len(lst), batch):
            mbs.append(lst[j:j + batch])
    steps = int(round(len(mbs) * epochs))
    order_rng = np.random.default_rng(seed + 1)
    schedule = []
    while len(schedule) < steps:
        schedule += list(order_rng.permutation(len(mbs)))
    schedule = schedule[:steps]
    bt = Batch(batch, max(SIZES), max(SIZES))
    t0 = time.time()
    hist = []
    for step, mi in enumerate(schedule):
        mb = mbs[mi]
        H, W = mb[0]["H"], mb[0]["W"]
        bt.set(len(mb), H, W)
        fill(bt, mb, fw, ow, tap_back)
        rate = lr_at(step, steps, lr)
        # the fine level's sample, then the coarse level's (fine-head labels
        # at each level's size; the coarse head is not trained)
        loss = net.train_step(bt.in_f, bt.zero_f, bt.lab_f, rate)
        loss += net.train_step(bt.in_c, bt.zero_c, bt.lab_c, rate)
        hist.append(loss)
        if step % log_every == 0 or step == steps - 1:
            el = time.time() - t0
            print("step %5d/%d  %4dx%-4d B=%d  loss %.5f (avg %.5f)  lr %.2e  %.0fs" %
                  (step, steps, H, W, len(mb), loss, float(np.mean(hist[-log_every:])), rate, el), flush=True)
        if time_limit and time.time() - t0 > time_limit:
            print("time limit reached at step %d" % step, flush=True)
            break
    if out:
        net.save(out, pairs=pairs, batch=batch, steps=len(hist), seed=seed, final_loss=float(np.mean(hist[-50:])),
                 loss_history=np.array(hist, np.float32), tap_back=np.array(tap_back, np.int32))
    return net


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--pairs", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--epochs", type=float, default=1.0)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--resume", default=None)
    ap.add_argument("--time-limit", type=float, default=None, help="stop after this many seconds")
    ap.add_argument("--out", default=None)
    ap.add_argument("--tap-back", type=int, nargs=2, default=None,
                    help="Liu-Shen sweeps before the plain stop at which labels are taken (coarse, fine)")
    a = ap.parse_args(argv)
    train(a.pairs, a.batch, a.epochs, a.lr, a.seed, a.out, a.resume, a.time_limit, tap_back=a.tap_back)


if __name__ == "__main__":
    main()
