"""Autotuner for the temporally blocked sweep kernels.

For an image size it times every compiled tile shape (variant) and fused
sweep depth of the Horn-Schunck and Liu-Shen kernels on the device (HIP
events on the context stream) and returns the fastest.  Results are cached
per GPU architecture and size in a JSON file (OFRI_TUNE_CACHE, default
~/.cache/ofri/tuning.json).  The library's built-in defaults were chosen
from these measurements on gfx950.

    python -m ofri.autotune 1024 512        # tune, print and cache

Every new context installs the cached results for its GPU architecture as
size classes (ofri_set_tuning_classes): a plain solve, or a batch by its
total pixel count, takes the shape tuned at the nearest tuned size (class
boundaries at the geometric means of neighbouring tuned sizes).
OFRI_TUNE_CACHE="" disables the cache.
"""
import functools
import json
import os
import sys

import numpy as np

from . import _capi
from .device import DeviceArray, context

HS_VARIANTS = 12    # number of compiled Horn-Schunck tile shapes
LS_VARIANTS = 7     # number of compiled Liu-Shen tile shapes
CACHE = os.environ.get("OFRI_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "ofri", "tuning.json"))


@functools.lru_cache(maxsize=4)
def _pair(n, seed=5):
    from . import synth
    rng = np.random.default_rng(seed)
    im1, im2, _, _ = synth.piv_pair(n, n, 0.05, lambda x, y: (1.2 + 0 * x, -0.4 + 0 * y), rng)
    return im1, im2


def time_hs(n, variant, depth, sweeps=120, reps=3):
    ctx = context()
    im1, im2 = _pair(n)
    d1, d2 = DeviceArray.from_host(im1), DeviceArray.from_host(im2)
    U, V = DeviceArray.zeros((n, n)), DeviceArray.zeros((n, n))
    ctx.set_tuning(hs_depth=depth, hs_variant=variant + 1)
    try:
        st = ctx.lib.ofri_hs(ctx.handle, d1.ptr, d2.ptr, n, n, 21.0, sweeps, U.ptr, V.ptr, None)
        if st != 0:
            return None
        ctx.timer_start()
        for _ in range(reps):
            _capi.check(ctx.lib.ofri_hs(ctx.handle, d1.ptr, d2.ptr, n, n, 21.0, sweeps, U.ptr, V.ptr, None))
        return ctx.timer_stop() / reps / sweeps
    finally:
        ctx.set_tuning()


def time_ls(n, variant, depth, sweeps=60, reps=3):
    ctx = context()
    im1, im2 = _pair(n)
    d1, d2 = DeviceArray.from_host(im1), DeviceArray.from_host(im2)
    u, v = DeviceArray.zeros((n, n)), DeviceArray.zeros((n, n))
    ctx.set_tuning(ls_depth=depth, ls_variant=variant + 1)
    try:
        st = ctx.lib.ofri_ls(ctx.handle, d1.ptr, d2.ptr, n, n, 5.0, sweeps, 0.0, u.ptr, v.ptr, None, None)
        if st != 0:
            return None
        ctx.timer_start()
        for _ in range(reps):
            _capi.check(ctx.lib.ofri_ls(ctx.handle, d1.ptr, d2.ptr, n, n, 5.0, sweeps, 0.0, u.ptr, v.ptr,
                                        None, None))
        return ctx.timer_stop() / reps / sweeps
    finally:
        ctx.set_tuning()


def tune(n, depths=(3, 4, 5, 6, 8, 10), verbose=True):
    """Returns {'hs': (variant, depth, ms/sweep), 'ls': (...)} for an n x n image;
    variant is the 0-based index of the compiled tile shape (tuning value - 1)."""
    out = {}
    for kind, nv, fn in (("hs", HS_VARIANTS, time_hs), ("ls", LS_VARIANTS, time_ls)):
        best = None
        for v in range(nv):
            for d in depths:
                t = fn(n, v, d)
                if t is None:
                    continue
                if verbose:
                    print("%s n=%d variant=%d depth=%d: %.4f ms/sweep" % (kind, n, v, d, t), flush=True)
                if best is None or t < best[2]:
                    best = (v, d, t)
        out[kind] = best
    return out


def save(results, arch=None, path=None):
    """Merges {n: tune(n)} into the cache under the GPU's architecture."""
    path = CACHE if path is None else path
    arch = arch or context().arch()
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    data.setdefault(arch, {}).update({str(k): v for k, v in results.items()})
    with open(path, "w") as f:
        json.dump(data, f, indent=1)


def classes(entries):
    """Size classes from cached results {n: {'hs': [variant, depth, ms], 'ls':
    [...]}} (0-based variants; an n x n image): (min_px, hs_variant,
    hs_depth, ls_variant, ls_depth) rows with 1-based variants, the class of
    size n starting at the geometric mean of n^2 and the next smaller tuned
    size's pixel count (at 0 for the smallest)."""
    sizes = sorted(int(k) for k in entries)
    rows = []
    for i, n in enumerate(sizes):
        r = entries[str(n)] if str(n) in entries else entries[n]
        lo = 0 if i == 0 else int(round(float(sizes[i - 1]) * n))
        hs, ls = r.get("hs"), r.get("ls")
        rows.append((lo, hs[0] + 1 if hs else 0, hs[1] if hs else 0, ls[0] + 1 if ls else 0, ls[1] if ls else 0))
    return rows


def apply(ctx, path=None):
    """Installs the cache's classes for ctx's architecture (none: nothing).
    Returns the rows installed."""
    path = CACHE if path is None else path
    if not path or not os.path.exists(path):
        return []
    try:
        with open(path) as f:
            data = json.load(f)
        entries = data.get(ctx.arch(), {})
        rows = classes(entries) if entries else []
    except (OSError, ValueError, KeyError, TypeError, IndexError):
        return []
    if rows:
        ctx.set_tuning_classes(rows)
    return rows


if __name__ == "__main__":
    sizes = [int(a) for a in sys.argv[1:]] or [1024, 512]
    res = {n: tune(n) for n in sizes}
    for n, r in res.items():
        print("best n=%d: %s" % (n, json.dumps(r)))
    save(res)
