"""The generic coarse-to-fine driver restated over the CPU oracle's
primitives.  TEST INFRASTRUCTURE ONLY: tests/ use it as the checker of the
drop-in driver with any adapter combination; the shipped library never
calls it.

generic_pyramid follows genericPyramidalOpticalFlow
(GenericPyramidalOpticalFlow.py:237-415) and updateNextPyramidalLevel
(:117-234) with bilinear or no warping:

  - level sizes: scale = 2^-(L-1), x2 per level; a level below the last is
    the Pillow-BICUBIC resize of the input to round(size * scale) (:335-347);
  - level > 1: the accumulated flow is spline-upsampled to the level size
    (RectBivariateSpline on arange(n)/n grids, clamped) and, with the level's
    scaling flag (intermediate levels: intermediateScaling, the last: scaling,
    :329-331), multiplied by the size ratio in float32 (:151-171); warping
    (bilinear) warps the pair by it symmetrically and the level starts from
    zero with the upsampled flow as the accumulator; without warping the
    upsampled flow is the level's initial flow and the accumulator restarts
    at zero (:223-231);
  - FILTER > 1e-3: the 3-tap pre-filter of sigma FILTER on the (warped) pair,
    else a copy (:367-377); with an optional adapter, FILTER_OPT > 1e-3: the
    5-tap pre-filter on the un-warped level pair (:379-385);
  - main(im1, im2, U, V) -> (U, V), then optional(opt1, opt2, U, V) -> (U, V),
    then the accumulator += U, V in float32 (:405-413);
  - biLinear False (the Liu-Shen warping mode, :188-220): the level's first
    frame is warped in place by the upsampled flow (`ls_warp`) and the second
    is left alone, so the 5-tap pre-filter and the re-warps below see the
    warped first frame (:353, :381, :393);
  - kLevels > 1 (:387-413): iteration k > 0 with warping warps a copy of the
    level's (possibly already warped) first frame and the second frame by the
    whole accumulated flow, restarts the solver from zero and pre-filters
    again when FILTER > 1; without warping the accumulated flow becomes the
    solver's initial flow and the accumulator restarts at zero, on the same
    pre-filtered pair.  The optional adapter's pair is formed once per level.

Adapters are callables over numpy arrays (prec selects the oracle instance
they use); the helpers below build the reference's HS, Liu-Shen, Farneback
and dense-LK adapters from the oracle.
"""
import numpy as np

from . import oracle as orc


def _weights(sigma, n, prec):
    """prepareGaussianKernel's weights (gaussian_filter.py:47-52): the Gaussian
    at integer offsets np.arange(-n/2, n/2, 1, dtype=int) evaluated in float64
    and stored into the instance's type, then divided by its own sum in that
    type."""
    dt = np.float32 if prec == "f32" else np.float64
    x = np.arange(-n / 2, n / 2, 1, dtype=int)
    k = np.zeros(n, dtype=dt)
    k[:] = 1.0 / np.sqrt(2.0 * np.pi * sigma ** 2) * np.exp(-x ** 2 / (2.0 * sigma ** 2))
    k /= np.sum(k)
    return k


def generic_pyramid(im1, im2, FILTER, main, levels=1, FILTER_OPT=None, optional=None, warping=True,
                    interScaling=True, finalScaling=False, prec="f32", filt_weights=None, biLinear=True, kLevels=1):
    """Returns (Uaccum, Vaccum).  filt_weights(sigma, n) -> weights
    (default: prepareGaussianKernel's formula in the instance's precision).
    The Liu-Shen warping mode (biLinear False) is restated in float32 only."""
    dt = np.float32 if prec == "f32" else np.float64
    wfun = filt_weights or (lambda s, n: _weights(s, n, prec))
    im1 = np.asarray(im1, dt)
    im2 = np.asarray(im2, dt)
    H, W = im1.shape
    scale = 1.0 / (2.0 ** (levels - 1))
    Uacc = Vacc = U = V = None
    prev = None
    for level in range(1, levels + 1):
        local = finalScaling if level == levels else interScaling
        if scale < 1.0 and level != levels:
            h, w = int(np.int32(np.round(H * scale))), int(np.int32(np.round(W * scale)))
            n1, n2 = orc.resize(im1, h, w, 3, prec), orc.resize(im2, h, w, 3, prec)
        else:
            h, w = H, W
            n1, n2 = im1, im2
        if level > 1:
            ph, pw = prev
            sx = dt(np.float32(w) / np.float32(pw)) if local else dt(1.0)
            sy = dt(np.float32(h) / np.float32(ph)) if local else dt(1.0)
            if (ph, pw) != (h, w):
                us = orc.spline_upsample(Uacc, h, w, sx, prec)
                vs = orc.spline_upsample(Vacc, h, w, sy, prec)
            else:
                us = (Uacc * sx).astype(dt)
                vs = (Vacc * sy).astype(dt)
            if warping:
                if biLinear:
                    w1, w2 = orc.warp_symmetric(n1, n2, us, vs, prec)
                else:
                    n1 = orc.ls_warp(n1, us, vs)      # the level's frame itself is moved
                    w1, w2 = n1, n2
                U, V = np.zeros((h, w), dt), np.zeros((h, w), dt)
                Uacc, Vacc = us, vs
            else:
                w1, w2 = n1, n2
                U, V = us, vs
                Uacc, Vacc = np.zeros((h, w), dt), np.zeros((h, w), dt)
        else:
            w1, w2 = n1, n2
            U, V = np.zeros((h, w), dt), np.zeros((h, w), dt)
            Uacc, Vacc = np.zeros((h, w), dt), np.zeros((h, w), dt)
        if FILTER > 1e-3:
            k3 = wfun(FILTER, 3)
            f1, f2 = orc.gauss_sep(w1, k3, prec), orc.gauss_sep(w2, k3, prec)
        else:
            f1, f2 = w1.copy(), w2
        if optional is not None:
            if FILTER_OPT is not None and FILTER_OPT > 1e-3:
                k5 = wfun(FILTER_OPT, 5)
                o1, o2 = orc.gauss_sep(n1, k5, prec), orc.gauss_sep(n2, k5, prec)
            else:
                o1, o2 = n1.copy(), n2
        for k in range(kLevels):
            if k > 0:
                if warping:
                    if biLinear:
                        w1, w2 = orc.warp_symmetric(n1, n2, Uacc, Vacc, prec)
                    else:
                        w1, w2 = orc.ls_warp(n1, Uacc, Vacc), n2
                    U, V = np.zeros((h, w), dt), np.zeros((h, w), dt)
                    if FILTER > 1:
                        k3 = wfun(FILTER, 3)
                        f1, f2 = orc.gauss_sep(w1, k3, prec), orc.gauss_sep(w2, k3, prec)
                    else:
                        f1, f2 = w1.copy(), w2
                else:
                    U, V = Uacc, Vacc
                    Uacc, Vacc = np.zeros((h, w), dt), np.zeros((h, w), dt)
            U, V = main(f1, f2, U, V, level)
            if optional is not None:
                U, V = optional(o1, o2, U, V, level)
            Uacc = (Uacc + np.asarray(U, dt)).astype(dt)
            Vacc = (Vacc + np.asarray(V, dt)).astype(dt)
        prev = (h, w)
        scale *= 2
    return Uacc, Vacc


def hs_adapter(alphas, niter, prec="f32"):
    """HSOpticalFlowAlgoAdapter (HornSchunck.py:28-49): alphas popped from the
    end, one per call."""
    al = list(alphas)

    def main(f1, f2, U, V, level):
        a = al.pop()
        u, v, _ = orc.hs(f1, f2, a, niter, U, V, prec)
        return u, v
    return main


def ls_adapter(alpha, prec="f32", maxnum=60, tol=1e-8):
    """LiuShenOpticalFlowAlgoAdapter (PhysicsBasedOpticalFlowLiuShen.py:33-45):
    the solver's u is the caller's V (the adapter swaps them)."""
    def opt(o1, o2, U, V, level):
        v, u, _, _ = orc.ls(o1, o2, alpha, V, U, maxnum, tol, prec)
        return u, v
    return opt


def farneback_adapter(consts, prec="f32"):
    """Farneback_PyCL with one internal level (Farneback_PyCL.py:461-603,
    pyramidalLevels=1): consts as the drop-in's _consts()."""
    def main(f1, f2, U, V, level):
        return orc.farneback(f1, f2, U, V, consts, prec)
    return main


def lk_adapter(half_win=13, niter=5, prec="f32"):
    """denseLucasKanade_PyCl without vorticity enhancement
    (denseLucasKanade_PyCL.py:113-169)."""
    def main(f1, f2, U, V, level):
        u, v, _ = orc.lk_dense(f1, f2, U, V, half_win, niter, prec)
        return u, v
    return main
