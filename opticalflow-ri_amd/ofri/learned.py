"""The learned form of LS(HS) 2-level: accelerated HS and a learned
Liu-Shen correction, solves to the plain method's residual targets, plain
fallback.

    r*  per level: the plain solver's update norm at its stopping point (HS
        after its 600 sweeps, Liu-Shen at tol or 60 sweeps), from
        ofri_pyr_lshs_taps on the same pair (plain_targets), or given.
    run per level: HS from zero (Chebyshev-accelerated, plain finish) until
        its norm <= r*_HS; the correction network maps the level's Liu-Shen
        pair and HS's flow to Liu-Shen's correction; Liu-Shen from HS's flow
        plus that correction until its norm <= r*_LS.
    fallback: when a level misses a target within the sweep budgets, the
        plain pyramid runs instead (its result is returned).
"""
import time

import numpy as np

from . import ops
from .device import DeviceArray, context

WORKLOAD = dict(levels=2, alphas=[45.0, 21.0], niter=600, ls_alpha=5.0, ls_maxnum=60, ls_tol=1e-8)
# Chebyshev spectrum bounds of the accelerated HS per level (coarse: alpha
# 45, fine: alpha 21 -- a larger alpha leaves the slowest mode closer to 1)
# and the factor over the target at which plain finishing sweeps take over.
# Measured on 60 held-out 1024^2 PIV pairs: with 0.99 / 0.985 and factors
# 4 / 2 the coarse level takes 60-96 sweeps and the fine one 80-90; a
# factor 4 at the fine level leaves a fifth of the pairs needing 110-160
# (plain sweeps from 4x the target are slow there), a bound of 0.99 there
# takes 90-100.
CHEB_BOUND = (0.99, 0.985)
CHEB_FINISH = (4.0, 2.0)
# Liu-Shen runs plain sweeps.  Its sweep's spectrum lies in about
# [-0.53, 0.999] on PIV pairs (alpha 5; power iteration on the oracle's
# sweep), and Chebyshev sweeps over [-0.55, 0.99] reach the plain residual
# in about half the sweeps -- but the plain method stops Liu-Shen far from
# its fixed point (60 sweeps), so an iterate with the same residual reached
# along another path lies 0.03-0.2 px from the plain result.  (lo, hi) here
# would switch the acceleration on; (0, 0) keeps the plain sweeps.
LS_CHEB = (0.0, 0.0)


def filters():
    from gaussian_filter import prepareGaussianKernel
    return prepareGaussianKernel(3.4, 3), prepareGaussianKernel(0.48, 5)


# A residual below the float32 resolution of the solution itself carries no
# information: the target of a level is the larger of the plain residual and
# the norm of a 2^-22 relative perturbation of the plain HS solution,
# (||U||_2 + ||V||_2) * 2^-22 / N, the same norm the residual uses.
F32_FLOOR = 2.0 ** -22


def plain_targets(im1, im2, fw=None, ow=None):
    """Plain run with residuals recorded.  Returns (U, V, hs_tol[2],
    ls_tol[2], stats); stats carry the raw plain residuals (hs_res, ls_err)."""
    if fw is None:
        fw, ow = filters()
    U, V, taps, st = ops.pyr_lshs_taps(im1, im2, filt_w=fw, opt_w=ow, **WORKLOAD)
    t = taps.download()
    hs_tol, off = [], 0
    for s in st:
        n = s["H"] * s["W"]
        u, v = t[off:off + n].astype(np.float64), t[off + n:off + 2 * n].astype(np.float64)
        floor = F32_FLOOR * (np.sqrt((u * u).sum()) + np.sqrt((v * v).sum())) / n
        hs_tol.append(float(max(s["hs_res"], floor)))
        off += 4 * n
    return U, V, hs_tol, [float(s["ls_err"]) for s in st], st


# Sweep budgets of the learned form per level (coarse, fine): HS 132 / 110
# sweeps in all, of them at most 36 / 20 plain finishing sweeps and the rest
# accelerated (on 60 held-out 1024^2 PIV pairs: 48-84 / 70-80 accelerated,
# one finishing launch of 12 / 10 on nearly every pair, three at most),
# Liu-Shen 20 (5-16 used after the network's correction; the plain method
# takes 60).  A level that misses its target within them sends the pair to
# the plain fallback; launches past a finished solve still cost their
# dispatch, so the budgets stay close to use.
HS_BUDGET = (132, 110)
HS_FINISH_BUDGET = (36, 20)
LS_BUDGET = 20


def _levels(x):
    return (x, x) if np.isscalar(x) else tuple(x)


def _finish_budgets(hs_budget, hs_finish_budget):
    return tuple(min(f, b // 2 or 1) for f, b in zip(_levels(hs_finish_budget), _levels(hs_budget)))


def solve(im1, im2, net, hs_tol, ls_tol, fw=None, ow=None, Uacc=None, Vacc=None, fallback=True, stats=True,
          hs_budget=HS_BUDGET, ls_budget=LS_BUDGET, ls_cheb=LS_CHEB, hs_finish_budget=HS_FINISH_BUDGET):
    """One pair through the learned form.  Returns (U, V, info) with info:
    converged, fell_back, per-level stats."""
    if fw is None:
        fw, ow = filters()
    U, V, st, conv = ops.pyr_lshs_learned(im1, im2, net=net, hs_tol=hs_tol, ls_tol=ls_tol, filt_w=fw, opt_w=ow,
                                          hs_budget=hs_budget, ls_budget=ls_budget, cheb_bound=CHEB_BOUND,
                                          cheb_finish=CHEB_FINISH, ls_cheb=ls_cheb, Uacc=Uacc, Vacc=Vacc,
                                          hs_finish_budget=_finish_budgets(hs_budget, hs_finish_budget),
                                          stats=stats or fallback,
                                          **{k: WORKLOAD[k] for k in ("levels", "alphas", "niter", "ls_alpha")})
    fell = False
    if fallback and not conv:
        ops.pyr_lshs(im1, im2, filt_w=fw, opt_w=ow, Uacc=U, Vacc=V, stats=False, **WORKLOAD)
        fell = True
    return U, V, dict(converged=conv, fell_back=fell, stats=st)


# Held-out evaluation pairs use seeds from here on (training uses 0 .. 19999)
HELDOUT_SEED = 1000000


def heldout_pairs(n, size=1024):
    """n device-rendered pairs of size x size outside the training seeds, flow
    families cycling Poiseuille, Lamb-Oseen, shear + turbulence.  Returns a
    list of (im1, im2) DeviceArrays and the list of their specs."""
    from .train import pair_spec
    out, specs, seed = [], [], HELDOUT_SEED
    for k in range(n):
        while True:
            sp = pair_spec(seed, size, size)
            seed += 1
            if sp["family"] == k % 3:
                break
        out.append(ops.render_pair(size, size, sp["family"], sp["density"], sp["seed"], sp["par"], sp["modes"]))
        specs.append(sp)
    return out, specs


def solve_many(jobs, net, fw=None, ow=None, flags=None, hs_budget=HS_BUDGET, ls_budget=LS_BUDGET,
               hs_finish_budget=HS_FINISH_BUDGET):
    """The learned form for many pairs with one host synchronisation: every
    pair is queued with its convergence flag written on the device, the flags
    are read once, and the pairs that missed a target run the plain pyramid.
    jobs: list of (im1, im2, hs_tol, ls_tol, Uacc, Vacc) with DeviceArrays
    (Uacc, Vacc receive the result).  flags: optional DeviceArray of at least
    len(jobs) 4-byte words.  Returns the list of fell-back flags."""
    if fw is None:
        fw, ow = filters()
    if flags is None or flags.nbytes // 4 < len(jobs):
        flags = DeviceArray((len(jobs),))
    for k, (a, b, ht, lt, U, V) in enumerate(jobs):
        ops.pyr_lshs_learned(a, b, net=net, hs_tol=ht, ls_tol=lt, filt_w=fw, opt_w=ow, hs_budget=hs_budget,
                             ls_budget=ls_budget, cheb_bound=CHEB_BOUND, cheb_finish=CHEB_FINISH, Uacc=U, Vacc=V,
                             stats=False, flag=flags.view(k, (1,)),
                             hs_finish_budget=_finish_budgets(hs_budget, hs_finish_budget),
                             **{key: WORKLOAD[key] for key in ("levels", "alphas", "niter", "ls_alpha")})
    ok = flags.download().reshape(-1)[:len(jobs)].view(np.int32)
    fell = []
    for k, (a, b, _, _, U, V) in enumerate(jobs):
        if not ok[k]:
            ops.pyr_lshs(a, b, filt_w=fw, opt_w=ow, Uacc=U, Vacc=V, stats=False, **WORKLOAD)
        fell.append(not ok[k])
    return fell


def epe(a, b):
    e = np.sqrt((a[0] - b[0]) ** 2 + (a[1] - b[1]) ** 2)
    return float(e.mean()), float(e.max())


def evaluate(pairs, net, reps=5, with_baseline=True, ls_cheb=LS_CHEB):
    """Per pair: plain time, learned time (network + solves to r*), sweeps,
    EPE of the learned result to the plain one; also the method without the
    network (zero warm starts) when with_baseline.  pairs: list of (im1, im2)."""
    ctx = context()
    fw, ow = filters()
    rows = []
    for im1, im2 in pairs:
        d1, d2 = (im1 if hasattr(im1, "ptr") else DeviceArray.from_host(im1),
                  im2 if hasattr(im2, "ptr") else DeviceArray.from_host(im2))
        H, W = d1.shape[-2:]
        Up, Vp, ht, lt, pst = plain_targets(d1, d2, fw, ow)
        plain = (Up.download(), Vp.download())
        row = dict(H=H, W=W, hs_tol=ht, ls_tol=lt, plain_ls_iters=[s["ls_iters"] for s in pst])
        Ua, Va = DeviceArray((H, W)), DeviceArray((H, W))

        def timed(fn):
            for _ in range(2):
                fn()
            ctx.sync()
            t = time.perf_counter()
            for _ in range(reps):
                fn()
            ctx.sync()
            return (time.perf_counter() - t) / reps * 1e3

        row["plain_ms"] = timed(lambda: ops.pyr_lshs(d1, d2, filt_w=fw, opt_w=ow, Uacc=Ua, Vacc=Va, stats=False,
                                                     **WORKLOAD))
        for name, nt in (("learned", net), ("nonet", None)) if with_baseline else (("learned", net),):
            U, V, info = solve(d1, d2, nt, ht, lt, fw, ow, Uacc=Ua, Vacc=Va, fallback=False, ls_cheb=ls_cheb)
            res = (U.download(), V.download())
            row[name] = dict(converged=info["converged"], epe=epe(res, plain),
                             hs_sweeps=[s["hs_iters"] for s in info["stats"]],
                             ls_sweeps=[s["ls_iters"] for s in info["stats"]],
                             ms=timed(lambda: solve(d1, d2, nt, ht, lt, fw, ow, Uacc=Ua, Vacc=Va, fallback=True,
                                                    stats=False, ls_cheb=ls_cheb)))
        rows.append(row)
    return rows
