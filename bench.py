"""Benchmark: LS(HS) 2-level dense flow at 1024x1024 on one MI355X -- the
learned form against the plain form.

The workload (default): LS(HS) 2-level on 1024^2 synthetic PIV pairs (the
workload's Lamb-Oseen pair, density 0.04, plus held-out rendered pairs of
the three flow families): bicubic 1/2 downsample, FILTER 3.4 (3 taps) and HS
(alpha 45 at 512^2, 21 at 1024^2), FILTER_OPT 0.48 (5 taps) and Liu-Shen
alpha 5, spline upsample x2, symmetric bilinear warp, accumulation.  The
plain form runs the reference's stopping rules (HS 600 sweeps, Liu-Shen to
1e-8 or 60 sweeps).  The learned form runs pair-independent rules (HS to a
relative residual of 2^-22, the correction network, a fixed number of
Liu-Shen sweeps with a relative-residual acceptance test), falls back to
the plain form when a rule fails, and lets the plain form judge every 32nd
pair (EPE to plain <= 0.01 px, else the plain result).  One step = one
batch of 32 pairs, judge and fallbacks included; inputs are resident in HBM
before timing.  Both forms spread a batch's pairs over the same lanes
(concurrent HIP streams of the library context, learned.LANES).

Prints ONE JSON line.  `value` is the learned form's Mpix-pairs/s
(finest-level H*W/1e6 per pair per second); `learned_vs_plain` holds the
plain form's rate, the method without the network, the EPE to the plain
result per pair and per flow family, the judge's record and the fallback
rate.  `roofline` is measured live (HIP events on the library stream) for
the learned form's dominant kernel, the Chebyshev HS sweep;
`roofline_plain_hs` for the plain form's HS sweep; `roofline_net` for the
network's forward pass against the bf16 MFMA peak.  `cpu_baseline` times the
CPU oracle (the reference dataflow restated in C) on this host at one
thread, at all threads and in float64.  Other workloads: --workload plain |
batch | fbls | lk | net.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "opticalflow-ri_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)
os.environ.setdefault("OFRI_VERBOSE", "0")

import numpy as np  # noqa: E402

METRIC = "Mpix-pairs/s to fixed residual, LiuShen-HS 1024² learned vs plain; HBM GB/s"
HBM_PEAK_GBS = 8000.0
HS_BYTES_PER_PX_SWEEP = 28.0     # read U, V, fx, fy, ft; write U, V
CHEB_BYTES_PER_PX_SWEEP = 36.0   # read U, V, previous U, V, fx, fy, ft; write U, V
HS_SETUP_BYTES_PER_PX = 20.0     # read 2 images, write fx, fy, ft
CALIB_BYTES = 64 << 20           # size of the counter-calibration copy


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--workload", default="lshs", choices=["lshs", "plain", "batch", "fbls", "lk", "net", "hs"],
                    help="lshs: 1024^2 LS(HS) 2-level learned vs plain (default); plain: the plain form alone; "
                         "batch: 64 uneven pairs LS(HS); "
                         "fbls: 1024^2 Farneback 2-level + Liu-Shen; lk: 2048^2 dense LK 2-level; "
                         "net: the correction network's forward pass over a 4-plane 1024^2 input; "
                         "hs: the HS sweep kernels alone (plain and Chebyshev, library stream)")
    ap.add_argument("--pairs", type=int, default=64, help="pairs in the batch workload")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every kernel eagerly instead of replaying HIP graphs (for tools that trace "
                         "kernels; the kernels are the same)")
    ap.add_argument("--lanes", type=int, default=0,
                    help="concurrent streams the default workload's pairs are spread over (0: the library's "
                         "default; 1: the library stream alone, as counter collection needs)")
    ap.add_argument("--eval-pairs", type=int, default=16,
                    help="pairs the default workload cycles over (its own pair + held-out rendered pairs)")
    return ap.parse_args()


def traffic_from_profile(name="hs_pass_1024_traffic.json"):
    """HBM bytes per launch from a committed rocprofv3 counter summary, if any."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            return float(json.load(f)["bytes_per_launch"])
    except Exception:
        return None


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(im1, im2, fw, ow, min_s=6.0):
    """The C oracle's LS(HS) 2-level (the reference dataflow: float32 fields,
    float64 stencil sums, HS 600 sweeps, Liu-Shen tol 1e-8 / 60 sweeps) on
    the same pair: at one thread (the reference is single-threaded: numba
    parallel=False, HornSchunck.py:51), at all of the host's threads, and the
    float64 instance at all threads.  Each leg repeats the pair until min_s
    seconds have run (one run at least)."""
    from oracle import oracle as orc
    H, W = im1.shape
    all_threads = orc.get_threads()

    def leg(threads, prec):
        orc.set_threads(threads)
        reps, t0 = 0, time.perf_counter()
        while True:
            orc.pyr_lshs(im1, im2, 2, [45.0, 21.0], 600, fw, ow, 5.0, prec=prec)
            reps += 1
            dt = time.perf_counter() - t0
            if dt >= min_s:
                return reps, dt

    try:
        r1, t1 = leg(1, "f32")
        ra, ta = leg(all_threads, "f32")
        r64, t64 = leg(all_threads, "f64")
    finally:
        orc.set_threads(all_threads)
    mpx = H * W / 1e6
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    return {"value": round(r1 * mpx / t1, 5), "unit": "Mpix-pairs/s", "cores": 1, "kind": "port",
            "sample": "%d run(s) of 1 pair %dx%d LS(HS) 2-level (plain form), C oracle (float32 fields, float64 "
                      "stencil sums), 1 thread, %.1f s" % (r1, H, W, t1),
            "all_cores": {"value": round(ra * mpx / ta, 5), "cores": all_threads,
                          "sample": "%d runs, OpenMP %d threads, %.1f s" % (ra, all_threads, ta)},
            "f64": {"value": round(r64 * mpx / t64, 5), "cores": all_threads,
                    "sample": "%d runs of the float64 instance, OpenMP %d threads, %.1f s" % (r64, all_threads, t64)},
            "cpu_model": _cpu_model(), "nproc": os.cpu_count(), "affinity": affinity,
            "opencl_cpu": "unavailable (no CPU OpenCL device in the image)"}


def _timed(step, ctx, steps, warmup):
    for _ in range(warmup):
        step()
    ctx.sync()
    ctx.timer_start()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    gpu_ms = ctx.timer_stop()
    return time.perf_counter() - t0, gpu_ms


def _calibration_copy(ctx):
    """Calibration dispatch for the HBM counters: one plain device-to-device
    copy of CALIB_BYTES (read CALIB_BYTES, write CALIB_BYTES)."""
    from ofri.device import DeviceArray
    DeviceArray((CALIB_BYTES // 4 // 1024, 1024)).copy()
    ctx.sync()


def _line(metric, value, unit, args, wall, gpu_ms, config, roofline, cpu=None, dtype="f32"):
    out = {"metric": metric, "value": round(value, 3), "unit": unit, "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 4),
           "gpu_ms_per_step": round(gpu_ms / args.steps, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config, "roofline": roofline}
    if cpu is not None:
        out["cpu_baseline"] = cpu
    return out


def bench_batch(args):
    """64 uneven LS(HS) 2-level pairs in one call per step (ofri_batch_lshs)."""
    from gaussian_filter import prepareGaussianKernel
    from ofri import ops, synth
    pairs = synth.batch_pairs(args.pairs, workers=16)   # rendered before the GPU is touched
    from ofri.device import DeviceArray, context
    ctx = context()
    fw, ow = prepareGaussianKernel(3.4, 3), prepareGaussianKernel(0.48, 5)
    prm = dict(levels=2, alphas=[45.0, 21.0], niter=600, filt_w=fw, opt_w=ow, ls_alpha=5.0)
    dev = [(DeviceArray.from_host(p[0]), DeviceArray.from_host(p[1])) for p in pairs]
    outs = [(DeviceArray(a.shape), DeviceArray(a.shape)) for a, _ in dev]
    mpx = sum(a.shape[0] * a.shape[1] for a, _ in dev) / 1e6

    def step():
        ops.batch_lshs(dev, out=outs, stats=False, **prm)

    wall, gpu_ms = _timed(step, ctx, args.steps, args.warmup)
    _, st = ops.batch_lshs(dev, out=outs, **prm)
    ls_sw = sum(s[-1]["ls_iters"] * s[-1]["H"] * s[-1]["W"] + s[0]["ls_iters"] * s[0]["H"] * s[0]["W"] for s in st)
    px = sum(s[-1]["H"] * s[-1]["W"] + s[0]["H"] * s[0]["W"] for s in st)
    alg = px * (600 * HS_BYTES_PER_PX_SWEEP + 160.0) + ls_sw * 48.0
    achieved = alg / (gpu_ms / args.steps * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        sizes = [p[0].size for p in pairs]
        a, b = pairs[int(np.argsort(sizes)[len(sizes) // 2])][:2]      # the median-size pair
        t0 = time.perf_counter()
        orc.pyr_lshs(a, b, 2, [45.0, 21.0], 600, fw, ow, 5.0)
        dt = time.perf_counter() - t0
        cpu = {"value": round(a.size / 1e6 / dt, 5), "unit": "Mpix-pairs/s", "cores": orc.get_threads(),
               "kind": "port", "sample": "the batch's median-size pair (%dx%d) through the C oracle's LS(HS) "
                                         "2-level, %.1f s" % (a.shape + (dt,))}
    return _line("Mpix-pairs/s, batched LS(HS) 2-level, %d uneven pairs" % len(pairs), mpx * args.steps / wall,
                 "Mpix-pairs/s", args, wall, gpu_ms,
                 {"workload": "%d pairs H,W in {256..2048}, density U[0.02,0.08], 3 flow families; LS(HS) "
                              "2-level as the default workload, all pairs per call (ofri_batch_lshs)" % len(pairs),
                  "pairs_per_step": len(pairs), "total_mpx": round(mpx, 3), "parallelism": "single GPU"},
                 {"bound": "hbm", "kernel": "whole batch (all launches)", "achieved": round(achieved, 1),
                  "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                  "algorithmic_bytes_per_step": round(alg)}, cpu)


def _oracle_whole(fn, min_s=0.0):
    """Times fn() (a whole-workload run of the CPU oracle) at all host
    threads, repeated until min_s seconds have run (once at least)."""
    from oracle import oracle as orc
    reps, t0 = 0, time.perf_counter()
    while True:
        fn()
        reps += 1
        dt = time.perf_counter() - t0
        if dt >= min_s:
            return reps, dt, orc.get_threads()


def bench_fbls(args):
    """1024^2 Farneback 2-level + Liu-Shen alpha 10 (the example script's
    LS(Farneback) pipeline) as one device-resident call (ofri_pyr_generic,
    the call the drop-in driver makes), inputs resident in HBM; the drop-in
    driver with numpy in and out is timed beside it."""
    from Farneback_PyCL import Farneback_PyCL
    from GenericPyramidalOpticalFlow import genericPyramidalOpticalFlow
    from PhysicsBasedOpticalFlowLiuShen import LiuShenOpticalFlowAlgoAdapter
    from gaussian_filter import prepareGaussianKernel
    from ofri import ops, synth
    from ofri.device import DeviceArray, context
    ctx = context()
    im1, im2, _, _ = synth.config_pair(1)
    n = im1.shape[0]
    consts = Farneback_PyCL()._consts()
    ow = prepareGaussianKernel(0.48, 5)
    d1, d2 = DeviceArray.from_host(im1), DeviceArray.from_host(im2)
    Ua, Va = DeviceArray((n, n)), DeviceArray((n, n))

    def step():
        ops.pyr_generic(d1, d2, 2, None, ow, False, True, True, ls_alpha=10.0, fb_consts=consts, Uacc=Ua, Vacc=Va,
                        stats=False)

    for _ in range(2):
        step()
    wall, gpu_ms = _timed(step, ctx, args.steps, args.warmup)
    _, _, st = ops.pyr_generic(d1, d2, 2, None, ow, False, True, True, ls_alpha=10.0, fb_consts=consts)
    reps = max(2, args.steps // 4)
    t0 = time.perf_counter()
    for _ in range(reps):
        genericPyramidalOpticalFlow(im1, im2, 0, Farneback_PyCL(), 2, 1, 0.48, LiuShenOpticalFlowAlgoAdapter(10))
    dropin_ms = (time.perf_counter() - t0) * 1e3 / reps
    px = n * n + (n // 2) ** 2
    ls_sw = sum(s["ls_iters"] * s["H"] * s["W"] for s in st)
    alg = px * (744.0 + 56.0) + ls_sw * 48.0
    achieved = alg / (gpu_ms / args.steps * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        from oracle import pyramid as opyr
        reps_c, dt, thr = _oracle_whole(lambda: opyr.generic_pyramid(
            im1, im2, 0, opyr.farneback_adapter(consts), 2, 0.48, opyr.ls_adapter(10.0), warping=False,
            interScaling=True, finalScaling=True))
        cpu = {"value": round(reps_c * n * n / 1e6 / dt, 5), "unit": "Mpix-pairs/s", "cores": thr, "kind": "port",
               "sample": "%d run(s) of the whole workload (1 pair %dx%d, Farneback 2-level + Liu-Shen), C oracle "
                         "composed by the driver loop, OpenMP %d threads, %.1f s" % (reps_c, n, n, thr, dt)}
    out = _line("Mpix-pairs/s, LS(Farneback) 2-level 1024^2", n * n / 1e6 * args.steps / wall, "Mpix-pairs/s", args,
                wall, gpu_ms,
                {"workload": "1 pair 1024x1024 Lamb-Oseen density 0.04; Farneback (defaults) FILTER 0, 2 levels, "
                             "Liu-Shen alpha 10, FILTER_OPT 0.48; one device-resident call (ofri_pyr_generic)",
                 "pairs_per_step": 1, "ls_sweeps": [s["ls_iters"] for s in st], "parallelism": "single GPU"},
                {"bound": "hbm", "kernel": "whole pipeline (all launches)", "achieved": round(achieved, 1),
                 "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                 "algorithmic_bytes_per_step": round(alg)}, cpu)
    out["dropin_ms_per_pair"] = round(dropin_ms, 3)
    out["dropin_value"] = round(n * n / 1e6 / (dropin_ms * 1e-3), 3)
    _calibration_copy(ctx)
    return out


def bench_lk(args):
    """2048^2 dense Lucas-Kanade 2-level (halfWindow 13, 5 iterations, FILTER
    2, no warping: the example script's pipeline) as one device-resident call
    (ofri_pyr_generic), inputs resident in HBM; the drop-in driver with numpy
    in and out is timed beside it.  The pair is rendered on the device."""
    from GenericPyramidalOpticalFlow import genericPyramidalOpticalFlow
    from denseLucasKanade_PyCL import denseLucasKanade_PyCl
    from gaussian_filter import prepareGaussianKernel
    from ofri import ops, synth
    from ofri.device import DeviceArray, context
    ctx = context()
    d1, d2, _, _ = synth.device_config_pair(2)
    n = d1.shape[0]
    fw = prepareGaussianKernel(2.0, 3)
    Ua, Va = DeviceArray((n, n)), DeviceArray((n, n))
    status = ops.ByteArray((n, n))

    def step():
        ops.pyr_generic(d1, d2, 2, fw, None, False, True, False, lk=(13, 5), Uacc=Ua, Vacc=Va, status=status,
                        stats=False)

    for _ in range(2):
        step()
    wall, gpu_ms = _timed(step, ctx, args.steps, args.warmup)
    im1, im2 = d1.download(), d2.download()
    reps = max(2, args.steps // 4)
    t0 = time.perf_counter()
    for _ in range(reps):
        genericPyramidalOpticalFlow(im1, im2, 2, denseLucasKanade_PyCl(Niter=5, halfWindow=13), 2, 1, 0.48, None,
                                    warping=False)
    dropin_ms = (time.perf_counter() - t0) * 1e3 / reps
    px = n * n + (n // 2) ** 2
    flops = px * (19000.0 + 8000.0 * 5)      # upper bound: every pixel runs all 5 iterations
    achieved = flops / (gpu_ms / args.steps * 1e-3) / 1e12
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        from oracle import pyramid as opyr

        def whole():
            opyr.generic_pyramid(im1, im2, 2.0, lambda a, b, u, v, lv: orc.lk_dense(a, b, u, v, 13, 5)[:2], 2,
                                 warping=False, interScaling=True, finalScaling=False)
        reps_c, dt, thr = _oracle_whole(whole)
        cpu = {"value": round(reps_c * n * n / 1e6 / dt, 5), "unit": "Mpix-pairs/s", "cores": thr, "kind": "port",
               "sample": "%d run(s) of the whole workload (1 pair %dx%d, dense LK 2-level), C oracle composed by "
                         "the driver loop, OpenMP %d threads, %.1f s" % (reps_c, n, n, thr, dt)}
    out = _line("Mpix-pairs/s, dense LK 2-level 2048^2", n * n / 1e6 * args.steps / wall, "Mpix-pairs/s", args,
                wall, gpu_ms,
                {"workload": "1 pair 2048x2048 shear + turbulence density 0.06 (device-rendered); dense LK "
                             "halfWindow 13, Niter 5, FILTER 2, 2 levels, no warping; one device-resident call "
                             "(ofri_pyr_generic)", "pairs_per_step": 1, "parallelism": "single GPU"},
                {"bound": "valu", "kernel": "whole pipeline (all launches; flops upper bound)",
                 "achieved": round(achieved, 3), "peak": 157.3, "unit": "TFLOP/s",
                 "frac": round(achieved / 157.3, 4), "traffic": None}, cpu)
    out["dropin_ms_per_pair"] = round(dropin_ms, 3)
    out["dropin_value"] = round(n * n / 1e6 / (dropin_ms * 1e-3), 3)
    _calibration_copy(ctx)
    return out


def bench_plain(args):
    """1024^2 LS(HS) 2-level, plain form only (the reference's sweep counts)."""
    from gaussian_filter import prepareGaussianKernel
    from ofri import ops, synth
    from ofri.device import DeviceArray, context

    ctx = context()
    n = args.size
    im1, im2, _, _ = synth.config_pair(4, size=n)
    fw, ow = prepareGaussianKernel(3.4, 3), prepareGaussianKernel(0.48, 5)
    d1, d2 = DeviceArray.from_host(im1), DeviceArray.from_host(im2)
    Ua, Va = DeviceArray((n, n)), DeviceArray((n, n))

    def step():
        ops.pyr_lshs(d1, d2, 2, [45.0, 21.0], 600, fw, ow, 5.0, Uacc=Ua, Vacc=Va, stats=False)

    for _ in range(args.warmup):
        step()
    ctx.sync()
    ctx.timer_start()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    gpu_ms = ctx.timer_stop()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    mpix = n * n / 1e6
    value = mpix * args.steps / wall

    # dominant kernel: HS sweeps at the finest level, timed alone on the context stream
    roof = hs_roofline(ctx, d1, d2, fw, n)

    # the headline run's per-level statistics (for the record)
    _, _, st = ops.pyr_lshs(d1, d2, 2, [45.0, 21.0], 600, fw, ow, 5.0)

    # calibration dispatch for the HBM counters: one plain device-to-device
    # copy of CALIB_BYTES (read CALIB_BYTES, write CALIB_BYTES)
    calib_src = DeviceArray((CALIB_BYTES // 4 // 1024, 1024))
    calib_src.copy()
    ctx.sync()

    out = {
        "metric": "Mpix-pairs/s, LS(HS) 2-level %dx%d plain form" % (n, n),
        "value": round(value, 3),
        "unit": "Mpix-pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "gpu_ms_per_step": round(gpu_ms / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "LS(HS) 2-level %dx%d plain form: HS 600 sweeps (alpha 45 / 21), LS alpha 5 "
                        "(tol 1e-8, max 60), FILTER 3.4/3 taps, FILTER_OPT 0.48/5 taps" % (n, n),
            "pairs_per_step": 1, "height": n, "width": n, "levels": 2,
            "ls_sweeps": [s["ls_iters"] for s in st], "parallelism": "single GPU",
        },
        "roofline": roof,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(im1, im2, fw, ow)
    return out


def hs_roofline(ctx, d1, d2, fw, n):
    """The temporally blocked HS sweep at n x n (600 sweeps, alpha 21) timed
    alone on the context stream with HIP events."""
    from ofri import ops
    from ofri.device import DeviceArray
    depth, _, hs_variant, _ = ctx.query_tuning(n, n)
    niter = 600
    launches = -(-niter // depth)
    f1 = ops.gauss_sep(d1, fw)
    f2 = ops.gauss_sep(d2, fw)
    U, V = DeviceArray.zeros((n, n)), DeviceArray.zeros((n, n))
    lib, h = ctx.lib, ctx.handle
    lib.ofri_hs(h, f1.ptr, f2.ptr, n, n, 21.0, niter, U.ptr, V.ptr, None)  # warm
    reps = 3
    ctx.timer_start()
    for _ in range(reps):
        lib.ofri_hs(h, f1.ptr, f2.ptr, n, n, 21.0, niter, U.ptr, V.ptr, None)
    hs_ms = ctx.timer_stop() / reps
    alg_bytes = (niter * HS_BYTES_PER_PX_SWEEP + HS_SETUP_BYTES_PER_PX) * n * n
    achieved = alg_bytes / (hs_ms * 1e-3) / 1e9
    return {"bound": "hbm",
            "kernel": "hs_pass_k (temporally blocked HS sweep, tile variant %d, %d sweeps/launch)" % (hs_variant, depth),
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profile(),
            "launch_ms": round(hs_ms / launches, 5), "algorithmic_bytes_per_launch": round(alg_bytes / launches),
            "sweeps_per_launch": depth}


def cheb_roofline(ctx, d1, d2, fw, n):
    """The Chebyshev-accelerated HS sweep at n x n (alpha 21, bound 0.99, stop
    test once per launch, a tolerance it never meets): HIP events around 16
    and 8 full launches, per-launch time from the difference."""
    from ofri import ops
    from ofri.device import DeviceArray
    f1 = ops.gauss_sep(d1, fw)
    f2 = ops.gauss_sep(d2, fw)
    U, V = DeviceArray.zeros((n, n)), DeviceArray.zeros((n, n))
    lib, h = ctx.lib, ctx.handle
    depth = 10

    def run(sweeps):
        lib.ofri_memset(h, U.ptr, 0, 4 * n * n)
        lib.ofri_memset(h, V.ptr, 0, 4 * n * n)
        ops.hs_accel(f1, f2, 21.0, sweeps, 1e-30, 0.99, U, V, finish=0.0, launch_test=True)

    ctx.set_tuning(hs_depth=depth, hs_variant=12)   # the residual-mode default at this size
    try:
        run(16 * depth)
        t = {}
        for k in (16, 8):
            reps = 5
            ctx.timer_start()
            for _ in range(reps):
                run(k * depth)
            t[k] = ctx.timer_stop() / reps
    finally:
        ctx.set_tuning()
    launch_ms = (t[16] - t[8]) / 8
    alg = CHEB_BYTES_PER_PX_SWEEP * depth * n * n
    achieved = alg / (launch_ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": "hs_pass_k<8, 16, true, true> (Chebyshev HS sweep, %d sweeps/launch, "
                                      "stop test at the launch's last sweep)" % depth,
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profile("hs_cheb_1024_traffic.json"),
            "launch_ms": round(launch_ms, 5), "algorithmic_bytes_per_launch": round(alg)}


def net_roofline(ctx, net, n):
    """One forward pass of the correction network over a 4-plane n x n input
    (all convolution launches plus the input and head kernels), timed with
    HIP events; flops on the real channels."""
    from ofri.device import DeviceArray
    rng = np.random.default_rng(0)
    x = DeviceArray.from_host(rng.standard_normal((1, 4, n, n)).astype(np.float32))
    c, f = DeviceArray((1, 4, n // 2, n // 2)), DeviceArray((1, 4, n, n))
    for _ in range(3):
        net.predict(x, c, f)
    reps = 20
    ctx.timer_start()
    for _ in range(reps):
        net.predict(x, c, f)
    ms = ctx.timer_stop() / reps
    fl = net.flops(n, n)
    achieved = fl / (ms * 1e-3) / 1e12
    return {"bound": "mfma", "kernel": "correction network forward %dx%d (conv3_lds_k bf16 MFMA + input/head "
                                       "kernels)" % (n, n),
            "achieved": round(achieved, 2), "peak": 2500.0, "unit": "TFLOP/s", "frac": round(achieved / 2500.0, 4),
            "traffic": None, "ms": round(ms, 4), "flops": fl}


def bench_net(args):
    """The correction network's forward pass (both heads) over a 4-plane
    n x n input, alone: TFLOP/s on the real channels."""
    from ofri.device import DeviceArray, context
    from ofri.net import Net
    ctx = context()
    n = args.size
    net = Net.load()
    rng = np.random.default_rng(0)
    x = DeviceArray.from_host(rng.standard_normal((1, 4, n, n)).astype(np.float32))
    c, f = DeviceArray((1, 4, n // 2, n // 2)), DeviceArray((1, 4, n, n))
    wall, gpu_ms = _timed(lambda: net.predict(x, c, f), ctx, args.steps, args.warmup)
    fl = net.flops(n, n)
    value = fl * args.steps / wall / 1e12
    ach = fl / (gpu_ms / args.steps * 1e-3) / 1e12
    roof = {"bound": "mfma", "kernel": "conv3_lds_k (bf16 MFMA) + input/head kernels", "achieved": round(ach, 2),
            "peak": 2500.0, "unit": "TFLOP/s", "frac": round(ach / 2500.0, 4), "traffic": None}
    return _line("TFLOP/s, correction network forward %dx%d" % (n, n), value, "TFLOP/s", args, wall, gpu_ms,
                 {"workload": "correction network forward, 4 input planes %dx%d, 16 layers (both heads)" % (n, n),
                  "height": n, "width": n, "flops_per_pass": fl, "parallelism": "single GPU"}, roof, dtype="bf16")


def bench_hs(args):
    """The HS sweep kernels alone on the workload pair's pre-filtered n x n
    frames, on the library stream: a plain 600-sweep solve (alpha 21) and the
    accelerated (Chebyshev) solve's full launches, each timed with HIP
    events.  A short target for counter collection; value = plain sweeps'
    Mpix-sweeps/s."""
    from gaussian_filter import prepareGaussianKernel
    from ofri import synth
    from ofri.device import DeviceArray, context
    ctx = context()
    n = args.size
    im1, im2, _, _ = synth.config_pair(4, size=n)
    fw = prepareGaussianKernel(3.4, 3)
    d1, d2 = DeviceArray.from_host(im1), DeviceArray.from_host(im2)
    t0 = time.perf_counter()
    plain = hs_roofline(ctx, d1, d2, fw, n)
    cheb = cheb_roofline(ctx, d1, d2, fw, n)
    wall = time.perf_counter() - t0
    _calibration_copy(ctx)
    value = n * n / 1e6 * plain["sweeps_per_launch"] / (plain["launch_ms"] * 1e-3)
    out = _line("Mpix-sweeps/s, plain HS sweep %dx%d" % (n, n), value, "Mpix-sweeps/s", args, wall,
                wall * 1e3, {"workload": "HS sweep kernels alone, %dx%d, alpha 21: plain 600-sweep solve and "
                             "the Chebyshev solve's full launches" % (n, n), "height": n, "width": n,
                             "parallelism": "single GPU, library stream"}, plain)
    out["steps"], out["warmup"] = 1, 0
    out["roofline_cheb"] = cheb
    return out


def bench_lshs(args):
    """The headline: LS(HS) 2-level at 1024^2, learned form against plain
    form.  A step is one judged batch of CHECK_EVERY (32) pairs: the learned
    form with its pair-independent rules, the judge's plain run and EPE check
    on one pair of the batch, and the plain fallback of every pair that
    missed a rule or failed its check -- all inside the timed region.  The
    steps cycle over the workload's pair (Lamb-Oseen, density 0.04) and
    held-out rendered pairs of the three flow families; the batch's first
    (checked) pair rotates through them."""
    from ofri import learned, synth
    from ofri.device import DeviceArray, context
    from ofri.net import Net
    ctx = context()
    n = args.size
    im1, im2, _, _ = synth.config_pair(4, size=n)
    fw, ow = learned.filters()
    pairs = [(DeviceArray.from_host(im1), DeviceArray.from_host(im2))]
    extra, specs = learned.heldout_pairs(args.eval_pairs - 1, n)
    pairs += extra
    fams = ["lamb_oseen"] + [("poiseuille", "lamb_oseen", "turbulence")[sp["family"]] for sp in specs]
    np_ = len(pairs)
    B = learned.CHECK_EVERY
    L = args.lanes or learned.LANES
    net = Net.load()
    # output slots: the t-th occurrence of pair p in a batch writes slot
    # p + np_ * t, and a slot always runs on lane slot % L, so every slot's
    # launch sequences stay captured on its lane
    nslot = np_ * ((B + np_ - 1) // np_)
    outs = [(DeviceArray((n, n)), DeviceArray((n, n))) for _ in range(nslot)]
    cnt = {"step": 0}

    def batch(k):
        seen, jobs, lanes = {}, [], []
        for j in range(B):
            p = (k + j) % np_
            t = seen.get(p, 0)
            seen[p] = t + 1
            slot = p + np_ * t
            jobs.append((pairs[p][0], pairs[p][1], outs[slot][0], outs[slot][1]))
            lanes.append(slot % L)
        return jobs, lanes

    def plain_step():
        jobs, lanes = batch(cnt["step"])
        learned.plain_many(jobs, fw, ow, lane_of=lanes)
        cnt["step"] += 1

    runs = {}

    def learned_step(nt, judge, key):
        jobs, lanes = batch(cnt["step"])
        res = learned.solve_many(jobs, nt, judge, fw, ow, lane_of=lanes)
        cnt["step"] += 1
        runs.setdefault(key, []).extend(res)

    # graph priming: every slot's learned and plain launch sequences, and the
    # judge's plain run of every pair at a batch's first position, are
    # captured on their second call
    for _ in range(2):
        jobs, lanes = batch(0)
        learned.plain_many(jobs, fw, ow, lane_of=lanes)
        for nt in (net, None):
            learned.solve_many(jobs, nt, None, fw, ow, lane_of=lanes)
        for k in range(np_):
            jobs, lanes = batch(k)
            learned.solve_many(jobs[:1], net, learned.Judge(every=1), fw, ow, lane_of=lanes[:1])
    ctx.sync()

    def timed(step):
        cnt["step"] = 0
        return _timed(step, ctx, args.steps, args.warmup)

    wall_p, gpu_p = timed(plain_step)
    judge = learned.Judge()
    runs.clear()
    wall_l, gpu_l = timed(lambda: learned_step(net, judge, "net"))
    judge_z = learned.Judge()
    wall_z, _ = timed(lambda: learned_step(None, judge_z, "nonet"))
    timed_runs = runs["net"][args.warmup * B:]
    fell = sum(r["fell_back"] for r in timed_runs)
    fell_z = sum(r["fell_back"] for r in runs["nonet"][args.warmup * B:])
    mpix = n * n / 1e6
    total = args.steps * B
    value = mpix * total / wall_l
    plain_value = mpix * total / wall_p
    # per pair, untimed: the learned result against the plain one
    rows = []
    for k, (a, b) in enumerate(pairs):
        Up, Vp, _ = learned.plain(a, b, fw, ow)
        P = (Up.download(), Vp.download())
        U, V, info = learned.solve(a, b, net, fw, ow, fallback=True)
        e = learned.epe((U.download(), V.download()), P)
        rows.append({"family": fams[k], "epe_mean": round(e[0], 5), "epe_max": round(e[1], 4),
                     "fell_back": info["fell_back"], "hs_sweeps": [s["hs_iters"] for s in info["stats"]],
                     "ls_sweeps": [s["ls_iters"] for s in info["stats"]],
                     "ls_r_over_n": [round(x[1], 5) for x in learned.residual_ratios(info["stats"])]})
    fam_epe = {f: round(float(np.mean([r["epe_mean"] for r in rows if r["family"] == f])), 5)
               for f in ("poiseuille", "lamb_oseen", "turbulence") if any(r["family"] == f for r in rows)}
    rules = learned.RULES
    out = _line(METRIC, value, "Mpix-pairs/s", args, wall_l, gpu_l,
                {"workload": "LS(HS) 2-level %dx%d: FILTER 3.4/3 taps, HS alpha 45 / 21, FILTER_OPT 0.48/5 taps, "
                             "Liu-Shen alpha 5; plain form = the reference's stopping rules (HS 600 sweeps, "
                             "Liu-Shen tol 1e-8 / max 60); learned form = Chebyshev HS to the relative residual "
                             "2^-22 (pair-independent), correction network, Liu-Shen %d / %d sweeps with a "
                             "relative-residual acceptance test, plain fallback, EPE-to-plain judge on every "
                             "%dth pair (0.01 px), all timed" % (n, n, rules.ls_sweeps[0], rules.ls_sweeps[1], B),
                 "pairs_per_step": B, "height": n, "width": n, "levels": 2, "eval_pairs": np_,
                 "hip_graphs": not args.no_graphs, "lanes": L,
                 "parallelism": "single GPU, pairs on %d concurrent HIP streams (plain form likewise)" % L},
                cheb_roofline(ctx, pairs[0][0], pairs[0][1], fw, n))
    out["roofline_plain_hs"] = hs_roofline(ctx, pairs[0][0], pairs[0][1], fw, n)
    out["learned_vs_plain"] = {
        "learned_value": round(value, 3), "plain_value": round(plain_value, 3),
        "no_network_value": round(mpix * total / wall_z, 3),
        "speedup_vs_plain": round(value / plain_value, 3), "plain_ms_per_pair": round(wall_p * 1e3 / total, 4),
        "learned_ms_per_pair": round(wall_l * 1e3 / total, 4),
        "fallback_rate": round(fell / max(1, len(timed_runs)), 4),
        "no_network_fallback_rate": round(fell_z / max(1, total), 4),
        "judge": judge.summary(), "no_network_judge": judge_z.summary(),
        "epe_to_plain_mean": round(float(np.mean([r["epe_mean"] for r in rows])), 5),
        "epe_to_plain_max_pair_mean": round(float(np.max([r["epe_mean"] for r in rows])), 5),
        "epe_to_plain_by_family": fam_epe,
        "rules": {"hs_rel": rules.hs_rel, "ls_sweeps": rules.ls_sweeps, "ls_accept": rules.ls_accept,
                  "hs_budget": rules.hs_budget},
        "arithmetic": "learned HS phase: Chebyshev sweeps in float32 (fused multiply-adds) to the relative "
                      "residual hs_rel, then the reference's update (float32 fields, float64 stencil sums); "
                      "where the plain 600 sweeps end below hs_rel the learned HS stops at a looser residual. "
                      "Liu-Shen: the reference's update on every tested sweep, float32 FMA sweeps between; "
                      "network: bfloat16 operands, float32 accumulation",
        "pairs": rows,
    }
    out["roofline_net"] = net_roofline(ctx, net, n)
    # calibration dispatch for the HBM counters: one plain device-to-device
    # copy of CALIB_BYTES (read CALIB_BYTES, write CALIB_BYTES)
    DeviceArray((CALIB_BYTES // 4 // 1024, 1024)).copy()
    ctx.sync()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(im1, im2, fw, ow)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (pairs are independent, "
              "see the sharding notes)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.no_graphs:
        from ofri.device import context
        context().set_graphs(False)
    if args.workload != "lshs":
        fn = {"plain": bench_plain, "batch": bench_batch, "fbls": bench_fbls, "lk": bench_lk,
              "net": bench_net, "hs": bench_hs}[args.workload]
        print(json.dumps(fn(args)))
        return
    print(json.dumps(bench_lshs(args)))


if __name__ == "__main__":
    main()
