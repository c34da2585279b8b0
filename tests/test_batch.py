"""Batched LS(HS) (ofri_batch_lshs): every pair of an uneven batch equals the
single-pair device pyramid bit for bit (which the parity tests tie to the
oracle), under the default and tuned tile shapes, eager and graph-replayed,
and the batched driver entry point equals per-pair driver calls."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

SIZES = [(256, 384), (130, 200), (512, 256), (64, 96), (384, 384), (100, 77)]


@pytest.fixture(scope="module")
def ops():
    from ofri import ops as _ops
    from ofri.device import context
    context()
    return _ops


def _pairs(sizes, seed=5):
    from ofri import synth
    rng = np.random.default_rng(seed)
    out = []
    for k, (H, W) in enumerate(sizes):
        fam = ("poiseuille", "lamb_oseen", "turbulence")[k % 3]
        im1, im2, _, _ = synth.piv_pair(H, W, 0.05, fam, rng)
        out.append((im1, im2))
    return out


def _params():
    from gaussian_filter import prepareGaussianKernel
    return dict(levels=2, alphas=[45.0, 21.0], niter=600, filt_w=prepareGaussianKernel(3.4, 3),
                opt_w=prepareGaussianKernel(0.48, 5), ls_alpha=5.0)


def _singles(ops, pairs, prm):
    res = []
    for a, b in pairs:
        U, V, st = ops.pyr_lshs(a, b, **prm)
        res.append((U.download(), V.download(), st))
    return res


def test_batch_equals_single_pair_pyramids(ops):
    pairs = _pairs(SIZES)
    prm = _params()
    ref = _singles(ops, pairs, prm)
    out, stats = ops.batch_lshs(pairs, **prm)
    for i, ((U, V), (Ur, Vr, str_)) in enumerate(zip(out, ref)):
        assert np.array_equal(U.download(), Ur) and np.array_equal(V.download(), Vr), SIZES[i]
        for s, r in zip(stats[i], str_):
            assert (s["H"], s["W"], s["ls_iters"]) == (r["H"], r["W"], r["ls_iters"])
            assert s["hs_err"] == r["hs_err"] and s["ls_err"] == r["ls_err"]


def test_batch_small_pair_matches_oracle(ops):
    """One pair of the batch against the oracle's whole pyramid directly."""
    pairs = _pairs([(96, 128), (64, 80)], seed=9)
    prm = _params()
    prm["niter"] = 80
    out, _ = ops.batch_lshs(pairs, **prm)
    a, b = pairs[1]
    U32, V32, _ = orc.pyr_lshs(a, b, 2, [45.0, 21.0], 80, prm["filt_w"], prm["opt_w"], 5.0)
    assert np.abs(out[1][0].download() - U32).max() <= 1e-5
    assert np.abs(out[1][1].download() - V32).max() <= 1e-5


@pytest.mark.parametrize("hs_variant,hs_depth,ls_variant,ls_depth", [(2, 3, 2, 2), (12, 10, 4, 4), (6, 7, 1, 5)])
def test_batch_tile_shapes_identical(ops, hs_variant, hs_depth, ls_variant, ls_depth):
    from ofri.device import context
    pairs = _pairs(SIZES[:4], seed=6)
    prm = _params()
    prm["niter"] = 47
    ref = _singles(ops, pairs, prm)
    ctx = context()
    ctx.set_tuning(hs_depth=hs_depth, ls_depth=ls_depth, hs_variant=hs_variant, ls_variant=ls_variant)
    try:
        out, _ = ops.batch_lshs(pairs, **prm)
    finally:
        ctx.set_tuning()
    for (U, V), (Ur, Vr, _) in zip(out, ref):
        assert np.array_equal(U.download(), Ur) and np.array_equal(V.download(), Vr)


def test_batch_graph_replays_and_eager_agree(ops):
    from ofri.device import DeviceArray, context
    pairs = _pairs(SIZES[:5], seed=7)
    prm = _params()
    prm["niter"] = 60
    dev = [(DeviceArray.from_host(a), DeviceArray.from_host(b)) for a, b in pairs]
    out = [(DeviceArray(a.shape), DeviceArray(a.shape)) for a, _ in pairs]
    ctx = context()
    ctx.set_graphs(False)
    try:
        ops.batch_lshs(dev, out=out, **prm)
    finally:
        ctx.set_graphs(True)
    eager = [(U.download(), V.download()) for U, V in out]
    for rep in range(4):            # eager, eager/capture, replay, replay
        for U, V in out:
            U.upload(np.full(U.shape, np.nan, np.float32))
            V.upload(np.full(V.shape, np.nan, np.float32))
        _, stats = ops.batch_lshs(dev, out=out, **prm)
        for (U, V), (Ue, Ve) in zip(out, eager):
            assert np.array_equal(U.download(), Ue) and np.array_equal(V.download(), Ve), rep
        assert all(s[-1]["ls_iters"] >= 1 for s in stats)


def test_batch_stops_each_pair_on_its_own(ops):
    """A tolerance that stops some pairs' Liu-Shen early and not others: each
    pair's sweep count and flow equal its single-pair run."""
    pairs = _pairs([(128, 160), (96, 96), (200, 120)], seed=8)
    prm = _params()
    prm["niter"] = 30
    prm["ls_tol"] = 4e-6
    ref = _singles(ops, pairs, prm)
    iters = {r[2][-1]["ls_iters"] for r in ref}
    out, stats = ops.batch_lshs(pairs, **prm)
    for i, ((U, V), (Ur, Vr, sr)) in enumerate(zip(out, ref)):
        assert [s["ls_iters"] for s in stats[i]] == [s["ls_iters"] for s in sr]
        assert np.array_equal(U.download(), Ur) and np.array_equal(V.download(), Vr)
    print("[batch stop] per-pair fine-level LS sweeps:", sorted(iters))


def test_batch_driver_entry_point(ops):
    from GenericPyramidalOpticalFlow import genericPyramidalOpticalFlow, genericPyramidalOpticalFlowBatch
    from HornSchunck import HSOpticalFlowAlgoAdapter
    from PhysicsBasedOpticalFlowLiuShen import LiuShenOpticalFlowAlgoAdapter
    pairs = _pairs([(128, 96), (64, 128)], seed=10)
    hs = HSOpticalFlowAlgoAdapter([21, 45], 40)
    res = genericPyramidalOpticalFlowBatch(pairs, 3.4, hs, 2, 1, 0.48, LiuShenOpticalFlowAlgoAdapter(5))
    assert hs.alphas == []
    for (a, b), (U, V) in zip(pairs, res):
        Ur, Vr = genericPyramidalOpticalFlow(a, b, 3.4, HSOpticalFlowAlgoAdapter([21, 45], 40), 2, 1, 0.48,
                                             LiuShenOpticalFlowAlgoAdapter(5))
        assert np.array_equal(U, Ur) and np.array_equal(V, Vr)


def test_full_uneven_batch_of_64_pairs(ops):
    """The batched configuration itself: 64 device-rendered pairs of uneven
    sizes (256..2048 per side, three flow families) in one ofri_batch_lshs
    call at the reference's sweep counts.  Every flow is finite and close to
    its rendered truth, a second call gives the same bits, the three smallest
    pairs equal the f32 oracle's pyramid and three more their single-pair
    call."""
    from ofri import synth
    from ofri.device import DeviceArray
    pairs, specs = synth.device_batch_pairs(64, truth=True)
    prm = _params()
    jobs = [(a, b) for a, b, _, _ in pairs]
    out, st = ops.batch_lshs(jobs, **prm)
    out2 = [(DeviceArray(a.shape), DeviceArray(a.shape)) for a, _ in jobs]
    ops.batch_lshs(jobs, out=out2, stats=False, **prm)
    epes = []
    for k, ((U, V), (U2, V2), (_, _, ut, vt)) in enumerate(zip(out, out2, pairs)):
        u, v = U.download(), V.download()
        assert np.isfinite(u).all() and np.isfinite(v).all(), k
        assert np.array_equal(u, U2.download()) and np.array_equal(v, V2.download()), k
        e = np.sqrt((u - ut.download()) ** 2 + (v - vt.download()) ** 2)[32:-32, 32:-32].mean()
        epes.append(float(e))
        assert [s["ls_iters"] for s in st[k]] and all(1 <= s["ls_iters"] <= 60 for s in st[k])
    print("[batch 64] total %.1f Mpx, sizes %d..%d px, mean EPE to truth per pair: min %.3f median %.3f max %.3f px"
          % (sum(a.shape[0] * a.shape[1] for a, _ in jobs) / 1e6, min(a.shape[0] * a.shape[1] for a, _ in jobs),
             max(a.shape[0] * a.shape[1] for a, _ in jobs), min(epes), float(np.median(epes)), max(epes)))
    assert max(epes) < 1.0
    order = np.argsort([a.shape[0] * a.shape[1] for a, _ in jobs])
    for k in order[:3]:
        a, b = jobs[k][0].download(), jobs[k][1].download()
        U32, V32, _ = orc.pyr_lshs(a, b, 2, [45.0, 21.0], 600, prm["filt_w"], prm["opt_w"], 5.0)
        d = max(np.abs(out[k][0].download() - U32).max(), np.abs(out[k][1].download() - V32).max())
        print("[batch 64] pair %d (%dx%d) max|d| vs oracle-f32 %.3g px" % ((k,) + a.shape + (d,)))
        assert d <= 1e-5
    for k in order[[20, 40, 63]]:
        U, V, _ = ops.pyr_lshs(jobs[k][0], jobs[k][1], **prm)
        assert np.array_equal(U.download(), out[k][0].download()) and np.array_equal(V.download(),
                                                                                      out[k][1].download())


def test_tuning_classes_pick_the_launch_shape(ops, tmp_path):
    """Size classes (the autotuner's cached per-size results): a plain solve
    takes its class's shape, ofri_set_tuning still overrides, a batch takes
    the class of its total pixel count per level (here both classes, with
    the single-pair bits), and a cache file for this GPU's architecture is
    installed as classes with boundaries at the geometric means."""
    from ofri import autotune
    from ofri.device import context
    ctx = context()
    try:
        ctx.set_tuning_classes([(0, 2, 3, 2, 2), (100000, 6, 7, 1, 5)])
        assert ctx.query_tuning(512, 512) == (7, 5, 6, 1)        # (hs_depth, ls_depth, hs_variant, ls_variant)
        assert ctx.query_tuning(200, 300) == (3, 2, 2, 2)
        ctx.set_tuning(hs_depth=4)
        assert ctx.query_tuning(512, 512) == (4, 5, 6, 1)
        ctx.set_tuning()
        pairs = _pairs([(256, 384), (130, 200)], seed=12)
        prm = _params()
        prm["niter"] = 80
        out, _ = ops.batch_lshs(pairs, **prm)
        ref = _singles(ops, pairs, prm)
        for (U, V), (Ur, Vr, _) in zip(out, ref):
            assert np.array_equal(U.download(), Ur) and np.array_equal(V.download(), Vr)
        cache = tmp_path / "tuning.json"
        autotune.save({512: {"hs": [4, 5, 0.1], "ls": [0, 5, 0.2]}, 1024: {"hs": [10, 10, 0.3], "ls": [2, 4, 0.4]}},
                      path=str(cache))
        rows = autotune.apply(ctx, path=str(cache))
        assert rows == [(0, 5, 5, 1, 5), (524288, 11, 10, 3, 4)]
        assert ctx.query_tuning(512, 512) == (5, 5, 5, 1) and ctx.query_tuning(1024, 1024) == (10, 4, 11, 3)
    finally:
        ctx.set_tuning_classes([])
        ctx.set_tuning()
