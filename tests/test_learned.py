"""The learned form: the device particle renderer, the pyramid's exported
per-level labels, and the learned LS(HS) form with the committed network."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ops():
    from ofri import ops as _ops
    from ofri.device import context
    context()
    return _ops


def _lshs(ops, levels=2):
    from gaussian_filter import prepareGaussianKernel
    return dict(levels=levels, alphas=[45.0, 21.0][-levels:] if levels == 1 else [45.0, 21.0], niter=600,
                filt_w=prepareGaussianKernel(3.4, 3), opt_w=prepareGaussianKernel(0.48, 5), ls_alpha=5.0)


def test_render_is_a_function_of_the_seed(ops):
    a1, a2 = ops.render_pair(192, 256, 1, 0.05, 77, [128.0, 96.0, 20.0, 3.0])
    b1, b2 = ops.render_pair(192, 256, 1, 0.05, 77, [128.0, 96.0, 20.0, 3.0])
    c1, _ = ops.render_pair(192, 256, 1, 0.05, 78, [128.0, 96.0, 20.0, 3.0])
    x1, x2, y1 = a1.download(), a2.download(), c1.download()
    assert np.array_equal(x1, b1.download()) and np.array_equal(x2, b2.download())
    assert not np.array_equal(x1, y1)
    # particle images: bright spots on a dark background, noise sigma 2
    assert 15 < x1.mean() < 120 and x1.max() > 200 and np.percentile(x1, 10) < 8


@pytest.mark.parametrize("family", [0, 1, 2])
def test_rendered_pairs_carry_their_flow(ops, family):
    """The plain pyramid recovers the rendered flow field (interior, mean EPE)."""
    H = W = 256
    rng = np.random.default_rng(family)
    modes = None
    if family == 0:
        par = [3.0]
    elif family == 1:
        par = [128.0, 128.0, 25.0, 3.0]
    else:
        n = 64
        # large-scale modes only (wavelengths 64..128 px): the smoothness
        # terms of HS / Liu-Shen cannot follow finer turbulence
        k = np.exp(rng.uniform(np.log(2 * np.pi / 128), np.log(2 * np.pi / 64), n))
        ang, phi = rng.uniform(0, 2 * np.pi, n), rng.uniform(0, 2 * np.pi, n)
        modes = np.stack([k * np.cos(ang), k * np.sin(ang), phi, k ** (-5.0 / 6.0) * np.sqrt(k) / k])
        par = [2.0, 3.0]
    im1, im2, u, v = ops.render_pair(H, W, family, 0.06, 1000 + family, par, modes, truth=True)
    U, V, _ = ops.pyr_lshs(im1, im2, **_lshs(ops))
    ut, vt = u.download(), v.download()
    assert np.sqrt(ut ** 2 + vt ** 2).max() == pytest.approx(3.0, rel=0.05)
    e = np.sqrt((U.download() - ut) ** 2 + (V.download() - vt) ** 2)[24:-24, 24:-24]
    print("[render family %d] mean EPE to truth %.3f px" % (family, e.mean()))
    assert e.mean() < 0.25


def test_taps_run_equals_the_plain_pyramid(ops):
    im1, im2 = ops.render_pair(256, 320, 1, 0.05, 5, [160.0, 128.0, 20.0, 3.0])
    prm = _lshs(ops)
    U, V, st = ops.pyr_lshs(im1, im2, **prm)
    Ut, Vt, taps, stt = ops.pyr_lshs_taps(im1, im2, **prm)
    assert np.array_equal(U.download(), Ut.download()) and np.array_equal(V.download(), Vt.download())
    for a, b in zip(st, stt):
        assert a["ls_iters"] == b["ls_iters"] and a["ls_err"] == b["ls_err"] and a["hs_err"] == b["hs_err"]
        assert b["hs_iters"] == 600 and 0 < b["hs_res"] < 1e-6
    # Liu-Shen planes tapped k sweeps before the stop: the pyramid's result is
    # unchanged, and the coarse level's planes are those of a run stopped
    # k sweeps earlier there
    Ub, Vb, tapb, stb = ops.pyr_lshs_taps(im1, im2, ls_tap_back=(5, 8), **prm)
    assert np.array_equal(U.download(), Ub.download()) and np.array_equal(V.download(), Vb.download())
    assert [s["ls_iters"] for s in stb] == [5, 8]
    _, _, tap55, _ = ops.pyr_lshs_taps(im1, im2, **dict(prm, ls_maxnum=55))
    nc = 128 * 160
    assert np.array_equal(tapb.download()[:4 * nc], tap55.download()[:4 * nc])
    assert not np.array_equal(tapb.download()[:4 * nc], taps.download()[:4 * nc])
    # one level: the accumulated flow is HS + (LS - HS)
    U1, V1, taps1, _ = ops.pyr_lshs_taps(im1, im2, **_lshs(ops, 1))
    t = taps1.download().reshape(4, 256, 320)
    assert np.abs(t[0] + t[2] - U1.download()).max() <= 1e-6 * max(1.0, np.abs(t[0]).max())
    assert np.abs(t[1] + t[3] - V1.download()).max() <= 1e-6 * max(1.0, np.abs(t[1]).max())


# ---------------------------------------------------------------- learned form

@pytest.fixture(scope="module")
def trained():
    """The committed correction network and held-out pairs (seeds outside
    the training and calibration sets) of the Lamb-Oseen and the
    shear + turbulence families, with their plain results."""
    from ofri import learned
    from ofri.net import Net
    net = Net.load()
    pairs, specs = learned.heldout_pairs(3, 1024)
    assert [s["family"] for s in specs] == [0, 1, 2]
    fw, ow = learned.filters()
    out = dict(net=net, fw=fw, ow=ow, pairs={}, plain={})
    for fam in (1, 2):
        a, b = pairs[fam]
        Up, Vp, st = learned.plain(a, b, fw, ow, stats=True)
        out["pairs"][fam] = (a, b)
        out["plain"][fam] = (Up.download(), Vp.download(), st)
    return out


@pytest.mark.parametrize("family", [1, 2])
def test_learned_form_meets_its_rules_and_lands_on_the_plain_result(trained, family):
    """With pair-independent rules (no plain solve of the pair): HS reaches
    its relative residual in a fraction of the plain 600 sweeps, Liu-Shen
    runs its fixed sweeps, the acceptance test passes, and the result lies
    within 0.01 px mean EPE of the plain result -- for the Lamb-Oseen and the
    turbulence family alike.  Without the network the same rules reject the
    pair (Liu-Shen from HS's flow alone is far from the plain iterate)."""
    from ofri import learned
    t = trained
    a, b = t["pairs"][family]
    U, V, info = learned.solve(a, b, t["net"], t["fw"], t["ow"], fallback=False)
    st = info["stats"]
    e_mean, e_max = learned.epe((U.download(), V.download()), t["plain"][family][:2])
    ratios = learned.residual_ratios(st)
    U0, V0, info0 = learned.solve(a, b, None, t["fw"], t["ow"], fallback=False)
    print("[learned family %d] HS sweeps %s, LS sweeps %s, r/n %s, EPE to plain mean %.5f max %.4f px; "
          "no network: accepted %s, r/n %s" % (family, [s["hs_iters"] for s in st], [s["ls_iters"] for s in st],
                                              ratios, e_mean, e_max, info0["converged"],
                                              learned.residual_ratios(info0["stats"])))
    assert info["converged"]
    assert [s["ls_iters"] for s in st] == list(learned.LS_SWEEPS)
    for (rh, rl), hs_rel, acc in zip(ratios, learned.HS_REL, learned.LS_ACCEPT):
        assert rh <= hs_rel * (1 + 1e-6) and rl <= acc * (1 + 1e-6)
    assert all(s["hs_iters"] <= 200 for s in st)
    assert e_mean <= 0.01
    assert not info0["converged"]


def test_learned_fallback_returns_the_plain_result(trained):
    """A sweep budget too small to reach the rules triggers the fallback,
    whose result is the plain pyramid's, bit for bit."""
    from ofri import learned
    t = trained
    a, b = t["pairs"][1]
    U, V, info = learned.solve(a, b, t["net"], t["fw"], t["ow"], fallback=True,
                               rules=learned.Rules(hs_budget=8))
    assert info["fell_back"] and not info["converged"]
    assert np.array_equal(U.download(), t["plain"][1][0]) and np.array_equal(V.download(), t["plain"][1][1])


def test_short_first_finishing_launch_moves_only_the_finishing_sweeps(trained):
    """The first plain finishing launch of f sweeps (HS_FINISH_FIRST) leaves
    the accelerated phase as it was: at each level the HS sweep count
    differs from that of the run whose first finishing launch has the fused
    depth d (12 at 512^2, 10 at 1024^2: the residual-mode defaults) by f - d
    (the short launch met the target) or by f (one more launch of d did).
    Both runs meet the rules and land within the EPE bar of each other."""
    from ofri import learned
    t = trained
    a, b = t["pairs"][2]
    f = learned.HS_FINISH_FIRST
    assert all(0 < x for x in f)
    U, V, info = learned.solve(a, b, t["net"], t["fw"], t["ow"], fallback=False)
    U0, V0, info0 = learned.solve(a, b, t["net"], t["fw"], t["ow"], fallback=False,
                                  rules=learned.Rules(hs_finish_first=(0, 0)))
    hs = [s["hs_iters"] for s in info["stats"]]
    hs0 = [s["hs_iters"] for s in info0["stats"]]
    print("[first finishing launch %s] HS sweeps %s, with launches of the fused depth %s" % (f, hs, hs0))
    assert info["converged"] and info0["converged"]
    for l, d in enumerate((12, 10)):
        assert hs[l] - hs0[l] in (f[l] - d, f[l])
    assert learned.epe((U.download(), V.download()), (U0.download(), V0.download()))[0] <= learned.EPE_BAR
    with pytest.raises(Exception, match="HS budget / finish"):
        learned.solve(a, b, t["net"], t["fw"], t["ow"], fallback=False, rules=learned.Rules(hs_finish_first=(-1, 2)))


def test_learned_form_replays_identically(trained):
    """Eager run, graph capture and replays of the learned form give the same bits."""
    from ofri import learned
    from ofri.device import DeviceArray
    t = trained
    a, b = t["pairs"][2]
    Ua, Va = DeviceArray((1024, 1024)), DeviceArray((1024, 1024))
    outs = []
    for _ in range(4):
        U, V, info = learned.solve(a, b, t["net"], t["fw"], t["ow"], Uacc=Ua, Vacc=Va, fallback=False)
        outs.append((U.download(), V.download(), [s["hs_iters"] for s in info["stats"]]))
    for o in outs[1:]:
        assert np.array_equal(o[0], outs[0][0]) and np.array_equal(o[1], outs[0][1]) and o[2] == outs[0][2]


def test_queued_pairs_the_judge_and_the_fallback(trained):
    """solve_many queues pairs with their acceptance flags on the device.  A
    checked pair whose EPE to plain is within the bar keeps the learned bits
    (those of a single call) and the judge records its EPE; with an
    impossible bar the checked pair takes the plain result and the judge
    checks the pairs after it; a pair whose budget is too small runs the
    plain pyramid."""
    from ofri import learned
    from ofri.device import DeviceArray
    t = trained
    a, b = t["pairs"][1]
    Us, Vs, info = learned.solve(a, b, t["net"], t["fw"], t["ow"], fallback=False)
    assert info["converged"]
    bufs = [(DeviceArray((1024, 1024)), DeviceArray((1024, 1024))) for _ in range(3)]
    judge = learned.Judge(every=2)
    res = learned.solve_many([(a, b, U, V) for U, V in bufs], t["net"], judge, t["fw"], t["ow"])
    assert [r["checked"] for r in res] == [True, False, True]
    assert not any(r["fell_back"] for r in res) and all(r["accepted"] for r in res)
    for U, V in bufs:
        assert np.array_equal(U.download(), Us.download()) and np.array_equal(V.download(), Vs.download())
    e = learned.epe((Us.download(), Vs.download()), t["plain"][1][:2])[0]
    assert judge.checks == 2 and judge.failed == 0 and abs(judge.epes[0] - e) <= 1e-6 + 1e-4 * e
    # an impossible bar: the checked pair falls back to the plain result and
    # the next pairs are all checked
    strict = learned.Judge(every=4, bar=0.0, after_fail=2)
    res = learned.solve_many([(a, b, U, V) for U, V in bufs], t["net"], strict, t["fw"], t["ow"])
    assert [r["fell_back"] for r in res] == [True, False, False] and strict.failed == 1
    assert np.array_equal(bufs[0][0].download(), t["plain"][1][0])
    assert strict.pick() and strict.pick() and not strict.pick()
    # a budget too small: every pair runs the plain pyramid
    res = learned.solve_many([(a, b, U, V) for U, V in bufs[:2]], t["net"], None, t["fw"], t["ow"],
                             rules=learned.Rules(hs_budget=8))
    assert [r["fell_back"] for r in res] == [True, True] and not any(r["accepted"] for r in res)
    assert np.array_equal(bufs[1][0].download(), t["plain"][1][0])


def test_lanes_give_the_bits_of_one_stream(trained):
    """Pairs spread over lanes (concurrent streams, each with its own scratch
    and captured graphs) give the bits of the same pairs queued one after
    another on the library stream -- on the eager call, the capture and the
    replays, with the default round-robin and with explicit lanes; the
    plain form likewise.  A lane cannot be destroyed on its own, has no lanes
    and the lane index is bounded."""
    import ctypes as C
    from ofri import learned, _capi
    from ofri.device import DeviceArray, context
    t = trained
    ref = {}
    for fam in (1, 2):
        a, b = t["pairs"][fam]
        U, V, _ = learned.solve(a, b, t["net"], t["fw"], t["ow"], fallback=False)
        ref[fam] = (U.download(), V.download())
    fams = [1, 2, 2, 1, 1, 2]
    bufs = [(DeviceArray((1024, 1024)), DeviceArray((1024, 1024))) for _ in fams]
    jobs = [t["pairs"][f] + bufs[k] for k, f in enumerate(fams)]
    for kw in [dict(lanes=1)] + [dict(lanes=4)] * 3 + [dict(lane_of=[3, 3, 0, 1, 0, 2])] * 3:
        for U, V in bufs:
            U.upload(np.full((1024, 1024), np.nan, np.float32))
            V.upload(np.full((1024, 1024), np.nan, np.float32))
        res = learned.solve_many(jobs, t["net"], None, t["fw"], t["ow"], **kw)
        assert all(r["accepted"] and not r["fell_back"] for r in res), kw
        for k, f in enumerate(fams):
            assert np.array_equal(bufs[k][0].download(), ref[f][0]), (kw, k)
            assert np.array_equal(bufs[k][1].download(), ref[f][1]), (kw, k)
    for rep in range(3):
        learned.plain_many(jobs[:4], t["fw"], t["ow"], lanes=4)
        for k, f in enumerate(fams[:4]):
            assert np.array_equal(bufs[k][0].download(), t["plain"][f][0]), (rep, k)
            assert np.array_equal(bufs[k][1].download(), t["plain"][f][1]), (rep, k)
    ctx = context()
    lib = ctx.lib
    lane = ctx.lane(0)
    h = C.c_void_p()
    assert lib.ofri_destroy(lane.handle) == _capi.OFRI_E_STATE
    assert lib.ofri_lane(lane.handle, 0, C.byref(h)) == _capi.OFRI_E_STATE
    assert lib.ofri_lane(ctx.handle, 8, C.byref(h)) == _capi.OFRI_E_ARG
    assert lib.ofri_fork(ctx.handle, 9) == _capi.OFRI_E_ARG
    assert lib.ofri_join(lane.handle, 1) == _capi.OFRI_E_STATE
