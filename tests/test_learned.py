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
    # one level: the accumulated flow is HS + (LS - HS)
    U1, V1, taps1, _ = ops.pyr_lshs_taps(im1, im2, **_lshs(ops, 1))
    t = taps1.download().reshape(4, 256, 320)
    assert np.abs(t[0] + t[2] - U1.download()).max() <= 1e-6 * max(1.0, np.abs(t[0]).max())
    assert np.abs(t[1] + t[3] - V1.download()).max() <= 1e-6 * max(1.0, np.abs(t[1]).max())


# ---------------------------------------------------------------- learned form

@pytest.fixture(scope="module")
def trained():
    """The committed correction network, a held-out Lamb-Oseen pair (seeds
    outside the training set) and its plain results and residual targets."""
    from ofri import learned
    from ofri.net import Net
    net = Net.load()
    pairs, specs = learned.heldout_pairs(2, 1024)
    assert specs[1]["family"] == 1
    pair = pairs[1]
    fw, ow = learned.filters()
    Up, Vp, ht, lt, st = learned.plain_targets(pair[0], pair[1], fw, ow)
    return dict(net=net, pair=pair, plain=(Up.download(), Vp.download()), ht=ht, lt=lt, st=st, fw=fw, ow=ow)


def test_learned_form_reaches_the_plain_residual_with_fewer_sweeps(trained):
    """The trained network is what lets Liu-Shen reach the plain residual:
    with it every level converges in a fraction of the plain 60 sweeps and
    lands within 0.01 px (mean EPE) of the plain result; the same method
    without the network misses the target within the budget or needs twice
    the sweeps."""
    from ofri import learned
    t = trained
    a, b = t["pair"]
    U, V, info = learned.solve(a, b, t["net"], t["ht"], t["lt"], t["fw"], t["ow"], fallback=False)
    assert info["converged"]
    ls = [s["ls_iters"] for s in info["stats"]]
    e_mean, e_max = learned.epe((U.download(), V.download()), t["plain"])
    U0, V0, info0 = learned.solve(a, b, None, t["ht"], t["lt"], t["fw"], t["ow"], fallback=False)
    ls0 = [s["ls_iters"] for s in info0["stats"]]
    print("[learned] LS sweeps %s (no network %s, plain %s), EPE to plain mean %.5f max %.4f px"
          % (ls, ls0, [s["ls_iters"] for s in t["st"]], e_mean, e_max))
    assert all(s["hs_res"] <= tol for s, tol in zip(info["stats"], t["ht"]))
    assert all(s["ls_err"] <= tol for s, tol in zip(info["stats"], t["lt"]))
    assert max(ls) <= 30 and 2 * sum(ls) <= sum(s["ls_iters"] for s in t["st"])
    assert not info0["converged"] or sum(ls0) >= 2 * sum(ls)
    assert e_mean <= 0.01


def test_learned_fallback_returns_the_plain_result(trained):
    """A sweep budget too small to reach the targets triggers the fallback,
    whose result is the plain pyramid's, bit for bit."""
    from ofri import learned, ops
    t = trained
    a, b = t["pair"]
    U, V, info = learned.solve(a, b, t["net"], t["ht"], t["lt"], t["fw"], t["ow"], fallback=True, hs_budget=8)
    assert info["fell_back"] and not info["converged"]
    assert np.array_equal(U.download(), t["plain"][0]) and np.array_equal(V.download(), t["plain"][1])
    Ur, Vr, _ = ops.pyr_lshs(a, b, filt_w=t["fw"], opt_w=t["ow"], **learned.WORKLOAD)
    assert np.array_equal(Ur.download(), t["plain"][0])


def test_learned_form_replays_identically(trained):
    """Eager run, graph capture and replays of the learned form give the same bits."""
    from ofri import learned
    from ofri.device import DeviceArray
    t = trained
    a, b = t["pair"]
    Ua, Va = DeviceArray((1024, 1024)), DeviceArray((1024, 1024))
    outs = []
    for _ in range(4):
        U, V, info = learned.solve(a, b, t["net"], t["ht"], t["lt"], t["fw"], t["ow"], Uacc=Ua, Vacc=Va,
                                   fallback=False)
        outs.append((U.download(), V.download(), [s["ls_iters"] for s in info["stats"]]))
    for o in outs[1:]:
        assert np.array_equal(o[0], outs[0][0]) and np.array_equal(o[1], outs[0][1]) and o[2] == outs[0][2]


def test_queued_pairs_flag_on_the_device_and_fall_back(trained):
    """solve_many queues pairs with their convergence flags written on the
    device: a pair within budget keeps the learned result (the same bits as
    a single call), a pair whose budget is too small gets the plain result."""
    from ofri import learned
    from ofri.device import DeviceArray
    t = trained
    a, b = t["pair"]
    U1, V1 = DeviceArray((1024, 1024)), DeviceArray((1024, 1024))
    U2, V2 = DeviceArray((1024, 1024)), DeviceArray((1024, 1024))
    Us, Vs, info = learned.solve(a, b, t["net"], t["ht"], t["lt"], t["fw"], t["ow"], fallback=False)
    assert info["converged"]
    fell = learned.solve_many([(a, b, t["ht"], t["lt"], U1, V1)], t["net"], t["fw"], t["ow"])
    assert fell == [False]
    assert np.array_equal(U1.download(), Us.download()) and np.array_equal(V1.download(), Vs.download())
    fell = learned.solve_many([(a, b, t["ht"], t["lt"], U1, V1), (a, b, t["ht"], t["lt"], U2, V2)], t["net"],
                              t["fw"], t["ow"], hs_budget=8)
    assert fell == [True, True]
    assert np.array_equal(U2.download(), t["plain"][0]) and np.array_equal(V2.download(), t["plain"][1])
