"""Whole pipelines as the reference's example scripts run them, through the
drop-in driver, against the oracle's composition of the same driver loop
(oracle/pyramid.py): LS(Farneback) 2-level (examples/LiuSE_Farneback_Fs0_0_
PyrLvls2.py:51-76), dense LK 2-level without warping (examples/denseLK_Fs2_0_
PyrLvls2.py:49-73) and LS(HS) 2-level at 1024^2 (examples/LiuSE_PyHSchunck_
Fs3_4_PyrLvls2.py:54-139), from 256^2 up to the configurations' full sizes.

The composition itself is pinned on the CPU: it equals the C oracle's own
LS(HS) pyramid bit for bit, its pre-filter weights equal the numpy fixture,
and a one-level run is the adapter alone."""
import numpy as np
import pytest

from oracle import oracle as orc
from oracle import pyramid as opyr

FLOW_TOL = 1e-3      # max |delta u| px against the oracle's reference dataflow
EPE_TOL = 1e-4       # mean end-point-error difference (to the synthetic truth), px


def _epe(u1, v1, u2, v2):
    return np.sqrt((u1 - u2) ** 2 + (v1 - v2) ** 2)


def _fb_consts():
    from Farneback_PyCL import Farneback_PyCL
    return Farneback_PyCL()._consts()


# ------------------------------------------------------------------ CPU pins
def test_composition_weights_are_prepare_gaussian_kernel(golden):
    g = golden("gauss_w")
    for key in g.files:
        sigma, n = key[1:].split("_n")
        assert np.array_equal(opyr._weights(float(sigma), int(n), "f32"), g[key]), key


@pytest.mark.parametrize("levels,size", [(2, 96), (3, 120)])
def test_composition_equals_the_c_pyramid(levels, size):
    """The Python composition of the driver loop and the C oracle's LS(HS)
    pyramid are two restatements of GenericPyramidalOpticalFlow.py:328-414:
    the same bits."""
    from ofri import synth
    im1, im2, _, _ = synth.config_pair(4, size=size)
    fw, ow = opyr._weights(3.4, 3, "f32"), opyr._weights(0.48, 5, "f32")
    alphas = [45.0] * (levels - 1) + [21.0]
    Uo, Vo, _ = orc.pyr_lshs(im1, im2, levels, alphas, 60, fw, ow, 5.0)
    Uc, Vc = opyr.generic_pyramid(im1, im2, 3.4, opyr.hs_adapter(alphas[::-1], 60), levels, 0.48,
                                  opyr.ls_adapter(5.0), warping=True, interScaling=True, finalScaling=True)
    assert np.array_equal(Uo, Uc) and np.array_equal(Vo, Vc)


def test_composition_one_level_is_the_adapter():
    from ofri import synth
    im1, im2, _, _ = synth.config_pair(1, size=96)
    c = _fb_consts()
    z = np.zeros_like(im1)
    U, V = opyr.generic_pyramid(im1, im2, 0, opyr.farneback_adapter(c), 1, warping=False)
    fx, fy = orc.farneback(im1, im2, z, z, c)
    assert np.array_equal(U, fx) and np.array_equal(V, fy)
    # no warping at level 2: the upsampled flow, scaled by 2, is the solver's start
    seen = []

    def spy(f1, f2, U0, V0, level):
        seen.append((level, U0.copy()))
        return U0 + np.float32(1.0), V0

    opyr.generic_pyramid(im1, im2, 0, spy, 2, warping=False, interScaling=True, finalScaling=True)
    assert seen[0][0] == 1 and not seen[0][1].any()
    ones = np.ones((48, 48), np.float32)
    assert np.array_equal(seen[1][1], orc.spline_upsample(ones, 96, 96, 2.0))
    assert np.allclose(seen[1][1], 2.0, atol=1e-6)
    seen.clear()
    opyr.generic_pyramid(im1, im2, 0, spy, 2, warping=False, interScaling=True, finalScaling=False)
    assert np.allclose(seen[1][1], 1.0, atol=1e-6)      # LK's defaults: no x2 at the last level


def test_composition_k_iterations_without_warping_continue_the_solver():
    """kLevels = 2 without warping hands the accumulated flow back to the
    solver as its start on the same pre-filtered pair
    (GenericPyramidalOpticalFlow.py:395-396, 223-227): HS 30 sweeps twice is
    HS 60 sweeps once, bit for bit (one level)."""
    from ofri import synth
    im1, im2, _, _ = synth.config_pair(4, size=64)
    U2, V2 = opyr.generic_pyramid(im1, im2, 3.4, opyr.hs_adapter([21.0, 21.0], 30), 1, warping=False,
                                  kLevels=2)
    U1, V1 = opyr.generic_pyramid(im1, im2, 3.4, opyr.hs_adapter([21.0], 60), 1, warping=False)
    assert np.array_equal(U1, U2) and np.array_equal(V1, V2)


def test_composition_k_iterations_with_warping_rewarp_the_level_frames():
    """kLevels = 2 with warping: the second iteration warps the level's pair
    by the whole accumulated flow and restarts the solver from zero
    (GenericPyramidalOpticalFlow.py:387-394).  A spy solver sees zero flow
    both times and, in the second, exactly the pre-filtered symmetric warp
    (bilinear) or Liu-Shen warp of the first frame (biLinear False) by the
    flow the first iteration returned (up and left: the Liu-Shen mode's
    targets then wrap instead of leaving the image)."""
    from ofri import synth
    im1, im2, _, _ = synth.config_pair(4, size=64)
    step = np.float32(0.75)
    for bil in (True, False):
        seen = []

        def spy(f1, f2, U0, V0, level):
            seen.append((f1.copy(), f2.copy(), U0.copy(), V0.copy()))
            return U0 - step, V0 - step

        Ua, Va = opyr.generic_pyramid(im1, im2, 3.4, spy, 1, warping=True, biLinear=bil, kLevels=2)
        assert len(seen) == 2 and not seen[0][2].any() and not seen[1][2].any() and not seen[1][3].any()
        acc_u, acc_v = np.full((64, 64), -step), np.full((64, 64), -step)
        k3 = opyr._weights(3.4, 3, "f32")
        if bil:
            w1, w2 = orc.warp_symmetric(im1, im2, acc_u, acc_v)
        else:
            w1, w2 = orc.ls_warp(im1, acc_u, acc_v), im2
        assert np.array_equal(seen[1][0], orc.gauss_sep(w1, k3)) and np.array_equal(seen[1][1], orc.gauss_sep(w2, k3))
        assert np.array_equal(Ua, acc_u - step) and np.array_equal(Va, acc_v - step)


# ------------------------------------------------------------------ GPU
@pytest.fixture(scope="module")
def ops():
    from ofri import ops as _ops
    from ofri.device import context
    context()
    return _ops


def _report(name, U, V, U32, V32, U64, V64, ut, vt, border=32):
    d32 = max(np.abs(U - U32).max(), np.abs(V - V32).max())
    d64 = max(np.abs(U - U64).max(), np.abs(V - V64).max())
    inner = (slice(border, -border), slice(border, -border))
    eg = _epe(U, V, ut, vt)[inner].mean()
    eo = _epe(U32, V32, ut, vt)[inner].mean()
    print("[%s] max|d| vs oracle-f32 %.3g px, vs oracle-f64 %.3g px; mean EPE to truth gpu %.4f oracle %.4f px"
          % (name, d32, d64, eg, eo))
    return d32, d64, abs(eg - eo)


def _fbls_dropin(im1, im2, route="device"):
    from Farneback_PyCL import Farneback_PyCL
    from GenericPyramidalOpticalFlow import genericPyramidalOpticalFlow
    from PhysicsBasedOpticalFlowLiuShen import LiuShenOpticalFlowAlgoAdapter

    class Loop(Farneback_PyCL):        # not the package's own class: the driver calls compute()
        pass

    fb = Farneback_PyCL(platformID=0) if route == "device" else Loop(platformID=0)
    return genericPyramidalOpticalFlow(im1, im2, 0, fb, 2, 1, 0.48, LiuShenOpticalFlowAlgoAdapter(10))


def _fbls_oracle(im1, im2, prec):
    return opyr.generic_pyramid(im1, im2, 0, opyr.farneback_adapter(_fb_consts(), prec), 2, 0.48,
                                opyr.ls_adapter(10.0, prec), warping=False, interScaling=True, finalScaling=True,
                                prec=prec)


@pytest.mark.gpu
@pytest.mark.parametrize("size", [256, 512])
def test_cfg1_farneback_liushen_pipeline_matches_oracle(ops, size):
    """LS(Farneback) 2-level as the example script runs it: FILTER 0,
    Farneback defaults (no warping, final scaling), Liu-Shen alpha 10 on the
    FILTER_OPT 0.48 pair.  The device-resident route and the adapter loop
    (numpy at every adapter call) give the same bits."""
    from ofri import synth
    im1, im2, ut, vt = synth.config_pair(1, size=size)
    U, V = _fbls_dropin(im1, im2)
    Ul, Vl = _fbls_dropin(im1, im2, route="loop")
    assert np.array_equal(U, Ul) and np.array_equal(V, Vl)
    U32, V32 = _fbls_oracle(im1, im2, "f32")
    U64, V64 = _fbls_oracle(im1, im2, "f64")
    d32, _, de = _report("LS(Farneback) %d^2" % size, U, V, U32, V32, U64, V64, ut, vt)
    assert d32 <= FLOW_TOL and de <= EPE_TOL


@pytest.mark.gpu
def test_cfg1_full_size_matches_oracle(ops):
    """The configuration itself: 1024^2 Lamb-Oseen, density 0.04."""
    from ofri import synth
    im1, im2, ut, vt = synth.config_pair(1)
    U, V = _fbls_dropin(im1, im2)
    U2, V2 = _fbls_dropin(im1, im2)
    assert np.array_equal(U, U2) and np.array_equal(V, V2) and np.isfinite(U).all() and np.isfinite(V).all()
    U32, V32 = _fbls_oracle(im1, im2, "f32")
    U64, V64 = _fbls_oracle(im1, im2, "f64")
    d32, _, de = _report("LS(Farneback) 1024^2", U, V, U32, V32, U64, V64, ut, vt, border=64)
    assert d32 <= FLOW_TOL and de <= EPE_TOL
    assert _epe(U, V, ut, vt)[64:-64, 64:-64].mean() < 0.5


def _lk_dropin(im1, im2, route="device"):
    from GenericPyramidalOpticalFlow import genericPyramidalOpticalFlow
    from denseLucasKanade_PyCL import denseLucasKanade_PyCl

    class Loop(denseLucasKanade_PyCl):
        pass

    cls = denseLucasKanade_PyCl if route == "device" else Loop
    lk = cls(Niter=5, halfWindow=13, platformID=0)
    U, V = genericPyramidalOpticalFlow(im1, im2, 2, lk, 2, 1, 0.48, None, warping=False)
    return U, V, lk.status


def _lk_oracle(im1, im2, prec, status=None):
    def main(f1, f2, U0, V0, level):
        u, v, st = orc.lk_dense(f1, f2, U0, V0, 13, 5, prec)
        if status is not None:
            status.append(st)
        return u, v
    return opyr.generic_pyramid(im1, im2, 2.0, main, 2, None, None, warping=False, interScaling=True,
                                finalScaling=False, prec=prec)


@pytest.mark.gpu
@pytest.mark.parametrize("size", [256, 512])
def test_cfg2_dense_lk_pipeline_matches_oracle(ops, size):
    """Dense LK 2-level as the example script runs it: FILTER 2, halfWindow
    13, Niter 5, warping False, and the adapter's defaults (intermediate
    scaling on, final scaling off: the last level's upsampled flow is not
    doubled).  Device route == adapter loop, status map included."""
    from ofri import synth
    im1, im2, ut, vt = synth.config_pair(2, size=size)
    U, V, st = _lk_dropin(im1, im2)
    Ul, Vl, stl = _lk_dropin(im1, im2, route="loop")
    assert np.array_equal(U, Ul) and np.array_equal(V, Vl) and np.array_equal(st, stl)
    sto = []
    U32, V32 = _lk_oracle(im1, im2, "f32", sto)
    U64, V64 = _lk_oracle(im1, im2, "f64")
    d32, _, de = _report("dense LK %d^2" % size, U, V, U32, V32, U64, V64, ut, vt)
    assert d32 <= FLOW_TOL and de <= EPE_TOL
    assert np.array_equal(st, sto[-1])
    print("  tracked %.4f of the pixels" % st.mean())


@pytest.mark.gpu
def test_cfg2_full_size_matches_oracle(ops):
    """The configuration itself: 2048^2 shear + turbulence, density 0.06
    (rendered on the device: the host renderer takes half a minute there)."""
    from ofri import synth
    im1, im2, ut, vt = (a.download() for a in synth.device_config_pair(2))
    U, V, st = _lk_dropin(im1, im2)
    U2, V2, _ = _lk_dropin(im1, im2)
    assert np.array_equal(U, U2) and np.array_equal(V, V2) and np.isfinite(U).all() and np.isfinite(V).all()
    U32, V32 = _lk_oracle(im1, im2, "f32")
    d32 = max(np.abs(U - U32).max(), np.abs(V - V32).max())
    inner = (slice(64, -64), slice(64, -64))
    eg, eo = _epe(U, V, ut, vt)[inner].mean(), _epe(U32, V32, ut, vt)[inner].mean()
    print("[dense LK 2048^2] max|d| vs oracle-f32 %.3g px; mean EPE to truth gpu %.4f oracle %.4f px; tracked %.4f"
          % (d32, eg, eo, st.mean()))
    assert d32 <= FLOW_TOL and abs(eg - eo) <= EPE_TOL
    assert eg < 1.0


@pytest.mark.gpu
def test_lshs_full_size_matches_oracle(ops):
    """The headline pipeline at its full size, 1024^2 LS(HS) 2-level, against
    the f32 oracle (and the f64 instance printed), with the Liu-Shen sweep
    counts equal."""
    from GenericPyramidalOpticalFlow import genericPyramidalOpticalFlow
    from HornSchunck import HSOpticalFlowAlgoAdapter
    from PhysicsBasedOpticalFlowLiuShen import LiuShenOpticalFlowAlgoAdapter
    from gaussian_filter import prepareGaussianKernel
    from ofri import synth
    im1, im2, ut, vt = synth.config_pair(4)
    U, V = genericPyramidalOpticalFlow(im1, im2, 3.4, HSOpticalFlowAlgoAdapter([21, 45], 600), 2, 1, 0.48,
                                       LiuShenOpticalFlowAlgoAdapter(5))
    fw, ow = prepareGaussianKernel(3.4, 3), prepareGaussianKernel(0.48, 5)
    _, _, st = ops.pyr_lshs(im1, im2, 2, [45.0, 21.0], 600, fw, ow, 5.0)
    U32, V32, s32 = orc.pyr_lshs(im1, im2, 2, [45.0, 21.0], 600, fw, ow, 5.0)
    U64, V64, _ = orc.pyr_lshs(im1, im2, 2, [45.0, 21.0], 600, fw, ow, 5.0, prec="f64")
    d32, _, de = _report("LS(HS) 1024^2", U, V, U32, V32, U64, V64, ut, vt, border=64)
    assert d32 <= FLOW_TOL and de <= EPE_TOL
    assert [s["ls_iters"] for s in st] == [int(x) for x in s32[:, 2]]


@pytest.mark.gpu
@pytest.mark.parametrize("warping,bil", [(True, True), (True, False), (False, True)])
def test_k_iterations_and_liu_shen_warping_match_oracle(ops, warping, bil):
    """genericPyramidalOpticalFlow with kLevels = 2 (re-warp or continue
    iterations) in each warping mode -- bilinear, the Liu-Shen mode (whose
    re-warp moves the already warped first frame again) and none -- against
    the oracle's composition of the same loop: the same bits."""
    from GenericPyramidalOpticalFlow import genericPyramidalOpticalFlow
    from HornSchunck import HSOpticalFlowAlgoAdapter
    from PhysicsBasedOpticalFlowLiuShen import LiuShenOpticalFlowAlgoAdapter
    from ofri import synth

    class NoDefaults(HSOpticalFlowAlgoAdapter):       # keep the caller's warping mode
        def hasGenericPyramidalDefaults(self):
            return False

    # a uniform shift up and left: the Liu-Shen mode's targets wrap rather
    # than leave the image
    rng = np.random.default_rng(61)
    im1, im2, _, _ = synth.piv_pair(96, 128, 0.05, lambda x, y: (-1.3 + 0 * x, -0.6 + 0 * y), rng)
    alphas = [21.0, 21.0, 45.0, 45.0]
    U, V = genericPyramidalOpticalFlow(im1.copy(), im2.copy(), 3.4, NoDefaults(list(alphas), 60), 2, 2, 0.48,
                                       LiuShenOpticalFlowAlgoAdapter(5), warping, bil, True, True)
    Uo, Vo = opyr.generic_pyramid(im1, im2, 3.4, opyr.hs_adapter(alphas, 60), 2, 0.48, opyr.ls_adapter(5.0),
                                  warping=warping, interScaling=True, finalScaling=True, biLinear=bil, kLevels=2)
    d = max(np.abs(U - Uo).max(), np.abs(V - Vo).max())
    print("[kLevels 2, warping %s, biLinear %s] max|d| vs oracle-f32 %.3g px" % (warping, bil, d))
    assert np.array_equal(U, Uo) and np.array_equal(V, Vo)

