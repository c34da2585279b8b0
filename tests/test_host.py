"""CPU-side checks of the host layer: the C ABI library loads and exports every
symbol include/ofri.h declares, the host-computed weights are bit-exact, and
the drop-in modules import and keep the reference's API surface.  No GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "opticalflow-ri_amd", "libofri.so")
HEADER = os.path.join(ROOT, "include", "ofri.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(ofri_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("ofri_create", "ofri_hs", "ofri_ls", "ofri_pyr_lshs", "ofri_gauss_sep", "ofri_resize",
              "ofri_spline_upsample", "ofri_warp_symmetric"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.skip("libofri.so not built")
    lib = ctypes.CDLL(LIB)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_capi_signatures_cover_header():
    from ofri import _capi
    syms = set(declared_symbols()) - {"ofri_last_error"}
    assert syms <= set(_capi.SIGNATURES), syms - set(_capi.SIGNATURES)


def test_library_reports_errors_without_gpu():
    """ofri_create on a machine without a GPU fails with a status, not a crash."""
    if not os.path.exists(LIB):
        pytest.skip("libofri.so not built")
    from ofri import _capi
    lib = _capi.load()
    if os.path.exists("/dev/kfd"):
        pytest.skip("GPU present")
    h = ctypes.c_void_p()
    st = lib.ofri_create(0, ctypes.byref(h))
    assert st != 0
    assert lib.ofri_last_error()


def test_prefilter_weights_bit_exact(golden):
    from gaussian_filter import prepareGaussianKernel
    g = golden("gauss_w")
    for key in g.files:
        s, n = key[1:].split("_n")
        w = prepareGaussianKernel(float(s), int(n))
        assert w.dtype == np.float32
        assert np.array_equal(w, g[key]), key


def test_gaussian_kernel_bitexact_known_answers(golden):
    from GaussianKernelBitExact import getGaussianKernelBitExact
    g = golden("bitexact")
    for n in (3, 5, 7, 9):
        s, k = getGaussianKernelBitExact(n, 0)
        assert np.array_equal(k, g["lit%d" % n]), n
        assert float(s) == 1.0
    for key in g.files:
        if not key.startswith("gen_"):
            continue
        n, sig = key[5:].split("_s")
        s, k = getGaussianKernelBitExact(int(n), float(sig))
        assert np.array_equal(k, g[key]), key
        assert abs(float(s) - 1.0) < 1e-15
    # sigma >= 0 is ignored for the general case (sigmaX = 0.15 n + 0.35)
    _, a = getGaussianKernelBitExact(33, 4.95)
    _, b = getGaussianKernelBitExact(33, 123.0)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("n", [11, 17, 33, 39, 159, 12])
def test_gaussian_kernel_bitexact_general_case_is_the_sampled_gaussian(n):
    """The general case against the mathematics it evaluates, in another
    arithmetic: normalised samples of exp(-k^2 / (2 sigma^2)) at the integer
    offsets k of the window (the reference's x = 2k runs 1-n .. n-1 in steps
    of 2 under exp(-0.125 x^2 / sigma^2)), sigma = 0.15 n + 0.35, formed in
    float64 with numpy: the float64 kernels agree to a few ulps, their
    float32 roundings (what every caller uses) bit for bit, and the window
    is symmetric, peaked at its centre and sums to 1."""
    from GaussianKernelBitExact import getGaussianKernelBitExact
    s, k = getGaussianKernelBitExact(n, 0.0 if n % 2 == 0 else 2.0)   # sigma >= 0 is ignored
    sigma = 0.15 * n + 0.35
    off = np.arange(n) - (n - 1) / 2.0
    if n % 2 == 0:
        # even windows: the reference's centre pair both take exp(0)
        off = np.where(np.abs(off) == 0.5, 0.0, off)
    g = np.exp(-off ** 2 / (2.0 * sigma * sigma))
    g = g / g.sum()
    assert np.abs(k - g).max() <= 8 * np.finfo(np.float64).eps * g.max()
    assert np.array_equal(k.astype(np.float32), g.astype(np.float32))
    assert np.array_equal(k, k[::-1]) and k.argmax() == (n - 1) // 2
    assert abs(float(s) - 1.0) < 1e-14 and abs(k.sum() - 1.0) < 1e-14


def test_dropin_api_surface():
    import GenericPyramidalOpticalFlow as G
    import HornSchunck as HSm
    import PhysicsBasedOpticalFlowLiuShen as LSm
    import gaussian_filter as gf
    for name in ("genericPyramidalOpticalFlow", "imresize", "doBiLinearWarping", "updateNextPyramidalLevel"):
        assert callable(getattr(G, name))
    hs = HSm.HSOpticalFlowAlgoAdapter([21, 45], 600)
    assert hs.getAlgoName() == "Horn-Schunck"
    assert hs.hasGenericPyramidalDefaults() is True
    assert hs.getGenericPyramidalDefaults() == {"warping": True, "biLinear": True, "scaling": True}
    assert HSm.HSOpticalFlowAlgoAdapter([1], 1, False).hasGenericPyramidalDefaults() is False
    ls = LSm.LiuShenOpticalFlowAlgoAdapter(5)
    assert ls.getAlgoName() == "Liu-Shen Physics based OF"
    with pytest.raises(TypeError):
        ls.hasGenericPyramidalDefaults()
    for name in ("prepareGaussianKernel", "convolveSeparableFilter", "gaussian_filter", "gaussian_filterPx"):
        assert callable(getattr(gf, name))


def test_level_shapes_follow_reference_rounding():
    import GenericPyramidalOpticalFlow as G
    assert G._level_shapes(1024, 1024, 2) == [(512, 512), (1024, 1024)]
    assert G._level_shapes(385, 257, 2) == [(192, 128), (385, 257)]   # round half to even
    assert G._level_shapes(100, 60, 3) == [(25, 15), (50, 30), (100, 60)]


def test_zero_levels_computes_nothing():
    """pyramidalLevels = 0: the reference's level loop is empty and it
    returns (None, None) without touching the adapter."""
    import GenericPyramidalOpticalFlow as G
    import HornSchunck as HSm
    hs = HSm.HSOpticalFlowAlgoAdapter([21], 10)
    out = G.genericPyramidalOpticalFlow(np.zeros((8, 8), np.float32), np.zeros((8, 8), np.float32), 3.4, hs, 0)
    assert out == (None, None) and hs.alphas == [21]


def test_synthetic_generator_is_deterministic():
    from ofri import synth
    a = synth.config_pair(0, size=64)
    b = synth.config_pair(0, size=64)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    im1, im2, u, v = a
    assert im1.dtype == np.float32 and im1.shape == (64, 64)
    assert np.abs(u).max() <= 3.5 + 1e-6 and np.all(v == 0)


def test_autotune_size_classes_from_cached_results():
    """The autotuner's cached per-size results become size classes with
    1-based variants, each class starting at the geometric mean of its
    size's pixel count and the next smaller tuned size's (the smallest at 0)."""
    from ofri import autotune
    rows = autotune.classes({"1024": {"hs": [10, 10, 0.3], "ls": [2, 4, 0.4]}, "256": {"hs": [1, 3, 0.01]},
                             "512": {"hs": [4, 5, 0.1], "ls": [0, 5, 0.2]}})
    assert rows == [(0, 2, 3, 0, 0), (131072, 5, 5, 1, 5), (524288, 11, 10, 3, 4)]
    assert autotune.apply(object(), path="") == []

