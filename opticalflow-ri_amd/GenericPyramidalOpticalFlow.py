"""Drop-in for the reference's GenericPyramidalOpticalFlow.py: the generic
coarse-to-fine driver, with every numeric step on the GPU.

Reference: GenericPyramidalOpticalFlow.py:66-415.  The public functions keep
their signatures and return types (numpy float32 arrays).

When the main adapter is this package's HSOpticalFlowAlgoAdapter and the
optional one is LiuShenOpticalFlowAlgoAdapter (or absent), with bilinear
warping, velocity scaling and kLevels = 1 -- the configuration of
examples/LiuSE_PyHSchunck_Fs3_4_PyrLvls2.py -- the whole pyramid runs as one
device-resident libofri call (ofri_pyr_lshs): the adapter's alpha list is
consumed exactly as the reference consumes it and the same numbers come out.
This package's Farneback (one internal level) and dense-LK (no vorticity
enhancement) adapters, with the Liu-Shen or no optional adapter, bilinear or
no warping and kLevels = 1 -- the configurations of
examples/LiuSE_Farneback_Fs0_0_PyrLvls2.py and examples/denseLK_Fs2_0_PyrLvls2.py
-- run as one device-resident call too (ofri_pyr_generic).
Every other combination (any adapter object following the reference's
adapter protocol, kLevels > 1, Liu-Shen warping) goes through the
level loop below,
which calls adapter.compute() with numpy arrays as the reference does and
runs resize / upsample / warp / pre-filter on the GPU.
"""
import numpy as np

from gaussian_filter import gaussian_filter, gaussian_filterPx, prepareGaussianKernel  # noqa: F401
from ofri import ops
from ofri.device import as_device

BICUBIC = ops.BICUBIC


def imresize(im, res):
    """Pillow BICUBIC resize of a float image to res = (width, height)
    (GenericPyramidalOpticalFlow.py:66-67)."""
    W2, H2 = int(res[0]), int(res[1])
    return ops.resize(np.ascontiguousarray(im, dtype=np.float32), H2, W2, BICUBIC).download()


def doBiLinearWarping(img, coordsY, coordsX, order=1, mode='nearest'):
    """Bilinear sample of img at (coordsY, coordsX) with round-half-even
    integer parts and clamped neighbours (GenericPyramidalOpticalFlow.py:69-115).
    Coordinates are taken as float32, as every reference caller passes them."""
    img = np.ascontiguousarray(img, dtype=np.float32)
    cy = np.ascontiguousarray(coordsY, dtype=np.float32)
    cx = np.ascontiguousarray(coordsX, dtype=np.float32)
    return ops.warp_coords(img, cy, cx).download()


def _upsample_flow(Uaccum, Vaccum, shape_prev, shape, scale):
    """Spline upsample to `shape` (or pass-through) and optional scaling
    (GenericPyramidalOpticalFlow.py:151-171).  Device arrays in and out."""
    (hp, wp), (h, w) = shape_prev, shape
    sx = np.float32(w) / np.float32(wp) if scale else np.float32(1.0)
    sy = np.float32(h) / np.float32(hp) if scale else np.float32(1.0)
    if scale:
        ops.log('Scaling  ')
    if (hp, wp) != (h, w):
        return ops.spline_upsample(Uaccum, h, w, sx), ops.spline_upsample(Vaccum, h, w, sy)
    # same size: no resampling, and the scaling factors are exactly 1
    return as_device(Uaccum).copy(), as_device(Vaccum).copy()


def updateNextPyramidalLevel(im1IterNext, im1IterPrev, im2IterNext, Uaccum, Vaccum, U, V, warping=True,
                             biLinear=True, scale=False):
    """GenericPyramidalOpticalFlow.py:117-234.  Returns (im1 warped, im2 warped,
    final Uaccum, final Vaccum, initial U, initial V) as numpy float32.  The
    Liu-Shen warping mode (biLinear False) warps im1IterNext in place, as the
    reference does, and returns it; it leaves the second frame as it is."""
    shape = np.shape(im1IterNext)
    us, vs = _upsample_flow(Uaccum, Vaccum, np.shape(im1IterPrev), shape, scale)
    if warping:
        z = np.zeros(shape, dtype='float32')
        if not biLinear:
            ops.log('Warping: Liu-Shen')
            w1 = ops.warp_liushen(im1IterNext, us, vs).download()
            if isinstance(im1IterNext, np.ndarray) and im1IterNext.flags.writeable:
                im1IterNext[...] = w1
                w1 = im1IterNext
            return w1, np.asarray(im2IterNext), us.download(), vs.download(), z, z.copy()
        ops.log('Warping: BiLinear')
        w1, w2 = ops.warp_symmetric(im1IterNext, im2IterNext, us, vs)
        return w1.download(), w2.download(), us.download(), vs.download(), z, z.copy()
    z = np.zeros(shape, dtype='float32')
    return (np.asarray(im1IterNext), np.asarray(im2IterNext), z, z.copy(), us.download(), vs.download())


def _native_lshs(main, optional):
    from HornSchunck import HSOpticalFlowAlgoAdapter
    from PhysicsBasedOpticalFlowLiuShen import LiuShenOpticalFlowAlgoAdapter
    return (type(main) is HSOpticalFlowAlgoAdapter and
            (optional is None or type(optional) is LiuShenOpticalFlowAlgoAdapter))


def _level_shapes(H, W, levels):
    scale = 1.0 / (2.0 ** (levels - 1))
    out = []
    for level in range(1, levels + 1):
        if scale < 1.0 and level != levels:
            out.append((int(np.int32(np.round(H * scale))), int(np.int32(np.round(W * scale)))))
        else:
            out.append((H, W))
        scale *= 2
    return out


def genericPyramidalOpticalFlow(im1, im2, FILTER, mainOFlowAlgoAdapter, pyramidalLevels=1, kLevels=1,
                                FILTER_OPT=None, optionalOFlowAlgoAdapter=None, warping=True, biLinear=True,
                                pyramidalIntermediateScaling=True, pyramidalScaling=False):
    """Coarse-to-fine optical flow with a main and an optional adapter
    (GenericPyramidalOpticalFlow.py:237-415).  Returns (Uaccum, Vaccum)."""
    main, optional = mainOFlowAlgoAdapter, optionalOFlowAlgoAdapter

    if main.hasGenericPyramidalDefaults():
        defaults = main.getGenericPyramidalDefaults()
        if defaults is not None:
            name = main.getAlgoName()
            for key, label in (('warping', 'warping'), ('biLinear', 'biLinear'),
                               ('intermediateScaling', 'intermediatePyramidalScaling'),
                               ('scaling', 'pyramidalScaling')):
                val = defaults.get(key)
                if val is None:
                    continue
                ops.log('Using algorithm ' + name + ' default value for ' + label + ' parameter:', val)
                if key == 'warping':
                    warping = val
                elif key == 'biLinear':
                    biLinear = val
                elif key == 'intermediateScaling':
                    pyramidalIntermediateScaling = val
                else:
                    pyramidalScaling = val

    if pyramidalLevels < 1:
        # the reference's level loop is empty: nothing is computed
        # (its 'Invalid scale level' check sits inside that loop)
        return None, None
    im1 = np.ascontiguousarray(im1, dtype=np.float32)
    im2 = np.ascontiguousarray(im2, dtype=np.float32)
    H, W = im1.shape
    shapes = _level_shapes(H, W, pyramidalLevels)

    if (_native_lshs(main, optional) and kLevels == 1 and warping and biLinear and
            (pyramidalLevels == 1 or (pyramidalIntermediateScaling and pyramidalScaling)) and
            len(main.alphas) >= pyramidalLevels and min(min(s) for s in shapes) >= 4):
        return _lshs_device(im1, im2, FILTER, main, pyramidalLevels, FILTER_OPT, optional)
    dev = _device_main(main, optional)
    if dev is not None and kLevels == 1 and (biLinear or not warping) and min(min(s) for s in shapes) >= 4:
        return _generic_device(im1, im2, FILTER, main, dev, pyramidalLevels, FILTER_OPT, optional, warping,
                               pyramidalIntermediateScaling, pyramidalScaling)
    return _generic_levels(im1, im2, FILTER, main, pyramidalLevels, kLevels, FILTER_OPT, optional, warping,
                           biLinear, pyramidalIntermediateScaling, pyramidalScaling, shapes)


def _lshs_device(im1, im2, FILTER, hs, levels, FILTER_OPT, ls):
    alphas = [hs.alphas.pop() for _ in range(levels)]
    filt = prepareGaussianKernel(FILTER, 3) if FILTER > 1e-3 else None
    opt = None
    if ls is not None and FILTER_OPT > 1e-3:
        opt = prepareGaussianKernel(FILTER_OPT, 5)
    Ud, Vd, stats = ops.pyr_lshs(im1, im2, levels, alphas, hs.Niter, filt, opt,
                                 ls.alpha if ls is not None else 0.0)
    for level, st in enumerate(stats, start=1):
        ops.log('Level=', level, ' kIter=', 0)
        ops.log(hs.getAlgoName() + ' estimated error for image registration: ' + str(st['hs_err']))
        if ls is not None:
            ops.log(ls.getAlgoName() + ' estimated error for image registration: ' + str(st['ls_err']))
    return Ud.download(), Vd.download()


def _device_main(main, optional):
    """('fb', constants) for this package's Farneback adapter with one internal
    level, ('lk', (halfWindow, Niter)) for its dense-LK adapter without
    vorticity enhancement, each with the Liu-Shen (or no) optional adapter;
    None otherwise (the level loop then calls the adapters)."""
    from PhysicsBasedOpticalFlowLiuShen import LiuShenOpticalFlowAlgoAdapter
    if optional is not None and type(optional) is not LiuShenOpticalFlowAlgoAdapter:
        return None
    from Farneback_PyCL import Farneback_PyCL
    from denseLucasKanade_PyCL import denseLucasKanade_PyCl
    if type(main) is Farneback_PyCL and main.pyramidalLevels == 0:
        # the reference's compute() checks (Farneback_PyCL.py:462-465)
        assert main.polyN == 5 or main.polyN == 7
        assert main.pyrScale < 1
        return ('fb', main._consts())
    if type(main) is denseLucasKanade_PyCl and not main.enableVorticityEnhancement:
        return ('lk', (main.windowHalfWidth, main.Niter))
    return None


def _generic_device(im1, im2, FILTER, main, dev, levels, FILTER_OPT, optional, warping, interScaling,
                    finalScaling):
    """The level loop below as one device-resident libofri call
    (ofri_pyr_generic): every adapter call runs on device buffers."""
    filt = prepareGaussianKernel(FILTER, 3) if FILTER > 1e-3 else None
    opt = None
    if optional is not None and FILTER_OPT is not None and FILTER_OPT > 1e-3:
        opt = prepareGaussianKernel(FILTER_OPT, 5)
    kind, arg = dev
    status = ops.ByteArray(im1.shape) if kind == 'lk' else None
    Ud, Vd, stats = ops.pyr_generic(im1, im2, levels, filt, opt, warping, interScaling, finalScaling,
                                    ls_alpha=optional.alpha if optional is not None else 0.0,
                                    fb_consts=arg if kind == 'fb' else None, lk=arg if kind == 'lk' else None,
                                    status=status)
    err = 'Unknown' if kind == 'fb' else True
    for level, st in enumerate(stats, start=1):
        ops.log('Level=', level, ' kIter=', 0)
        ops.log(main.getAlgoName() + ' estimated error for image registration: ' + str(err))
        if optional is not None:
            ops.log(optional.getAlgoName() + ' estimated error for image registration: ' + str(st['ls_err']))
    if status is not None:
        main.status = status.download()
    return Ud.download(), Vd.download()


def genericPyramidalOpticalFlowBatch(pairs, FILTER, mainOFlowAlgoAdapter, pyramidalLevels=1, kLevels=1,
                                     FILTER_OPT=None, optionalOFlowAlgoAdapter=None, warping=True, biLinear=True,
                                     pyramidalIntermediateScaling=True, pyramidalScaling=False):
    """genericPyramidalOpticalFlow for many (im1, im2) pairs of any sizes with one
    set of parameters.  Returns a list of (Uaccum, Vaccum), each equal to what
    genericPyramidalOpticalFlow returns for that pair alone.

    With the Horn-Schunck main adapter and the Liu-Shen (or no) optional
    adapter on the adapters' default options, the whole batch is one device
    call (ofri_batch_lshs) and the main adapter's alphas are popped once for
    the batch, as one driver call pops them.  Any other combination runs the
    driver per pair on a copy of each adapter."""
    import copy
    main, optional = mainOFlowAlgoAdapter, optionalOFlowAlgoAdapter
    pairs = [(np.ascontiguousarray(a, dtype=np.float32), np.ascontiguousarray(b, dtype=np.float32))
             for a, b in pairs]
    for a, b in pairs:
        if a.shape != b.shape:
            raise ValueError("images differ in shape: %s vs %s" % (a.shape, b.shape))
    # the HS adapter's defaults switch warping, biLinear and final scaling on
    native = (_native_lshs(main, optional) and main.hasGenericPyramidalDefaults() and kLevels == 1 and
              pyramidalLevels >= 1 and (pyramidalLevels == 1 or pyramidalIntermediateScaling) and
              len(main.alphas) >= pyramidalLevels and
              all(min(min(s) for s in _level_shapes(a.shape[0], a.shape[1], pyramidalLevels)) >= 4
                  for a, _ in pairs))
    if not native:
        return [genericPyramidalOpticalFlow(a, b, FILTER, copy.deepcopy(main), pyramidalLevels, kLevels,
                                            FILTER_OPT, copy.deepcopy(optional), warping, biLinear,
                                            pyramidalIntermediateScaling, pyramidalScaling) for a, b in pairs]
    alphas = [main.alphas.pop() for _ in range(pyramidalLevels)]
    filt = prepareGaussianKernel(FILTER, 3) if FILTER > 1e-3 else None
    opt = None
    if optional is not None and FILTER_OPT is not None and FILTER_OPT > 1e-3:
        opt = prepareGaussianKernel(FILTER_OPT, 5)
    out, _ = ops.batch_lshs(pairs, pyramidalLevels, alphas, main.Niter, filt, opt,
                            optional.alpha if optional is not None else 0.0, stats=False)
    return [(U.download(), V.download()) for U, V in out]


def _generic_levels(im1, im2, FILTER, main, levels, kLevels, FILTER_OPT, optional, warping, biLinear,
                    interScaling, finalScaling, shapes):
    Uaccum = Vaccum = U = V = None
    work1 = None
    d1, d2 = as_device(im1), as_device(im2)
    for level in np.arange(1, levels + 1):
        localScaling = finalScaling if level == levels else interScaling
        h, w = shapes[level - 1]
        prev_shape = work1.shape if work1 is not None else None
        if (h, w) != im1.shape:
            new1, new2 = ops.resize(d1, h, w), ops.resize(d2, h, w)
        else:
            new1, new2 = d1, d2
        if level > 1:
            n1 = new1.download()
            w1, w2, Uaccum, Vaccum, U, V = updateNextPyramidalLevel(
                n1, np.empty(prev_shape, np.float32), new2.download(), Uaccum, Vaccum, U, V,
                warping, biLinear, localScaling)
            warp1, warp2 = as_device(w1), as_device(w2)
            if warping and not biLinear:
                # the reference's Liu-Shen warp moves im1IterNew's pixels in
                # place: the optional pre-filter and the re-warp iterations
                # below see the warped frame (GenericPyramidalOpticalFlow.py:206,353,381,393)
                new1 = as_device(n1)
        else:
            warp1, warp2 = new1, new2
            U = np.zeros((h, w), dtype='float32')
            V = np.zeros((h, w), dtype='float32')
            Uaccum = np.zeros((h, w), dtype='float32')
            Vaccum = np.zeros((h, w), dtype='float32')
        if FILTER > 1e-3:
            k3 = prepareGaussianKernel(FILTER, 3)
            work1, work2 = ops.gauss_sep(warp1, k3), ops.gauss_sep(warp2, k3)
        else:
            work1, work2 = warp1.copy(), warp2
        opt1 = opt2 = None
        if optional is not None and FILTER_OPT > 1e-3:
            k5 = prepareGaussianKernel(FILTER_OPT, 5)
            opt1, opt2 = ops.gauss_sep(new1, k5), ops.gauss_sep(new2, k5)
        elif optional is not None:
            opt1, opt2 = new1.copy(), new2
        for k in np.arange(0, kLevels):
            ops.log('Level=', level, ' kIter=', k)
            if k > 0:
                if warping:
                    w1, w2, Uaccum, Vaccum, U, V = updateNextPyramidalLevel(
                        new1.download(), np.empty((h, w), np.float32), new2.download(), Uaccum, Vaccum, U, V,
                        warping, biLinear, False)
                    warp1, warp2 = as_device(w1), as_device(w2)
                    if FILTER > 1:
                        k3 = prepareGaussianKernel(FILTER, 3)
                        work1, work2 = ops.gauss_sep(warp1, k3), ops.gauss_sep(warp2, k3)
                    else:
                        work1, work2 = warp1.copy(), warp2
                else:
                    w1, w2, Uaccum, Vaccum, U, V = updateNextPyramidalLevel(
                        work1.download(), np.empty((h, w), np.float32), work2.download(), Uaccum, Vaccum, U, V,
                        warping, biLinear, False)
                    work1, work2 = as_device(w1), as_device(w2)
            U, V, error = main.compute(work1.download(), work2.download(), U, V)
            ops.log(main.getAlgoName() + ' estimated error for image registration: ' + str(error))
            if optional is not None:
                U, V, errorOptional = optional.compute(opt1.download(), opt2.download(), U, V)
                ops.log(optional.getAlgoName() + ' estimated error for image registration: ' +
                        str(errorOptional))
            Uaccum = ops.accumulate(as_device(Uaccum), as_device(np.asarray(U, dtype=np.float32))).download()
            Vaccum = ops.accumulate(as_device(Vaccum), as_device(np.asarray(V, dtype=np.float32))).download()
    return Uaccum, Vaccum
