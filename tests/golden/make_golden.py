"""Writes the library-probed golden vectors in tests/golden/*.npz.

The reference ships no expected outputs, so every fixture here is produced by
the third-party libraries the reference calls, on small seeded inputs, with
the call pattern the reference uses:

  stencils.npz   scipy.ndimage.convolve with the reference's kernels and modes
                 (HornSchunck.py:86-88,109-124; PhysicsBasedOpticalFlowLiuShen.py:
                 48-65,110-128,142-148)
  hs_small.npz   a few Horn-Schunck sweeps driven by scipy convolve + numpy
                 float32 element-wise updates (HornSchunck.py:51-104)
  ls_small.npz   a few Liu-Shen sweeps, same construction
                 (PhysicsBasedOpticalFlowLiuShen.py:82-158)
  gauss_w.npz    float32 pre-filter weights formed by numpy
                 (gaussian_filter.py:47-52) for the sigma / tap counts used
  pil.npz        Pillow float-image resizes, BICUBIC and BILINEAR
                 (GenericPyramidalOpticalFlow.py:66-67, Farneback_PyCL.py:62)
  spline.npz     scipy RectBivariateSpline on arange(n)/n grids
                 (GenericPyramidalOpticalFlow.py:151-161)
  warp.npz       numpy round-half-even bilinear warp incl. exact .5 ties
                 (GenericPyramidalOpticalFlow.py:69-115,199-200)
  bitexact.npz   getGaussianKernelBitExact known answers: the sigma <= 0
                 literal kernels (GaussianKernelBitExact.py:58-96) and the
                 general case restated with the stdlib decimal module
                 (GaussianKernelBitExact.py:98-144)
  piv_pair.npz   the reference's only data fixture, the 8-bit 512x512 PIV
                 pair examples/testImages/Bits08/Ni06/parabolic01_{0,1}.tif,
                 stored as uint8 (read from the reference checkout only when
                 this script runs).

Run:  python tests/golden/make_golden.py   (needs scipy and Pillow; not run by
the test-suite, whose tests only read the .npz files).
"""
import decimal
import os

import numpy as np
from PIL import Image
from scipy.interpolate import RectBivariateSpline
from scipy.ndimage import convolve

HERE = os.path.dirname(os.path.abspath(__file__))
REF_IMAGES = "/root/reference/examples/testImages/Bits08/Ni06"


def flip(k):
    return np.flipud(np.fliplr(k))


def stencils(rng):
    out = {}
    X = (rng.random((9, 11)) * 255).astype(np.float32)
    Y = (rng.standard_normal((9, 11))).astype(np.float32)
    out["X"], out["Y"] = X, Y
    kx = np.array([[-1, 1], [-1, 1]], np.float32) * .25
    ky = np.array([[-1, -1], [1, 1]], np.float32) * .25
    kt = np.ones((2, 2), np.float32) * .25
    out["hs_cx"] = convolve(X, kx, mode="mirror")
    out["hs_cy"] = convolve(X, ky, mode="mirror")
    out["hs_ct"] = convolve(X, kt, mode="mirror")
    K = np.array([[1 / 12, 1 / 6, 1 / 12], [1 / 6, 0, 1 / 6], [1 / 12, 1 / 6, 1 / 12]], np.float32)
    out["hs_avg"] = convolve(Y, K, mode="mirror")
    D = flip(np.array([[0, -1, 0], [0, 0, 0], [0, 1, 0]], np.float32) / 2)
    M = flip(np.array([[1, 0, -1], [0, 0, 0], [-1, 0, 1]], np.float32) / 4)
    F = flip(np.array([[0, 1, 0], [0, 0, 0], [0, 1, 0]], np.float32))
    D2 = flip(np.array([[0, 1, 0], [0, -2, 0], [0, 1, 0]], np.float32))
    Hk = flip(np.array([[1, 1, 1], [1, 0, 1], [1, 1, 1]], np.float32))
    out["ls_dr"] = convolve(Y, D, mode="nearest")
    out["ls_dc"] = convolve(Y, D.T, mode="nearest")
    out["ls_m"] = convolve(Y, M, mode="nearest")
    out["ls_fr"] = convolve(Y, F, mode="nearest")
    out["ls_fc"] = convolve(Y, F.T, mode="nearest")
    out["ls_d2r"] = convolve(Y, D2, mode="nearest")
    out["ls_d2c"] = convolve(Y, D2.T, mode="nearest")
    out["ls_h0"] = convolve(Y, Hk, mode="constant")
    out["ls_cmtx"] = convolve(np.ones(Y.shape, np.float32), Hk, mode="constant")
    np.savez_compressed(os.path.join(HERE, "stencils.npz"), **out)


def hs_small(rng):
    """Horn-Schunck on a 14x12 pair, 7 sweeps from a non-zero initial flow."""
    I0 = (rng.random((14, 12)) * 255).astype(np.float32)
    I1 = (rng.random((14, 12)) * 255).astype(np.float32)
    U0 = rng.standard_normal((14, 12)).astype(np.float32)
    V0 = rng.standard_normal((14, 12)).astype(np.float32)
    alpha, niter = np.float32(21.0), 7
    kx = np.array([[-1, 1], [-1, 1]], np.float32) * .25
    ky = np.array([[-1, -1], [1, 1]], np.float32) * .25
    kt = np.ones((2, 2), np.float32) * .25
    # HS(im2=I0, im1=I1): computeDerivatives(I1, I0)
    fx = convolve(I1, kx, mode="mirror") + convolve(I0, kx, mode="mirror")
    fy = convolve(I1, ky, mode="mirror") + convolve(I0, ky, mode="mirror")
    ft = convolve(I0, kt, mode="mirror") + convolve(I1, -kt, mode="mirror")
    K = np.array([[1 / 12, 1 / 6, 1 / 12], [1 / 6, 0, 1 / 6], [1 / 12, 1 / 6, 1 / 12]], np.float32)
    U, V = U0.copy(), V0.copy()
    for _ in range(niter):
        ub = convolve(U, K, mode="mirror")
        vb = convolve(V, K, mode="mirror")
        der = (fx * ub + fy * vb + ft) / (alpha * alpha + fx * fx + fy * fy)
        U = ub - fx * der
        V = vb - fy * der
    np.savez_compressed(os.path.join(HERE, "hs_small.npz"), I0=I0, I1=I1, U0=U0, V0=V0,
                        alpha=alpha, niter=niter, fx=fx, fy=fy, ft=ft, U=U, V=V)


def ls_small(rng):
    """Liu-Shen on a 13x15 pair (solver orientation), 6 sweeps."""
    im1 = (rng.random((13, 15)) * 200 + 20).astype(np.float32)
    im2 = (rng.random((13, 15)) * 200 + 20).astype(np.float32)
    u0 = (rng.standard_normal((13, 15)) * .5).astype(np.float32)
    v0 = (rng.standard_normal((13, 15)) * .5).astype(np.float32)
    h, nsweep = 5, 6
    I = im1 / np.max(im1)
    I2 = im2 / np.max(im2)
    D = flip(np.array([[0, -1, 0], [0, 0, 0], [0, 1, 0]], np.float32) / 2)
    M = flip(np.array([[1, 0, -1], [0, 0, 0], [-1, 0, 1]], np.float32) / 4)
    F = flip(np.array([[0, 1, 0], [0, 0, 0], [0, 1, 0]], np.float32))
    D2 = flip(np.array([[0, 1, 0], [0, -2, 0], [0, 1, 0]], np.float32))
    Hk = flip(np.array([[1, 1, 1], [1, 0, 1], [1, 1, 1]], np.float32))
    IIx = I * convolve(I, D, mode="nearest")
    IIy = I * convolve(I, D.T, mode="nearest")
    II = I * I
    Ixt = I * convolve(I2 - I, D, mode="nearest")
    Iyt = I * convolve(I2 - I, D.T, mode="nearest")
    hf = np.float32(h)
    cm = convolve(np.ones(I.shape, np.float32), Hk, mode="constant")
    A11 = I * (convolve(I, D2, mode="nearest") - 2 * I) - hf * cm
    A22 = I * (convolve(I, D2.T, mode="nearest") - 2 * I) - hf * cm
    A12 = I * convolve(I, M, mode="nearest")
    det = A11 * A22 - A12 * A12
    B11, B12, B22 = A22 / det, -A12 / det, A11 / det
    u, v = u0.copy(), v0.copy()
    errs = []
    for _ in range(nsweep):
        bu = 2 * IIx * convolve(u, D, mode="nearest") + IIx * convolve(v, D.T, mode="nearest") + \
            IIy * convolve(v, D, mode="nearest") + II * convolve(u, F, mode="nearest") + \
            II * convolve(v, M, mode="nearest") + h * convolve(u, Hk, mode="constant") + Ixt
        bv = IIy * convolve(u, D, mode="nearest") + IIx * convolve(u, D.T, mode="nearest") + \
            2 * IIy * convolve(v, D.T, mode="nearest") + II * convolve(u, M, mode="nearest") + \
            II * convolve(v, F.T, mode="nearest") + h * convolve(v, Hk, mode="constant") + Iyt
        un = -(B11 * bu + B12 * bv)
        vn = -(B12 * bu + B22 * bv)
        errs.append((np.linalg.norm((un - u).astype(np.float64)) +
                     np.linalg.norm((vn - v).astype(np.float64))) / u.size)
        u, v = un, vn
    np.savez_compressed(os.path.join(HERE, "ls_small.npz"), im1=im1, im2=im2, u0=u0, v0=v0, h=h,
                        nsweep=nsweep, IIx=IIx, IIy=IIy, II=II, Ixt=Ixt, Iyt=Iyt, B11=B11, B12=B12,
                        B22=B22, u=u, v=v, errs=np.array(errs))


def gauss_w():
    out = {}
    for sigma, n in ((3.4, 3), (0.48, 5), (2.0, 3), (1.0, 5), (2.5, 7), (0.8, 3)):
        k = np.zeros(n, np.float32)
        xs = np.arange(-n / 2, n / 2, 1, dtype=int)
        k[:] = 1.0 / np.sqrt(2.0 * np.pi * sigma ** 2) * np.exp(-xs ** 2 / (2.0 * sigma ** 2))
        k /= np.sum(k)
        out["s%g_n%d" % (sigma, n)] = k
    np.savez_compressed(os.path.join(HERE, "gauss_w.npz"), **out)


def pil(rng):
    out = {}
    cases = [(16, 16, 8, 8), (37, 29, 18, 14), (24, 40, 12, 20), (10, 7, 5, 4), (64, 48, 32, 24),
             (33, 65, 16, 32), (9, 9, 9, 9), (20, 12, 40, 24)]
    for t, (H, W, H2, W2) in enumerate(cases):
        im = (rng.random((H, W)) * 255).astype(np.float32)
        out["in%d" % t] = im
        out["bicubic%d" % t] = np.array(Image.fromarray(im).resize((W2, H2), Image.BICUBIC))
        out["bilinear%d" % t] = np.array(Image.fromarray(im).resize((W2, H2), Image.BILINEAR))
        # impulse responses of one column pass
    imp = np.zeros((1, 21), np.float32)
    imp[0, 10] = 1.0
    out["impulse"] = imp
    out["impulse_bicubic"] = np.array(Image.fromarray(imp).resize((10, 1), Image.BICUBIC))
    out["impulse_bilinear"] = np.array(Image.fromarray(imp).resize((10, 1), Image.BILINEAR))
    out["cases"] = np.array(cases, np.int32)
    np.savez_compressed(os.path.join(HERE, "pil.npz"), **out)


def spline(rng):
    out = {}
    cases = [(6, 8, 12, 16), (9, 7, 18, 14), (16, 16, 32, 32), (5, 11, 10, 21), (12, 10, 25, 19)]
    for t, (hc, wc, h, w) in enumerate(cases):
        C = rng.standard_normal((hc, wc)).astype(np.float32)
        yl = np.arange(hc) / np.float32(hc)
        xl = np.arange(wc) / np.float32(wc)
        y = np.arange(h) / np.float32(h)
        x = np.arange(w) / np.float32(w)
        out["in%d" % t] = C
        out["out%d" % t] = RectBivariateSpline(yl, xl, C)(y, x)
    out["cases"] = np.array(cases, np.int32)
    np.savez_compressed(os.path.join(HERE, "spline.npz"), **out)


def bilinear_warp(img, cy, cx):
    iy = np.int32(np.round(cy)); ix = np.int32(np.round(cx))
    dy = cy - iy; dx = cx - ix
    iyn = iy + 1; ixn = ix + 1
    r = np.where(dy < 0); iyn[r] = iy[r] - 1
    r = np.where(dx < 0); ixn[r] = ix[r] - 1
    dy = np.abs(dy); dx = np.abs(dx)
    H, W = img.shape
    iy = np.clip(iy, 0, H - 1); iyn = np.clip(iyn, 0, H - 1)
    ix = np.clip(ix, 0, W - 1); ixn = np.clip(ixn, 0, W - 1)
    out = np.zeros(img.shape, np.float32)
    out[:, :] = (1 - dy) * (1 - dx) * img[iy, ix] + (1 - dy) * dx * img[iy, ixn] + \
        dy * (1 - dx) * img[iyn, ix] + dy * dx * img[iyn, ixn]
    return out


def warp(rng):
    H, W = 17, 23
    im1 = (rng.random((H, W)) * 255).astype(np.float32)
    im2 = (rng.random((H, W)) * 255).astype(np.float32)
    U = (rng.standard_normal((H, W)) * 3).astype(np.float32)
    V = (rng.standard_normal((H, W)) * 3).astype(np.float32)
    # exact half-pixel ties (U/2 = k + .5) and shifts past the border
    U[0, :5] = [1.0, 3.0, -1.0, -3.0, 5.0]
    V[1, :5] = [1.0, -1.0, 7.0, -7.0, 0.0]
    U[-1, -3:] = [40.0, -40.0, 0.5]
    ys, xs = np.meshgrid(np.arange(H, dtype=np.int32), np.arange(W, dtype=np.int32), indexing="ij")
    o1 = bilinear_warp(im1, np.float32(ys - V / 2.0), np.float32(xs - U / 2.0))
    o2 = bilinear_warp(im2, np.float32(ys + V / 2.0), np.float32(xs + U / 2.0))
    np.savez_compressed(os.path.join(HERE, "warp.npz"), im1=im1, im2=im2, U=U, V=V, o1=o1, o2=o2)


def bitexact():
    """Known answers of getGaussianKernelBitExact (GaussianKernelBitExact.py:55-144).
    sigma <= 0, n in {3,5,7,9}: the fixed binomial kernels the reference stores
    as hex literals; general case: stdlib decimal restatement (prec 28)."""
    out = {
        "lit3": np.array([0.25, 0.5, 0.25]),
        "lit5": np.array([1, 4, 6, 4, 1]) / 16.0,
        "lit7": np.array([1, 3.5, 7, 9, 7, 3.5, 1]) / 32.0,
        "lit9": np.array([4, 13, 30, 51, 60, 51, 30, 13, 4]) / 256.0,
    }
    D = decimal.Decimal
    for n, sigma in ((33, 4.95), (11, 0.0), (15, -2.0), (7, 1.0), (33, 5.3)):
        ctx = decimal.Context(prec=28)
        with decimal.localcontext(ctx):
            sx = D(sigma) if sigma < 0 else D(n) * D("0.15") + D("0.35")
            s2 = D("-0.125") / (sx * sx)
            half = (n - 1) // 2
            vals, tot, x = [], D("0.0"), 1 - n
            for _ in range(half):
                t = (D(x * x) * s2).exp()
                vals.append(t)
                tot += t
                x += 2
            tot = tot * D(2.0) + D("1.0")
            mul = D("1.0") / tot
            res = [None] * n
            for i in range(half):
                res[i] = res[n - 1 - i] = vals[i] * mul
            res[half] = D("1.0") * mul
            out["gen_n%d_s%g" % (n, sigma)] = np.array([float(r) for r in res])
    np.savez_compressed(os.path.join(HERE, "bitexact.npz"), **out)


def piv_pair():
    a = np.array(Image.open(os.path.join(REF_IMAGES, "parabolic01_0.tif")))
    b = np.array(Image.open(os.path.join(REF_IMAGES, "parabolic01_1.tif")))
    assert a.dtype == np.uint8 and a.shape == (512, 512)
    np.savez_compressed(os.path.join(HERE, "piv_pair.npz"), im0=a, im1=b)


if __name__ == "__main__":
    rng = np.random.default_rng(20261203)
    stencils(rng)
    hs_small(rng)
    ls_small(rng)
    gauss_w()
    pil(rng)
    spline(rng)
    warp(rng)
    bitexact()
    if os.path.isdir(REF_IMAGES):
        piv_pair()
    print("golden fixtures written to", HERE)
