"""Synthetic PIV image pairs with known displacement fields.

Particles: positions uniform at `density` per pixel, diameter d ~ N(2.5, 0.5)
clipped to [1, 4], peak I0 ~ U[150, 255], rendered as I0*exp(-8 r^2 / d^2),
summed and clipped to [0, 255].  Frame 2 renders the same particles moved by
the flow evaluated at their positions (one explicit step).  Additive Gaussian
noise (sigma 2 grey levels by default) is added to both frames.  Output is
float32, not quantised.  Every draw comes from one numpy Generator.

Flow families (u along x / columns, v along y / rows, in pixels):
  poiseuille   u = 4*umax*y(H-y)/H^2, v = 0
  lamb_oseen   vortex at the centre, core radius rc, peak speed scaled to vmax
  turbulence   uniform shear + 256 random Fourier modes with a k^-5/3
               spectrum, made divergence-free through a stream function
"""
import numpy as np

MASTER_SEED = 20261203


def flow_poiseuille(H, W, umax=3.5):
    y = np.arange(H, dtype=np.float64)[:, None]
    u = 4.0 * umax * y * (H - y) / (H * H) * np.ones((1, W))
    return u, np.zeros((H, W))


def _lamb_oseen_at(x, y, H, W, rc, vmax):
    cx, cy = (W - 1) / 2.0, (H - 1) / 2.0
    dx, dy = x - cx, y - cy
    r2 = dx * dx + dy * dy
    r2s = np.maximum(r2, 1e-12)
    # tangential speed ~ (1 - exp(-r^2/rc^2)) / r, peak at r ~ 1.1209 rc
    vt = (1.0 - np.exp(-r2s / (rc * rc))) / np.sqrt(r2s)
    rp = 1.1209 * rc
    peak = (1.0 - np.exp(-(rp * rp) / (rc * rc))) / rp
    vt = vt * (vmax / peak)
    r = np.sqrt(r2s)
    return -vt * dy / r, vt * dx / r


def flow_lamb_oseen(H, W, rc=80.0, vmax=4.0):
    y, x = np.mgrid[0:H, 0:W].astype(np.float64)
    return _lamb_oseen_at(x, y, H, W, rc, vmax)


class Turbulence:
    """Shear u = shear*(y/H - 0.5) plus a divergence-free random field
    (stream function psi = sum a_k sin(k.x + phi_k), u = dpsi/dy, v = -dpsi/dx)
    with E(k) ~ k^-5/3, rescaled so max |displacement| = dmax."""

    def __init__(self, H, W, rng, modes=256, shear=2.0, dmax=4.0):
        self.H, self.W, self.shear = H, W, shear
        kmag = np.exp(rng.uniform(np.log(2 * np.pi / max(H, W) * 2), np.log(2 * np.pi / 8.0), modes))
        ang = rng.uniform(0, 2 * np.pi, modes)
        self.kx, self.ky = kmag * np.cos(ang), kmag * np.sin(ang)
        self.phi = rng.uniform(0, 2 * np.pi, modes)
        # velocity amplitude per mode ~ sqrt(E(k) dk) ~ k^(-5/6) (log-spaced dk ~ k)
        self.amp = kmag ** (-5.0 / 6.0) * np.sqrt(kmag) / kmag
        self.scale = 1.0
        y, x = np.mgrid[0:H:4, 0:W:4].astype(np.float64)
        u, v = self._raw(x, y)
        self.scale = dmax / max(np.sqrt(u * u + v * v).max(), 1e-12)
        u, v = self._raw(x, y)
        self.scale *= dmax / max(np.sqrt(u * u + v * v).max(), 1e-12)

    def _raw(self, x, y):
        u = np.zeros_like(x)
        v = np.zeros_like(x)
        for kx, ky, ph, a in zip(self.kx, self.ky, self.phi, self.amp):
            c = np.cos(kx * x + ky * y + ph) * a
            u += c * ky
            v -= c * kx
        u = u * self.scale + self.shear * (y / self.H - 0.5) * self.scale
        return u, v * self.scale

    def at(self, x, y):
        return self._raw(x, y)


def render(H, W, xs, ys, d, I0, rng, noise=2.0):
    """Sum of Gaussian particle images, clipped to [0, 255], plus noise."""
    img = np.zeros(H * W, np.float64)
    R = 4
    ix = np.floor(xs).astype(np.int64)
    iy = np.floor(ys).astype(np.int64)
    inv = 8.0 / (d * d)
    for oy in range(-R, R + 2):
        py = iy + oy
        oky = (py >= 0) & (py < H)
        ddy = py - ys
        for ox in range(-R, R + 2):
            px = ix + ox
            ok = oky & (px >= 0) & (px < W)
            ddx = px - xs
            val = I0 * np.exp(-(ddx * ddx + ddy * ddy) * inv)
            img += np.bincount((py * W + px)[ok], weights=val[ok], minlength=H * W)
    img = np.clip(img, 0, 255).reshape(H, W)
    if noise > 0:
        img = img + rng.normal(0.0, noise, (H, W))
    return img.astype(np.float32)


def piv_pair(H, W, density, flow, rng, noise=2.0, truth=True):
    """Returns (im1, im2, u, v) with float32 images and the dense flow
    (u = v = None when truth is False: the dense field of the turbulence
    family costs far more than rendering).

    flow: 'poiseuille' | 'lamb_oseen' | 'turbulence' or a callable
    (x, y) -> (u, v) evaluated at particle positions."""
    n = rng.poisson(density * H * W)
    xs = rng.uniform(0, W, n)
    ys = rng.uniform(0, H, n)
    d = np.clip(rng.normal(2.5, 0.5, n), 1.0, 4.0)
    I0 = rng.uniform(150, 255, n)
    if flow == "poiseuille":
        def f(x, y):
            return 4.0 * 3.5 * y * (H - y) / (H * H), np.zeros_like(x)
    elif flow == "lamb_oseen":
        def f(x, y):
            return _lamb_oseen_at(x, y, H, W, 80.0 * min(H, W) / 1024.0 if min(H, W) < 1024 else 80.0, 4.0)
    elif flow == "turbulence":
        f = Turbulence(H, W, rng).at
    else:
        f = flow
    up, vp = f(xs, ys)
    im1 = render(H, W, xs, ys, d, I0, rng, noise)
    im2 = render(H, W, xs + up, ys + vp, d, I0, rng, noise)
    if not truth:
        return im1, im2, None, None
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    u, v = f(xx, yy)
    return im1, im2, np.asarray(u, np.float32), np.asarray(v, np.float32)


def config_pair(cfg, seed=MASTER_SEED, size=None):
    """The synthetic pair of a workload configuration:
      0  256x256 Poiseuille, density 0.05
      1  1024x1024 Lamb-Oseen (core 80 px, peak 4 px), density 0.04
      2  2048x2048 shear + turbulence, density 0.06
      4  1024x1024 LS(HS) pair (Lamb-Oseen, density 0.04)
    `size` overrides the edge length (for reduced CPU checks)."""
    rng = np.random.default_rng(np.random.SeedSequence(seed).spawn(8)[cfg])
    if cfg == 0:
        n = size or 256
        return piv_pair(n, n, 0.05, "poiseuille", rng)
    if cfg in (1, 4):
        n = size or 1024
        return piv_pair(n, n, 0.04, "lamb_oseen", rng)
    if cfg == 2:
        n = size or 2048
        return piv_pair(n, n, 0.06, "turbulence", rng)
    raise ValueError("no single-pair configuration %r" % cfg)


def turbulence_modes(H, W, rng, modes=256):
    """(4, modes) float32 rows kx, ky, phase, amplitude of the turbulence
    family (as Turbulence draws them) for the device renderer."""
    kmag = np.exp(rng.uniform(np.log(2 * np.pi / max(H, W) * 2), np.log(2 * np.pi / 8.0), modes))
    ang = rng.uniform(0, 2 * np.pi, modes)
    phi = rng.uniform(0, 2 * np.pi, modes)
    amp = kmag ** (-5.0 / 6.0) * np.sqrt(kmag) / kmag
    return np.stack([kmag * np.cos(ang), kmag * np.sin(ang), phi, amp]).astype(np.float32)


def device_config_pair(cfg, seed=MASTER_SEED, size=None):
    """The workload configuration's pair rendered on the GPU (ofri_render_pair:
    the same particle model and flow families, drawn by the device from one
    seed; rendering a 2048^2 turbulence pair on the host takes half a
    minute).  Returns (im1, im2, u, v) DeviceArrays.
      1  1024x1024 Lamb-Oseen (core 80 px, peak 4 px), density 0.04
      2  2048x2048 shear 2.0 + turbulence (256 modes, max 4 px), density 0.06"""
    from . import ops
    rng = np.random.default_rng(np.random.SeedSequence(seed).spawn(8)[cfg])
    dev_seed = int(rng.integers(0, 2 ** 62))
    if cfg == 1:
        n = size or 1024
        rc = 80.0 * n / 1024.0 if n < 1024 else 80.0
        return ops.render_pair(n, n, 1, 0.04, dev_seed, [(n - 1) / 2.0, (n - 1) / 2.0, rc, 4.0], truth=True)
    if cfg == 2:
        n = size or 2048
        return ops.render_pair(n, n, 2, 0.06, dev_seed, [2.0, 4.0], turbulence_modes(n, n, rng), truth=True)
    raise ValueError("no device-rendered configuration %r" % cfg)


def _batch_item(args):
    H, W, dens, fam, child, truth = args
    return piv_pair(H, W, dens, fam, np.random.default_rng(child), truth=truth)


# the device-rendered uneven batch draws its pairs' flow parameters from
# seeds here on (training 0 .. 19,999; held-out evaluation 1,000,000 on;
# calibration 3,000,000 on)
BATCH_SEED = 5000000


def device_batch_pairs(n_pairs=64, seed=MASTER_SEED, sizes=None, truth=False):
    """The uneven batch rendered on the device: H, W drawn from {256, 384,
    ..., 2048}, flow family, density and flow parameters as the training
    pairs draw them (ofri.train.pair_spec, seeds from BATCH_SEED).  Returns
    a list of (im1, im2) or (im1, im2, u, v) DeviceArrays and the specs."""
    from . import ops
    from .train import pair_spec
    rng = np.random.default_rng(np.random.SeedSequence(seed).spawn(8)[4])
    choices = sizes if sizes is not None else np.arange(256, 2049, 128)
    out, specs = [], []
    for k in range(n_pairs):
        H, W = int(rng.choice(choices)), int(rng.choice(choices))
        sp = pair_spec(BATCH_SEED + k, H, W)
        out.append(ops.render_pair(H, W, sp["family"], sp["density"], sp["seed"], sp["par"], sp["modes"],
                                   truth=truth))
        specs.append(sp)
    return out, specs


def batch_pairs(n_pairs=64, seed=MASTER_SEED, sizes=None, truth=False, workers=1):
    """Uneven batch: H, W from {256, 384, ..., 2048}, density U[0.02, 0.08],
    flow family uniform over the three families (max 4 px).  Items are
    (im1, im2, u, v); u, v are None unless truth is True.  workers > 1 renders
    the pairs in that many processes (same pairs: every pair has its own
    child seed)."""
    ss = np.random.SeedSequence(seed).spawn(8)[3]
    rng = np.random.default_rng(ss)
    choices = sizes if sizes is not None else np.arange(256, 2049, 128)
    specs = []
    for child in ss.spawn(n_pairs):
        H, W = int(rng.choice(choices)), int(rng.choice(choices))
        dens = rng.uniform(0.02, 0.08)
        fam = ["poiseuille", "lamb_oseen", "turbulence"][rng.integers(0, 3)]
        specs.append((H, W, dens, fam, child, truth))
    if workers <= 1:
        return [_batch_item(a) for a in specs]
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(_batch_item, specs))
