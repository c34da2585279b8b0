"""CPU restatement of the warm-start network (test infrastructure only).

The network has no counterpart in the reference: it is this build's learned
form.  This module restates its forward pass, loss and backward pass in
numpy (float64), in two modes:

  bf16=True   the device's rounding points: network input, stored
              activations, weights and the operands of every data-gradient
              convolution rounded to bfloat16 (round to nearest even),
              accumulation in float64 where the device accumulates in float32;
  bf16=False  plain float64, for finite-difference checks of the restated
              backward pass itself.

Pins: the float64 mode's gradients are checked against central finite
differences of its own loss (tests/test_cnn.py); the device is checked
against the bf16 mode.  Layer table, packing and tensor layout follow the
device's (NHWC, space-to-depth channel = (2*dy+dx)*C + c).
"""
import numpy as np

SLOPE = 0.1

# tensor: (level, channels); level L has size H / 2^(L+1)
TENSORS = {"X0": (0, 8), "T0": (0, 32), "CAT0": (0, 64), "T1": (1, 32), "CAT1": (1, 64), "T2": (2, 48),
           "CAT2": (2, 96), "T3": (3, 64), "T3B": (3, 64), "R2": (2, 48), "R1": (1, 32), "R0": (0, 32),
           "HC": (0, 8), "HF": (0, 16)}
F32_TENSORS = ("HC", "HF")
# name, src, src_c0, s2d, dst, dst_c0, d2s, cin, cout, act, level
# (enc0a's input width is 4 x the input planes: see layers())
LAYERS = [
    ("enc0a", "X0", 0, 0, "T0", 0, 0, 8, 32, 1, 0),
    ("enc0b", "T0", 0, 0, "CAT0", 32, 0, 32, 32, 1, 0),
    ("enc1a", "CAT0", 32, 1, "T1", 0, 0, 128, 32, 1, 1),
    ("enc1b", "T1", 0, 0, "CAT1", 32, 0, 32, 32, 1, 1),
    ("enc2a", "CAT1", 32, 1, "T2", 0, 0, 128, 48, 1, 2),
    ("enc2b", "T2", 0, 0, "CAT2", 48, 0, 48, 48, 1, 2),
    ("enc3a", "CAT2", 48, 1, "T3", 0, 0, 192, 64, 1, 3),
    ("enc3b", "T3", 0, 0, "T3B", 0, 0, 64, 64, 1, 3),
    ("dec2u", "T3B", 0, 0, "CAT2", 0, 1, 64, 192, 1, 3),
    ("dec2", "CAT2", 0, 0, "R2", 0, 0, 96, 48, 1, 2),
    ("dec1u", "R2", 0, 0, "CAT1", 0, 1, 48, 128, 1, 2),
    ("dec1", "CAT1", 0, 0, "R1", 0, 0, 64, 32, 1, 1),
    ("dec0u", "R1", 0, 0, "CAT0", 0, 1, 32, 128, 1, 1),
    ("dec0", "CAT0", 0, 0, "R0", 0, 0, 64, 32, 1, 0),
    ("head_c", "R0", 0, 0, "HC", 0, 0, 32, 8, 0, 0),
    ("head_f", "R0", 0, 0, "HF", 0, 0, 32, 16, 0, 0),
]


def layers(planes=2):
    """The layer table for a network with `planes` input planes."""
    out = [list(L) for L in LAYERS]
    out[0][7] = 4 * planes
    return [tuple(L) for L in out]


def num_params(planes=2):
    return sum(9 * L[7] * L[8] + L[8] for L in layers(planes))


def unpack(flat, planes=2):
    """Packed vector -> list of (W [cout, 9, cin], b [cout])."""
    out, q = [], 0
    for L in layers(planes):
        cin, cout = L[7], L[8]
        W = flat[q:q + 9 * cin * cout].reshape(cout, 9, cin)
        q += 9 * cin * cout
        b = flat[q:q + cout]
        q += cout
        out.append((W, b))
    assert q == len(flat)
    return out


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _s2d(a):
    """[B, 2h, 2w, C] -> [B, h, w, 4C], channel (2*dy+dx)*C + c."""
    B, H2, W2, C = a.shape
    return a.reshape(B, H2 // 2, 2, W2 // 2, 2, C).transpose(0, 1, 3, 2, 4, 5).reshape(B, H2 // 2, W2 // 2, 4 * C)


def _d2s(a):
    """[B, h, w, 4C] -> [B, 2h, 2w, C] (inverse of _s2d)."""
    B, h, w, C4 = a.shape
    C = C4 // 4
    return a.reshape(B, h, w, 2, 2, C).transpose(0, 1, 3, 2, 4, 5).reshape(B, 2 * h, 2 * w, C)


def _im2col(x):
    """[B, h, w, C] -> [B*h*w, 9*C], tap-major (tap = 3*(dy+1) + dx+1), zero padding."""
    B, h, w, C = x.shape
    xp = np.zeros((B, h + 2, w + 2, C))
    xp[:, 1:-1, 1:-1] = x
    cols = [xp[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w] for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
    return np.concatenate(cols, axis=3).reshape(B * h * w, 9 * C)


def _col2im(c, shape):
    """Adjoint of _im2col."""
    B, h, w, C = shape
    c = c.reshape(B, h, w, 9, C)
    xp = np.zeros((B, h + 2, w + 2, C))
    t = 0
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            xp[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w] += c[:, :, :, t]
            t += 1
    return xp[:, 1:-1, 1:-1]


def _alloc(B, H, W, planes=2):
    T = {k: np.zeros((B, H >> (lv + 1), W >> (lv + 1), C)) for k, (lv, C) in TENSORS.items()}
    T["X0"] = np.zeros((B, H >> 1, W >> 1, 4 * planes))
    return T


def _read(T, L):
    name, src, c0, s2d = L[0], L[1], L[2], L[3]
    cin = L[7]
    a = T[src][..., c0:c0 + (cin // 4 if s2d else cin)]
    return _s2d(a) if s2d else a


def _write(T, L, y, add=False):
    dst, c0, d2s, cout = L[4], L[5], L[6], L[8]
    v = _d2s(y) if d2s else y
    C = cout // 4 if d2s else cout
    if add:
        T[dst][..., c0:c0 + C] += v
    else:
        T[dst][..., c0:c0 + C] = v


def forward(params, inp, bf=True):
    """inp: [B, P, H, W] planar input (planes 0, 1 frames, 2, 3 flow).
    Returns the tensor dict (all activations)."""
    rnd = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    inp = np.asarray(inp, np.float32)
    B, P_, H, W = inp.shape
    T = _alloc(B, H, W, P_)
    planes = []
    for k in range(P_):
        x = inp[:, k]
        planes.append((x - np.float32(128)) * np.float32(1.0 / 64.0) if k < 2 else x)
    st = np.stack(planes, axis=-1)                         # [B, H, W, P]
    T["X0"][:] = rnd(_s2d(st))                             # channel (2dy+dx)*P + plane
    P = unpack(np.asarray(params, np.float64), P_)
    for L, (Wt, b) in zip(layers(P_), P):
        x = _read(T, L)
        Wm = rnd(Wt).reshape(L[8], -1)
        B_, h, w, _ = x.shape
        y = (_im2col(x) @ Wm.T + b).reshape(B_, h, w, L[8])
        if L[9]:
            y = np.where(y > 0, y, SLOPE * y)
        if L[4] not in F32_TENSORS:
            y = rnd(y)
        else:
            y = y.astype(np.float32).astype(np.float64) if bf else y
        _write(T, L, y)
    return T


def heads(T):
    """Planar outputs: coarse [B, 4, h, w], fine [B, 4, H, W]."""
    hc, hf = T["HC"], T["HF"]
    coarse = hc[..., :4].transpose(0, 3, 1, 2)
    B, h, w, _ = hf.shape
    fine = _d2s(hf.reshape(B, h, w, 4, 4).reshape(B, h, w, 16)).transpose(0, 3, 1, 2)
    return coarse, fine


def loss_and_head_grads(T, lab_c, lab_f, fw, eps2=1e-6):
    """Weighted Charbonnier loss of the device (mean per level), and its
    gradients with respect to HC (8 channels) and HF (16 channels)."""
    hc, hf = T["HC"], T["HF"]
    B, h, w, _ = hc.shape
    n = B * h * w
    pc = hc[..., :4].transpose(0, 3, 1, 2)
    dc = pc - lab_c
    qc = np.sqrt(dc * dc + eps2)
    loss = sum(fw[f] * qc[:, f].sum() / n for f in range(4))
    ghc = np.zeros_like(hc)
    ghc[..., :4] = (np.array(fw[:4]).reshape(1, 4, 1, 1) * dc / qc / n).transpose(0, 2, 3, 1)
    # fine: HF channel s*4+f at (y, x) is field f at (2y + s//2, 2x + s%2)
    pf = _d2s(hf).transpose(0, 3, 1, 2)                  # [B, 4, H, W]
    df = pf - lab_f
    qf = np.sqrt(df * df + eps2)
    loss += sum(fw[4 + f] * qf[:, f].sum() / (4 * n) for f in range(4))
    gpf = np.array(fw[4:]).reshape(1, 4, 1, 1) * df / qf / (4 * n)
    ghf = _s2d(gpf.transpose(0, 2, 3, 1))
    return loss, ghc, ghf


def backward(params, T, ghc, ghf, bf=True):
    """Packed parameter gradients, following the device's rounding points."""
    rnd = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    planes = T["X0"].shape[-1] // 4
    LAY = layers(planes)
    P = unpack(np.asarray(params, np.float64), planes)
    B = T["X0"].shape[0]
    G = {k: np.zeros_like(v) for k, v in T.items()}
    G["HC"][:] = ghc
    G["HF"][:] = ghf
    if bf:
        G["HC"][:] = ghc.astype(np.float32)
        G["HF"][:] = ghf.astype(np.float32)
    have = {"HC", "HF"}
    grads = [None] * len(LAY)
    for li in range(len(LAY) - 1, -1, -1):
        L = LAY[li]
        Wt, b = P[li]
        cout, cin = L[8], L[7]
        # output gradient and forward output in conv space
        dst, c0, d2s = L[4], L[5], L[6]
        C = cout // 4 if d2s else cout
        dy = G[dst][..., c0:c0 + C]
        yv = T[dst][..., c0:c0 + C]
        if d2s:
            dy, yv = _s2d(dy), _s2d(yv)
        dz = dy * np.where(yv > 0, 1.0, SLOPE) if L[9] else dy
        x = _read(T, L)
        _, h, w, _ = x.shape
        cols = _im2col(x)
        dzf = dz.reshape(-1, cout)
        gw = (dzf.T @ cols).reshape(cout, 9, cin)
        gb = dzf.sum(axis=0)
        grads[li] = (gw, gb)
        if li == 0:
            break
        Wm = rnd(Wt).reshape(cout, -1)
        dx = _col2im(rnd(dzf) @ Wm, x.shape)
        src, sc0, s2d = L[1], L[2], L[3]
        Cs = cin // 4 if s2d else cin
        v = _d2s(dx) if s2d else dx
        if src in have:
            G[src][..., sc0:sc0 + Cs] = rnd(G[src][..., sc0:sc0 + Cs] + v)
        else:
            G[src][..., sc0:sc0 + Cs] = rnd(v)
        have.add(src)
    flat = []
    for gw, gb in grads:
        flat.append(gw.reshape(-1))
        flat.append(gb)
    return np.concatenate(flat)


def init_params(seed=0, scale=1.0, planes=2):
    """He-style random initialisation (test parameters)."""
    rng = np.random.default_rng(seed)
    flat = []
    for L in layers(planes):
        cin, cout = L[7], L[8]
        std = scale * np.sqrt(2.0 / (9 * cin)) if L[9] else scale * 0.1 * np.sqrt(1.0 / (9 * cin))
        flat.append(rng.standard_normal(9 * cin * cout) * std)
        flat.append(np.zeros(cout))
    return np.concatenate(flat).astype(np.float32)
