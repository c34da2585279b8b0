"""Correction network: the CPU restatement against the same network written
with torch's operators (forward) and torch's autograd (backward), and
against finite differences of its own loss; the device's forward pass,
loss and gradients against the restatement at the device's bfloat16
rounding points; the fused correction step against the forward pass."""
import numpy as np
import pytest

from oracle import cnn


def _batch(B, H, W, seed, planes=4):
    """Planar input (frames in [0, 255], then flow planes) and labels."""
    rng = np.random.default_rng(seed)
    inp = np.concatenate([rng.uniform(0, 255, (B, 2, H, W)),
                          rng.standard_normal((B, planes - 2, H, W)) * 2.0], axis=1).astype(np.float32)
    lab_c = rng.standard_normal((B, 4, H // 2, W // 2))
    lab_f = rng.standard_normal((B, 4, H, W)) * 0.3
    return inp, lab_c, lab_f


FW = np.array([1.0, 1.0, 2.0, 2.0, 1.0, 1.0, 2.0, 2.0])


def test_packing_and_size():
    for planes in (2, 4):
        assert cnn.num_params(planes) == sum(9 * L[7] * L[8] + L[8] for L in cnn.layers(planes))
        assert 450_000 < cnn.num_params(planes) < 650_000
    p = np.arange(cnn.num_params(4), dtype=np.float64)
    W0, b0 = cnn.unpack(p, 4)[0]
    assert W0.shape == (32, 9, 16) and b0[0] == 9 * 16 * 32


def test_space_to_depth_round_trip_and_order():
    a = np.arange(2 * 4 * 6 * 3, dtype=np.float64).reshape(2, 4, 6, 3)
    s = cnn._s2d(a)
    assert s.shape == (2, 2, 3, 12)
    # channel (2*dy + dx) * C + c at (y, x) is a[2y+dy, 2x+dx, c]
    assert s[1, 1, 2, (2 * 1 + 0) * 3 + 2] == a[1, 3, 4, 2]
    assert np.array_equal(cnn._d2s(s), a)


def test_im2col_adjoint():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 5, 7, 3))
    y = rng.standard_normal((2 * 5 * 7, 27))
    assert abs((cnn._im2col(x) * y).sum() - (x * cnn._col2im(y, x.shape)).sum()) < 1e-9


def test_restated_backward_matches_finite_differences():
    inp, lc, lf = _batch(1, 32, 32, 1)
    p = cnn.init_params(3, scale=1.0, planes=4).astype(np.float64)
    T = cnn.forward(p, inp, bf=False)
    loss, ghc, ghf = cnn.loss_and_head_grads(T, lc, lf, FW)
    g = cnn.backward(p, T, ghc, ghf, bf=False)
    rng = np.random.default_rng(5)
    # a spread of indices over every layer's weights and biases
    idx, q = [], 0
    for L in cnn.layers(4):
        nw = 9 * L[7] * L[8]
        idx += list(q + rng.integers(0, nw, 2)) + [q + nw + int(rng.integers(0, L[8]))]
        q += nw + L[8]
    h = 1e-6   # the Charbonnier penalty is sharply curved near zero residuals
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp = cnn.loss_and_head_grads(cnn.forward(pp, inp, bf=False), lc, lf, FW)[0]
        lm = cnn.loss_and_head_grads(cnn.forward(pm, inp, bf=False), lc, lf, FW)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-4 * max(1e-3, abs(fd)) + 1e-9, (i, fd, g[i])


def _torch_network(params, inp, planes=4):
    """The network of the layer table written with torch's own operators
    (NCHW float64; F.conv2d is cross-correlation with zero padding, the
    weight [cout][tap][cin] with tap = 3 (dy + 1) + dx + 1 becomes
    [cout][cin][dy + 1][dx + 1]; space-to-depth channel (2 dy + dx) C + c).
    Returns the planar heads and the parameter tensors (for autograd)."""
    import torch
    import torch.nn.functional as F

    def s2d(a):
        B, C, H2, W2 = a.shape
        return a.reshape(B, C, H2 // 2, 2, W2 // 2, 2).permute(0, 3, 5, 1, 2, 4).reshape(B, 4 * C, H2 // 2, W2 // 2)

    def d2s(a):
        B, C4, h, w = a.shape
        C = C4 // 4
        return a.reshape(B, 2, 2, C, h, w).permute(0, 3, 4, 1, 5, 2).reshape(B, C, 2 * h, 2 * w)

    x = torch.tensor(np.asarray(inp, np.float32))
    # frames scaled in float32, as the network's input stage scales them
    planes_t = [((x[:, k] - 128.0) * (1.0 / 64.0) if k < 2 else x[:, k]).double() for k in range(x.shape[1])]
    pieces = {"X0": {0: s2d(torch.stack(planes_t, dim=1))}}

    def read(name, c0, n):
        ps = pieces[name]
        cs = sorted(ps)
        full = torch.cat([ps[c] for c in cs], dim=1)
        off = cs[0]
        assert all(cs[i] + ps[cs[i]].shape[1] == cs[i + 1] for i in range(len(cs) - 1))
        return full[:, c0 - off:c0 - off + n]

    flat = torch.tensor(np.asarray(params, np.float64), requires_grad=True)
    q = 0
    for L in cnn.layers(planes):
        name, src, c0, s2d_in, dst, d0, d2s_out, cin, cout, act = L[:10]
        w = flat[q:q + 9 * cin * cout].reshape(cout, 3, 3, cin).permute(0, 3, 1, 2)
        q += 9 * cin * cout
        b = flat[q:q + cout]
        q += cout
        xin = read(src, c0, cin // 4 if s2d_in else cin)
        if s2d_in:
            xin = s2d(xin)
        y = F.conv2d(xin, w, b, padding=1)
        if act:
            y = F.leaky_relu(y, cnn.SLOPE)
        if d2s_out:
            y = d2s(y)
        pieces.setdefault(dst, {})[d0] = y
    hc, hf = pieces["HC"][0], pieces["HF"][0]
    return hc[:, :4], d2s(hf), flat


def test_restatement_matches_torch_forward_and_autograd():
    """The numpy restatement (float64 mode) against the same network written
    with torch's conv2d / leaky_relu and differentiated by torch's autograd:
    a flipped kernel, a transposed tap or channel order, a wrong routing or
    a wrong gradient term fails here."""
    import torch
    inp, lc, lf = _batch(2, 32, 48, 21)
    p = cnn.init_params(17, scale=1.0, planes=4).astype(np.float64)
    T = cnn.forward(p, inp, bf=False)
    rc, rf = cnn.heads(T)
    tc, tf, flat = _torch_network(p, inp)
    for r, t in ((rc, tc), (rf, tf)):
        t = t.detach().numpy()
        assert r.shape == t.shape
        assert np.abs(r - t).max() <= 1e-10 * max(1.0, np.abs(t).max())
    loss, ghc, ghf = cnn.loss_and_head_grads(T, lc, lf, FW)
    g = cnn.backward(p, T, ghc, ghf, bf=False)
    # the same weighted Charbonnier loss in torch
    n = tc.shape[0] * tc.shape[2] * tc.shape[3]
    eps2 = 1e-6
    dc = tc - torch.tensor(lc)
    df = tf - torch.tensor(lf)
    fw = torch.tensor(FW)
    tl = ((torch.sqrt(dc * dc + eps2).sum(dim=(0, 2, 3)) * fw[:4]).sum() / n +
          (torch.sqrt(df * df + eps2).sum(dim=(0, 2, 3)) * fw[4:]).sum() / (4 * n))
    assert abs(float(tl) - loss) <= 1e-10 * abs(loss)
    tl.backward()
    tg = flat.grad.numpy()
    q = 0
    for L in cnn.layers(4):
        k = 9 * L[7] * L[8] + L[8]
        a, b = g[q:q + k], tg[q:q + k]
        assert np.linalg.norm(a - b) <= 1e-9 * np.linalg.norm(b) + 1e-14, L[0]
        q += k


# ------------------------------------------------------------------ GPU
@pytest.fixture(scope="module")
def net():
    from ofri.net import Net
    return Net(planes=4)


@pytest.mark.gpu
def test_device_parameter_count_and_round_trip(net):
    from ofri.net import Net
    assert net.num_params == cnn.num_params(4)
    assert Net(planes=2).num_params == cnn.num_params(2)
    p = cnn.init_params(11, planes=4)
    net.set_params(p)
    assert np.array_equal(net.get_params(), p)


@pytest.mark.gpu
@pytest.mark.parametrize("B,H,W", [(1, 32, 32), (2, 32, 48), (1, 64, 80)])
def test_device_forward_matches_restatement(net, B, H, W):
    inp, _, _ = _batch(B, H, W, 2)
    p = cnn.init_params(7, planes=4)
    net.set_params(p)
    coarse, fine = net.predict(inp)
    rc, rf = cnn.heads(cnn.forward(p.astype(np.float64), inp, bf=True))
    gc, gf = coarse.download(), fine.download()
    assert gc.shape == rc.shape and gf.shape == rf.shape
    for g, r in ((gc, rc), (gf, rf)):
        err = np.abs(g - r)
        scale = np.abs(r).max()
        print("[cnn fwd %dx%dx%d] max |d| %.3g of %.3g, mean %.3g" % (B, H, W, err.max(), scale, err.mean()))
        assert err.max() <= 2e-2 * scale and err.mean() <= 2e-3 * scale


@pytest.mark.gpu
def test_device_forward_batch_equals_single_images(net):
    """24 images of 256 x 256 in one pass run the full-resolution layers on
    16 x 16 tiles with 32-channel groups; one image alone runs them on 8 x 8
    tiles.  Every image's heads equal, bit for bit, those of the image passed
    alone: an output is the same sequence of MFMAs whichever tile shape and
    workgroup compute it."""
    B, H, W = 24, 256, 256
    inp, _, _ = _batch(B, H, W, 17)
    net.set_params(cnn.init_params(3, planes=4))
    coarse, fine = net.predict(inp)
    gc, gf = coarse.download(), fine.download()
    assert np.isfinite(gc).all() and np.isfinite(gf).all()
    for i in (0, 5, 11, 23):
        c1, f1 = net.predict(inp[i:i + 1])
        assert np.array_equal(c1.download()[0], gc[i]) and np.array_equal(f1.download()[0], gf[i]), i


@pytest.mark.gpu
def test_device_gradients_match_restatement(net):
    """Per layer, the device gradient's distance from the bf16 restatement is
    measured against the bf16 rounding noise itself (the distance between the
    bf16 and the float64 restatements): a wrong index or missing term shows
    up far above that noise, rounding differences stay within a few times it."""
    B, H, W = 2, 64, 64
    inp, lc, lf = _batch(B, H, W, 4)
    lc, lf = lc.astype(np.float32).astype(np.float64), lf.astype(np.float32).astype(np.float64)
    p = cnn.init_params(9, planes=4)
    net.set_params(p)
    loss = net.train_step(inp, lc.astype(np.float32), lf.astype(np.float32), lr=0.0, field_w=FW)
    g = net.get_grads()
    T = cnn.forward(p.astype(np.float64), inp, bf=True)
    rl, ghc, ghf = cnn.loss_and_head_grads(T, lc, lf, FW)
    rg = cnn.backward(p.astype(np.float64), T, ghc, ghf, bf=True)
    T64 = cnn.forward(p.astype(np.float64), inp, bf=False)
    _, h64c, h64f = cnn.loss_and_head_grads(T64, lc, lf, FW)
    rg64 = cnn.backward(p.astype(np.float64), T64, h64c, h64f, bf=False)
    assert abs(loss - rl) <= 1e-3 * abs(rl)
    q, bad = 0, []
    for L in cnn.layers(4):
        n = 9 * L[7] * L[8] + L[8]
        a, b, c = g[q:q + n], rg[q:q + n], rg64[q:q + n]
        dev = np.linalg.norm(a - b) / np.linalg.norm(b)
        noise = np.linalg.norm(c - b) / np.linalg.norm(b)
        cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30))
        print("[cnn grad %-6s] |dev - bf16| %.3g   bf16 noise |f64 - bf16| %.3g   cos %.6f" % (L[0], dev, noise, cos))
        if not (dev <= 3.0 * noise + 2e-3 and cos >= 0.99):
            bad.append(L[0])
        q += n
    assert not bad, bad


@pytest.mark.gpu
def test_device_adam_steps_reduce_loss(net):
    B, H, W = 2, 32, 32
    inp, lc, lf = _batch(B, H, W, 6)
    lc = (lc * 0.5).astype(np.float32)
    lf = (lf * 0.5).astype(np.float32)
    net.set_params(cnn.init_params(13, planes=4))
    losses = [net.train_step(inp, lc, lf, lr=2e-3, field_w=FW) for _ in range(40)]
    print("[cnn adam] loss %.4f -> %.4f" % (losses[0], losses[-1]))
    assert losses[-1] < 0.8 * losses[0]


@pytest.mark.gpu
def test_device_correction_step_equals_forward_plus_add(net):
    """ofri_net_correct (separate input planes, coarse head skipped, fine
    fields 0, 1 added into the flow in place) gives exactly the forward
    pass's fine fields added to the flow, also when the flow planes are the
    network's own inputs."""
    from ofri.device import DeviceArray
    net.init(5)
    inp, _, _ = _batch(1, 64, 80, 31)
    _, fine = net.predict(inp)
    f = fine.download()[0]
    planes = [DeviceArray.from_host(inp[0, k]) for k in range(4)]
    u, v = planes[2], planes[3]            # the flow inputs, updated in place
    want_u = (inp[0, 2] + f[0]).astype(np.float32)
    want_v = (inp[0, 3] + f[1]).astype(np.float32)
    net.correct(planes, u, v)
    assert np.array_equal(u.download(), want_u) and np.array_equal(v.download(), want_v)
