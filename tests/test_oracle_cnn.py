"""Pins of the U-Net oracle (oracle/cnn_ref.py) by mathematics: an
independent convolution routine, adjoint (dot-product) identities, torch
autograd in FP64, central finite differences, and torch's own BF16
rounding.  The network has no counterpart in the reference."""
import numpy as np
import pytest
import torch
from scipy.signal import correlate2d

from oracle import cnn_ref as R


def test_bf16_rounding_matches_torch_and_ties_to_even():
    rng = np.random.default_rng(0)
    a = np.concatenate([rng.normal(0, 1, 20000), rng.normal(0, 1e-30, 100), rng.normal(0, 1e30, 100)]).astype(np.float32)
    want = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(R.bf16(a), want)
    one = np.float32(1.0)
    assert R.bf16(np.array([one + np.float32(2.0 ** -8)]))[0] == 1.0             # tie -> even
    assert R.bf16(np.array([one + np.float32(3 * 2.0 ** -8)]))[0] == 1.0 + 2.0 ** -6  # tie -> even


def test_conv3x3_matches_scipy_correlate():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 7, 9, 3))
    w = rng.normal(size=(4, 9, 3))
    y = R.conv3x3(x, w)
    for n in range(2):
        for co in range(4):
            ref = sum(correlate2d(x[n, :, :, ci], w[co, :, ci].reshape(3, 3), mode="same") for ci in range(3))
            assert np.allclose(y[n, :, :, co], ref, rtol=1e-12, atol=1e-12)


def test_conv_gradients_are_the_adjoints():
    # <conv(x, w), g> = <x, xgrad(g, w)> = <w, wgrad(g, x)>
    rng = np.random.default_rng(2)
    x = rng.normal(size=(2, 6, 5, 8))
    w = rng.normal(size=(16, 9, 8))
    g = rng.normal(size=(2, 6, 5, 16))
    a = np.sum(R.conv3x3(x, w) * g)
    b = np.sum(x * R.conv3x3_xgrad(g, w, 6, 5))
    c = np.sum(w * R.conv3x3_wgrad(g, x))
    assert abs(a - b) <= 1e-10 * abs(a) and abs(a - c) <= 1e-10 * abs(a)


def test_pool_and_upsample():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(2, 6, 4, 5))
    assert np.allclose(R.pool2(x), x.reshape(2, 3, 2, 2, 2, 5).mean(axis=(2, 4)), rtol=1e-14)
    u = R.up2(R.pool2(x))
    assert u.shape == x.shape and np.allclose(R.pool2(u), R.pool2(x))


def _torch_unet(lay, p, x, target, mask):
    """the same network in torch FP64 (conv2d is cross-correlation)"""
    P = torch.tensor(p.astype(np.float64), requires_grad=True)

    def conv(h, L):
        w = P[L["woff"]:L["woff"] + L["cout"] * 9 * L["cin_pad"]].reshape(L["cout"], 3, 3, L["cin_pad"])
        w = w.permute(0, 3, 1, 2)[:, :L["cin"]]
        b = P[L["boff"]:L["boff"] + L["cout"]]
        return torch.relu(torch.nn.functional.conv2d(h, w, b, padding=1))

    h, skips = torch.from_numpy(x).permute(0, 3, 1, 2), []
    for lv in range(lay.levels):
        h = conv(conv(h, lay.enc[lv][0]), lay.enc[lv][1])
        skips.append(h)
        if lv < lay.levels - 1:
            h = torch.nn.functional.avg_pool2d(h, 2)
    for lv in range(lay.levels - 2, -1, -1):
        h = torch.cat([torch.nn.functional.interpolate(h, scale_factor=2, mode="nearest"), skips[lv]], dim=1)
        h = conv(conv(h, lay.dec[lv][0]), lay.dec[lv][1])
    out = torch.einsum("nchw,c->nhw", h, P[lay.head_w:lay.head_w + lay.base]) + P[lay.head_b]
    m = torch.from_numpy(mask.astype(bool))
    err = torch.where(m, out - torch.from_numpy(target), torch.zeros_like(out))
    loss = (err ** 2).sum() / m.sum()
    loss.backward()
    return float(loss.detach()), out.detach().numpy(), P.grad.numpy()


@pytest.mark.parametrize("in_ch,base,levels", [(4, 4, 3), (3, 8, 2)])
def test_oracle_fp64_matches_torch_autograd(in_ch, base, levels):
    rng = np.random.default_rng(4)
    net = R.UNetRef(in_ch, base, levels, emulate=False)
    p = R.init_params(in_ch, base, levels, 5).astype(np.float64)
    p[np.arange(p.size) % 7 == 3] += 0.01  # nonzero biases too
    x = rng.normal(size=(2, 8, 8, in_ch))
    t = rng.normal(size=(2, 8, 8))
    mask = rng.random((2, 8, 8)) < 0.7
    loss, out, g = net.loss_and_grad(x, t, mask, p)
    tl, tout, tg = _torch_unet(net.lay, p, x, t, mask)
    assert np.allclose(out, tout, rtol=1e-12, atol=1e-12)
    assert abs(loss - tl) <= 1e-12 * abs(tl)
    # padded input channels: zero gradient in the oracle, unused in torch
    assert np.allclose(g, tg, rtol=1e-9, atol=1e-12)


def test_oracle_fp64_gradient_by_finite_differences():
    rng = np.random.default_rng(6)
    net = R.UNetRef(2, 4, 2, emulate=False)
    p = R.init_params(2, 4, 2, 7).astype(np.float64)
    x = rng.normal(size=(1, 4, 4, 2))
    t = rng.normal(size=(1, 4, 4))
    _, _, g = net.loss_and_grad(x, t, None, p)
    idx = rng.choice(p.size, 40, replace=False)
    idx = np.concatenate([idx, [net.lay.head_b, net.lay.head_w]])
    eps = 1e-6
    for i in idx:
        if g[i] == 0.0 and p[i] == 0.0 and (i % 8) >= 2:
            continue
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (net.loss_and_grad(x, t, None, pp)[0] - net.loss_and_grad(x, t, None, pm)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_emulated_forward_rounds_where_the_device_rounds():
    # every stored activation is a BF16 value; the emulated result stays
    # close to the FP64 one
    rng = np.random.default_rng(8)
    p = R.init_params(4, 8, 3, 9)
    x = rng.normal(size=(2, 16, 16, 4)).astype(np.float32)
    e = R.UNetRef(4, 8, 3, emulate=True)
    f = R.UNetRef(4, 8, 3, emulate=False)
    oe, ce = e.forward(x, p)
    of, _ = f.forward(x, p)
    for k in ("ea", "eb", "p"):
        for a in ce[k]:
            assert np.array_equal(R.bf16(a), a)
    assert np.linalg.norm(oe - of) <= 0.05 * np.linalg.norm(of)
