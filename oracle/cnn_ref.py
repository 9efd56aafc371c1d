"""CPU oracle of the warm-start U-Net (test infrastructure only: imported
by tests/, never by the product path).

A NumPy restatement of the network liblsdtt3_gpu.so trains and runs
(lsdtt3_amd/csrc/cnn.hip); the reference has no learned component, so its
parity is pinned here by mathematics instead of by reference vectors:
  * convolution against scipy.signal.correlate (an independent routine),
    pooling against a reshape mean;
  * the whole backward pass against central finite differences of the
    FP64 loss (emulate=False);
and the GPU is compared with this oracle with emulate=True, which rounds
exactly where the device rounds to BF16 (inputs, weights, every
convolution's post-ReLU output, pooled outputs, and the masked output
gradient fed to the backward products), accumulating in FP64 where the
device accumulates in FP32.

Parameter layout (include/lsdtt3_gpu.h): per 3x3 layer W[cout][9][cin_pad]
(tap = 3 (dy + 1) + dx + 1) then bias[cout]; per level from the finest the
encoder's two layers, then per level below the coarsest the decoder's two
layers, then the head's weights[base] and bias.
"""
from __future__ import annotations

import numpy as np


def bf16(a):
    """round to the nearest BF16 (ties to even), returned as float32"""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


class Layout:
    def __init__(self, in_channels, base, levels):
        self.in_channels, self.base, self.levels = in_channels, base, levels
        self.n = 0
        self.enc, self.dec = [], []
        for lv in range(levels):
            c = base << lv
            a = self._layer(in_channels if lv == 0 else base << (lv - 1), c)
            b = self._layer(c, c)
            self.enc.append((a, b))
        for lv in range(levels - 1):
            c = base << lv
            a = self._layer((base << (lv + 1)) + c, c)
            b = self._layer(c, c)
            self.dec.append((a, b))
        self.head_w = self.n
        self.n += base
        self.head_b = self.n
        self.n += 1

    def _layer(self, cin, cout):
        cp = (cin + 7) // 8 * 8
        L = {"cin": cin, "cin_pad": cp, "cout": cout, "woff": self.n}
        self.n += cout * 9 * cp
        L["boff"] = self.n
        self.n += cout
        return L

    def weights(self, p, L):
        return p[L["woff"]:L["woff"] + L["cout"] * 9 * L["cin_pad"]].reshape(L["cout"], 9, L["cin_pad"])

    def bias(self, p, L):
        return p[L["boff"]:L["boff"] + L["cout"]]


def conv3x3(x, w):
    """y[n, y, x, co] = sum_t,ci x[n, y + dy, x + dx, ci] w[co, t, ci], zero
    padding; FP64"""
    N, H, W, _ = x.shape
    xp = np.pad(x.astype(np.float64), ((0, 0), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((N, H, W, w.shape[0]))
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        y += xp[:, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W, :] @ w[:, t, :].astype(np.float64).T
    return y


def conv3x3_wgrad(g, x):
    """dW[co, t, ci] = sum_p g[p, co] x[p + off(t), ci]"""
    N, H, W, C = x.shape
    xp = np.pad(x.astype(np.float64), ((0, 0), (1, 1), (1, 1), (0, 0)))
    g2 = g.reshape(-1, g.shape[-1]).astype(np.float64)
    dw = np.zeros((g.shape[-1], 9, C))
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        dw[:, t, :] = g2.T @ xp[:, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W, :].reshape(-1, C)
    return dw


def conv3x3_xgrad(g, w, H, W):
    """gradient at the input of conv3x3 (the transpose)"""
    N = g.shape[0]
    gp = np.zeros((N, H + 2, W + 2, w.shape[2]))
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        gp[:, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W, :] += g.astype(np.float64) @ w[:, t, :].astype(np.float64)
    return gp[:, 1:-1, 1:-1, :]


def pool2(x):
    """2x2 average, as the device sums it: ((a + b) + (c + d)) * 0.25"""
    a, b = x[:, 0::2, 0::2], x[:, 0::2, 1::2]
    c, d = x[:, 1::2, 0::2], x[:, 1::2, 1::2]
    return ((a + b) + (c + d)) * 0.25


def up2(x):
    return x.repeat(2, axis=1).repeat(2, axis=2)


class UNetRef:
    """forward / backward of the U-Net for a parameter vector p"""

    def __init__(self, in_channels, base, levels, emulate=True):
        self.lay = Layout(in_channels, base, levels)
        self.emulate = emulate

    def _r(self, a):
        return bf16(a.astype(np.float32)) if self.emulate else a.astype(np.float64)

    @staticmethod
    def _pad(x, L):
        if x.shape[-1] == L["cin_pad"]:
            return x
        return np.concatenate([x, np.zeros(x.shape[:-1] + (L["cin_pad"] - x.shape[-1],), x.dtype)], axis=-1)

    def _conv(self, x, L, p):
        x = self._pad(x, L)
        w = self.lay.weights(p, L)
        if self.emulate:
            w = bf16(w)
        y = conv3x3(x, w) + self.lay.bias(p, L).astype(np.float64)
        y = np.maximum(y, 0.0)
        return self._r(y), w

    def forward(self, x, p):
        lay, L = self.lay, self.lay.levels
        N, H, W, C = x.shape
        x0 = np.zeros((N, H, W, lay.enc[0][0]["cin_pad"]), np.float64)
        x0[..., :C] = x
        x0 = self._r(x0)
        cache = {"x0": x0, "ea": [], "eb": [], "p": [], "u": [], "da": [None] * (L - 1), "db": [None] * (L - 1)}
        inp = x0
        for lv in range(L):
            ea, _ = self._conv(inp, lay.enc[lv][0], p)
            eb, _ = self._conv(ea, lay.enc[lv][1], p)
            cache["ea"].append(ea)
            cache["eb"].append(eb)
            if lv < L - 1:
                pl = pool2(eb.astype(np.float32) if self.emulate else eb)
                pl = self._r(pl)
                cache["p"].append(pl)
                inp = pl
        cache["u"] = [None] * (L - 1)
        coarse = cache["eb"][L - 1]
        for lv in range(L - 2, -1, -1):
            u = np.concatenate([up2(coarse), cache["eb"][lv]], axis=-1)
            cache["u"][lv] = u
            da, _ = self._conv(u, lay.dec[lv][0], p)
            db, _ = self._conv(da, lay.dec[lv][1], p)
            cache["da"][lv], cache["db"][lv] = da, db
            coarse = db
        hw = p[lay.head_w:lay.head_w + lay.base].astype(np.float64)
        out = cache["db"][0].astype(np.float64) @ hw + float(p[lay.head_b])
        cache["out"] = out
        return out, cache

    def loss_and_grad(self, x, target, mask, p):
        """MSE over mask != 0 and its gradient with respect to p (FP64
        vector in p's layout)"""
        lay, L = self.lay, self.lay.levels
        out, c = self.forward(x, p)
        m = np.ones(out.shape, bool) if mask is None else (np.asarray(mask) != 0)
        err = np.where(m, out - target, 0.0)
        cnt = m.sum()
        loss = float((err ** 2).sum() / cnt) if cnt else 0.0
        d = 2.0 * err / cnt if cnt else np.zeros_like(err)
        if self.emulate:
            d = d.astype(np.float32)
        g = np.zeros(lay.n)
        hw = p[lay.head_w:lay.head_w + lay.base].astype(np.float64)
        g[lay.head_w:lay.head_w + lay.base] = np.einsum("nhw,nhwc->c", d.astype(np.float64),
                                                        c["db"][0].astype(np.float64))
        g[lay.head_b] = d.sum()
        gy = d.astype(np.float64)[..., None] * hw

        def conv_bwd(Lr, xin, yout, gout, need_in):
            gm = np.where(yout > 0, gout, 0.0)
            if self.emulate:
                gm = bf16(gm.astype(np.float32))
            xin = self._pad(xin, Lr)
            g[Lr["woff"]:Lr["woff"] + Lr["cout"] * 9 * Lr["cin_pad"]] += conv3x3_wgrad(gm, xin).ravel()
            g[Lr["boff"]:Lr["boff"] + Lr["cout"]] += gm.reshape(-1, Lr["cout"]).sum(axis=0)
            if not need_in:
                return None
            w = lay.weights(p, Lr)
            if self.emulate:
                w = bf16(w)
            return conv3x3_xgrad(gm, w, xin.shape[1], xin.shape[2])[..., :Lr["cin"]]

        gskip = [None] * (L - 1)
        for lv in range(L - 1):
            ga = conv_bwd(lay.dec[lv][1], c["da"][lv], c["db"][lv], gy, True)
            gu = conv_bwd(lay.dec[lv][0], c["u"][lv], c["da"][lv], ga, True)
            cs = lay.base << (lv + 1)
            gskip[lv] = gu[..., cs:]
            gc = gu[..., :cs]
            gy = (gc[:, 0::2, 0::2] + gc[:, 0::2, 1::2]) + (gc[:, 1::2, 0::2] + gc[:, 1::2, 1::2])
        gcur = gy
        for lv in range(L - 1, -1, -1):
            ga = conv_bwd(lay.enc[lv][1], c["ea"][lv], c["eb"][lv], gcur, True)
            if lv == 0:
                conv_bwd(lay.enc[0][0], c["x0"], c["ea"][0], ga, False)
            else:
                gp = conv_bwd(lay.enc[lv][0], c["p"][lv - 1], c["ea"][lv], ga, True)
                gcur = gskip[lv - 1] + 0.25 * up2(gp)
        return loss, out, g


def init_params(in_channels, base, levels, seed):
    """He-normal weights, zero biases (any initialisation serves the tests;
    the device draws its own)"""
    lay = Layout(in_channels, base, levels)
    rng = np.random.default_rng(seed)
    p = np.zeros(lay.n, np.float32)
    for pair in lay.enc + lay.dec:
        for L in pair:
            w = rng.normal(0.0, np.sqrt(2.0 / (9 * L["cin"])), (L["cout"], 9, L["cin"]))
            W = np.zeros((L["cout"], 9, L["cin_pad"]), np.float32)
            W[..., :L["cin"]] = w
            p[L["woff"]:L["woff"] + W.size] = W.ravel()
    p[lay.head_w:lay.head_w + base] = rng.normal(0.0, np.sqrt(1.0 / base), base)
    return p
