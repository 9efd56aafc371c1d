"""Learned hillslope warm start (no counterpart in the reference): the
BF16 U-Net of liblsdtt3_gpu.so, trained on pairs captured from this
package's own LEM steps, and the LEM hook that starts each step's first
Newton solve from its prediction (include/lsdtt3_gpu.h, lsdtt_gpu_cnn_*).

The reference starts the hillslope Newton solve from z^n
(src/cpu/lsdtt_hillslope_newton_cpu.cpp:196-260); with a warm start the
predicted point is kept only when it lowers ||F||inf, and Newton still
converges to the same relative residual, so the result meets the same
tolerance as without it.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

WARM_NONE, WARM_EXPLICIT, WARM_LEARNED, WARM_PREVIOUS = 0, 1, 2, 3


class UNet:
    """U-Net with ``levels`` levels of ``base_channels`` doubling per level.

    Parameters are one FP32 vector (layout in lsdtt3_gpu.h); ``forward`` and
    ``backward`` take host arrays (NHWC input, [n, h, w] output)."""

    def __init__(self, in_channels=4, base_channels=32, levels=4, seed=0, tile=256, halo=16):
        self.in_channels, self.base_channels, self.levels = int(in_channels), int(base_channels), int(levels)
        self.tile, self.halo = int(tile), int(halo)
        h = N.lib().lsdtt_gpu_cnn_create(self.in_channels, self.base_channels, self.levels, int(seed), self.tile,
                                         self.halo)
        if not h:
            raise N.GpuError(N.ERR_ARG, (N.lib().lsdtt_gpu_last_error() or b"").decode())
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            N.lib().lsdtt_gpu_cnn_destroy(self._h)
            self._h = None

    @property
    def num_params(self) -> int:
        return int(N.lib().lsdtt_gpu_cnn_num_params(self._h))

    def get_params(self) -> np.ndarray:
        p = np.empty(self.num_params, np.float32)
        N.check(N.lib().lsdtt_gpu_cnn_get_params(self._h, N.ptr(p)))
        return p

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.num_params:
            raise ValueError("parameter vector has the wrong size")
        N.check(N.lib().lsdtt_gpu_cnn_set_params(self._h, N.ptr(p)))

    def get_grads(self) -> np.ndarray:
        g = np.empty(self.num_params, np.float32)
        N.check(N.lib().lsdtt_gpu_cnn_get_grads(self._h, N.ptr(g)))
        return g

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        if x.ndim != 4 or x.shape[3] != self.in_channels:
            raise ValueError("x must be [n, h, w, in_channels]")
        n, hh, ww, _ = x.shape
        y = np.empty((n, hh, ww), np.float32)
        N.check(N.lib().lsdtt_gpu_cnn_forward(self._h, N.ptr(x), n, hh, ww, N.ptr(y)))
        return y

    def backward(self, target, mask=None) -> float:
        t = np.ascontiguousarray(target, np.float32)
        m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        loss = C.c_double()
        N.check(N.lib().lsdtt_gpu_cnn_backward(self._h, N.ptr(t), N.ptr(m), C.byref(loss)))
        return loss.value

    def zero_grad(self) -> None:
        N.check(N.lib().lsdtt_gpu_cnn_zero_grad(self._h))

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8) -> None:
        N.check(N.lib().lsdtt_gpu_cnn_adam_step(self._h, lr, beta1, beta2, eps))

    def benchmark(self, n, h, w, reps=5) -> dict:
        f, c, t, fl = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        N.check(N.lib().lsdtt_gpu_cnn_benchmark(self._h, n, h, w, reps, C.byref(f), C.byref(c), C.byref(t),
                                                C.byref(fl)))
        return {"forward_ms": f.value, "conv_ms": c.value, "train_ms": t.value, "conv_flops": fl.value}

    def train(self, dataset, n_iterations, batch=32, lr=1e-3, seed=0) -> np.ndarray:
        losses = np.zeros(int(n_iterations), np.float32)
        N.check(N.lib().lsdtt_gpu_cnn_train(self._h, dataset._h, int(n_iterations), int(batch), float(lr), int(seed),
                                            N.ptr(losses)))
        return losses

    def evaluate(self, dataset, n_batches=4, batch=32, seed=1):
        """(network MSE, zero-prediction MSE) on random windows"""
        a, b = C.c_double(), C.c_double()
        N.check(N.lib().lsdtt_gpu_cnn_evaluate(self._h, dataset._h, int(n_batches), int(batch), int(seed),
                                               C.byref(a), C.byref(b)))
        return a.value, b.value


def conv3x3_layer(x, w, b, relu=True, gy=None, input_grad=True):
    """one 3x3 layer through the network's kernels: x [n, h, w, cin], w
    [cout, 9, cin], b [cout]; returns y, or (y, gx, gw, gb) with gy"""
    x = np.ascontiguousarray(x, np.float32)
    w = np.ascontiguousarray(w, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    n, hh, ww, cin = x.shape
    cout = w.shape[0]
    if w.shape != (cout, 9, cin) or b.shape != (cout,):
        raise ValueError("w must be [cout, 9, cin] and b [cout]")
    y = np.empty((n, hh, ww, cout), np.float32)
    if gy is None:
        N.check(N.lib().lsdtt_gpu_cnn_conv3x3(N.ptr(x), N.ptr(w), N.ptr(b), n, hh, ww, cin, cout, int(relu), N.ptr(y),
                                              None, None, None, None))
        return y
    gy = np.ascontiguousarray(gy, np.float32)
    gx = np.empty_like(x) if input_grad else None
    gw = np.empty_like(w)
    gb = np.empty_like(b)
    N.check(N.lib().lsdtt_gpu_cnn_conv3x3(N.ptr(x), N.ptr(w), N.ptr(b), n, hh, ww, cin, cout, int(relu), N.ptr(y),
                                          N.ptr(gy), N.ptr(gx), N.ptr(gw), N.ptr(gb)))
    return y, gx, gw, gb


class Dataset:
    """Device-resident (post-incision surface, hillslope change) pairs
    captured from LEM steps on one grid."""

    def __init__(self):
        self._h = N.lib().lsdtt_gpu_cnn_dataset_create()
        if not self._h:
            raise N.GpuError(N.ERR_RUNTIME, (N.lib().lsdtt_gpu_last_error() or b"").decode())

    def __del__(self):
        if getattr(self, "_h", None):
            N.lib().lsdtt_gpu_cnn_dataset_destroy(self._h)
            self._h = None

    def __len__(self) -> int:
        return int(N.lib().lsdtt_gpu_cnn_dataset_size(self._h))

    def add(self, model) -> None:
        """the pair of `model`'s last step (capture must be on)"""
        N.check(N.lib().lsdtt_gpu_cnn_dataset_add(self._h, model._h))


def capture(model, enable=True) -> None:
    """keep each later step's post-incision surface and hillslope change"""
    N.check(N.lib().lsdtt_gpu_lem_capture_hillslope(model._h, int(enable)))


def set_warm_start(model, mode, net=None) -> None:
    """mode WARM_NONE (z^n), WARM_EXPLICIT (z^n - dt F(z^n)), WARM_LEARNED
    (the network `net`, kept alive by the model while set) or WARM_PREVIOUS
    (z^n plus the previous step's hillslope change)."""
    N.check(N.lib().lsdtt_gpu_lem_set_warm_start(model._h, net._h if net is not None else None, int(mode)))
    model._warm_net = net if mode == WARM_LEARNED else None


def generate(models, dataset, n_steps, every=1) -> None:
    """step each LandscapeModel in `models` n_steps times with capture on,
    adding every `every`-th step's pair to `dataset`"""
    for m in models:
        capture(m, True)
    for k in range(int(n_steps)):
        for m in models:
            m.step()
            if (k + 1) % every == 0:
                dataset.add(m)
    for m in models:
        capture(m, False)
