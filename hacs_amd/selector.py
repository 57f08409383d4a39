"""The learned kernel selector: a 2-layer MLP cost model trained on the autotuner's timings.

The network (64 hidden units, FP32) runs entirely in the library's own HIP kernels
(hx_mlp_forward / hx_mlp_backward / hx_adam_step); torch tensors only hold its
parameters in device memory.  It predicts log-time per (variant, tile) choice from job
features; the cheapest predicted choice is taken.  Every choice computes the same bytes,
so a misprediction only costs time.  The exhaustive autotuner (autotune.py) remains the
fallback (unknown features, an untrained model) and the reference it is scored against (it
produced the labels).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import check, lib


def _p(t: torch.Tensor):
    return C.c_void_p(t.data_ptr())


def _stream():
    return torch.cuda.current_stream().cuda_stream


class MLP:
    """hidden = relu(X W1^T + b1), y = hidden W2^T + b2, all FP32 on the device."""

    def __init__(self, n_in: int, n_hidden: int = 64, n_out: int = 1, seed: int = 0):
        g = np.random.default_rng(seed)
        lim1 = np.sqrt(6.0 / (n_in + n_hidden))
        lim2 = np.sqrt(6.0 / (n_hidden + n_out))
        dev = "cuda"
        self.F, self.H, self.O = n_in, n_hidden, n_out
        self.W1 = torch.as_tensor(g.uniform(-lim1, lim1, (n_hidden, n_in)), dtype=torch.float32, device=dev)
        self.b1 = torch.zeros(n_hidden, dtype=torch.float32, device=dev)
        self.W2 = torch.as_tensor(g.uniform(-lim2, lim2, (n_out, n_hidden)), dtype=torch.float32, device=dev)
        self.b2 = torch.zeros(n_out, dtype=torch.float32, device=dev)
        self.m1 = [torch.zeros_like(t) for t in self.params()]
        self.m2 = [torch.zeros_like(t) for t in self.params()]
        self.t = 0

    def params(self):
        return [self.W1, self.b1, self.W2, self.b2]

    def forward(self, X: torch.Tensor):
        X = X.contiguous().float()
        B = X.shape[0]
        hidden = torch.empty((B, self.H), dtype=torch.float32, device="cuda")
        y = torch.empty((B, self.O), dtype=torch.float32, device="cuda")
        check(lib().hx_mlp_forward(_p(X), B, self.F, _p(self.W1), _p(self.b1), self.H, _p(self.W2), _p(self.b2),
                                   self.O, _p(hidden), _p(y), _stream()))
        return hidden, y

    def backward(self, X: torch.Tensor, T: torch.Tensor, mask: torch.Tensor | None = None):
        X = X.contiguous().float()
        T = T.contiguous().float()
        hidden, y = self.forward(X)
        grads = [torch.empty_like(t) for t in self.params()]
        loss = torch.empty(1, dtype=torch.float32, device="cuda")
        mk = None if mask is None else _p(mask.contiguous().float())
        check(lib().hx_mlp_backward(_p(X), X.shape[0], self.F, _p(self.W1), self.H, _p(self.W2), self.O, _p(hidden),
                                    _p(y), _p(T), mk, _p(grads[0]), _p(grads[1]), _p(grads[2]), _p(grads[3]),
                                    _p(loss), _stream()))
        return float(loss.item()), grads

    def train_step(self, X, T, mask=None, lr: float = 1e-3) -> float:
        loss, grads = self.backward(X, T, mask)
        self.t += 1
        for w, g, a, b in zip(self.params(), grads, self.m1, self.m2):
            check(lib().hx_adam_step(_p(w), _p(g), _p(a), _p(b), w.numel(), lr, 0.9, 0.999, 1e-8, self.t, _stream()))
        return loss

    def predict(self, X: torch.Tensor) -> torch.Tensor:
        return self.forward(X)[1]
