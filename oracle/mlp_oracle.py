"""FP64 restatement of the learned-start network (test infrastructure only).

The device network (mpflow_amd/csrc/mlp.hip) has no counterpart in the reference; this
module is the checker for it: dense layers with exact (erf) GELU between them and a
linear output, MSE loss, Adam.  With ``bf16=True`` every GEMM operand is rounded to
bfloat16 (round to nearest even) exactly as the device stages them, and the products
are accumulated in FP64, so device and oracle differ only by FP32 accumulation order.
Parity of the backward pass itself is pinned by central finite differences
(tests/test_mlp_oracle.py).
"""
from __future__ import annotations

import math

import numpy as np

__all__ = ["bf16_round", "gelu", "gelu_grad", "MlpOracle"]


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round float32 values to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    nan = np.isnan(a)
    if nan.any():
        out[nan] = np.nan
    return out


_erf = np.vectorize(math.erf)


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + _erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


class MlpOracle:
    """Parameters in the device's flat order: W_0 (in x out, row-major), b_0, W_1, b_1, ..."""

    def __init__(self, dims, flat):
        self.dims = list(dims)
        self.W, self.b = [], []
        off = 0
        flat = np.asarray(flat, dtype=np.float64)
        for l in range(len(dims) - 1):
            n = dims[l] * dims[l + 1]
            self.W.append(flat[off:off + n].reshape(dims[l], dims[l + 1]).copy())
            off += n
            self.b.append(flat[off:off + dims[l + 1]].copy())
            off += dims[l + 1]

    def flat(self):
        return np.concatenate([np.concatenate([w.ravel(), b]) for w, b in zip(self.W, self.b)])

    def _mm(self, a, b, bf16):
        if bf16:
            a, b = bf16_round(a), bf16_round(b)
        return a @ b

    def forward(self, X, bf16=False):
        """Returns Y and the cache (inputs of each layer, pre-activations)."""
        h = np.asarray(X, dtype=np.float64)
        ins, zs = [], []
        L = len(self.W)
        for l in range(L):
            ins.append(h)
            z = self._mm(h, self.W[l], bf16) + self.b[l]
            zs.append(z)
            h = z if l == L - 1 else gelu(z)
            if bf16 and l < L - 1:
                h = h.astype(np.float32).astype(np.float64)   # the device stores H in FP32
        return h, (ins, zs)

    def loss_and_grads(self, X, T, bf16=False):
        Y, (ins, zs) = self.forward(X, bf16)
        T = np.asarray(T, dtype=np.float64)
        n = Y.size
        loss = float(np.sum((Y - T) ** 2) / n)
        dz = 2.0 * (Y - T) / n
        gW = [None] * len(self.W)
        gb = [None] * len(self.W)
        for l in range(len(self.W) - 1, -1, -1):
            gW[l] = self._mm(ins[l].T, dz, bf16)
            gb[l] = dz.sum(axis=0)
            if l > 0:
                dh = self._mm(dz, self.W[l].T, bf16)
                dz = dh * gelu_grad(zs[l - 1])
        return loss, gW, gb

    def grads_flat(self, X, T, bf16=False):
        loss, gW, gb = self.loss_and_grads(X, T, bf16)
        return loss, np.concatenate([np.concatenate([w.ravel(), b]) for w, b in zip(gW, gb)])


def adam_step(w, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015), bias correction folded into the step size as on the
    device.  Returns (w, m, v)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    lr_t = lr * math.sqrt(1.0 - b2 ** t) / (1.0 - b1 ** t)
    return w - lr_t * m / (np.sqrt(v) + eps), m, v
