"""Learned Gauss-Newton start: a network maps measured voltages to an initial
conductivity, Gauss-Newton refines it, and plain Gauss-Newton from the homogeneous start
remains the fallback.  No counterpart in the reference (its Solver always starts from
gamma = 1, src/solver/solver.cpp:55-62).

* Network: ``Mlp`` over the C ABI (mpf_mlp_*): dense layers with GELU, MSE, Adam; every
  GEMM on MFMA with BF16 operands and FP32 accumulation.
* Features: f(U) = sign(U) log1p(|U| / U0) per measurement (U0 = median |U| of the
  training set), then standardised per feature with training-set mean and deviation.
* Target: ln gamma per element (conductivity relative to the reference value).
* Training data come from the device forward solver itself (``generate``).
* ``LearnedReconstruction``: network start + ``gn_steps`` Gauss-Newton steps; the start
  is accepted only if the final forward-model residual is no worse than the noise-level
  bound (discrepancy principle, tau * ||noise||) or than the residual plain Gauss-Newton
  reaches from the homogeneous start; otherwise the frame is reconstructed by plain GN
  (one frame at a time on ``fallback_eit`` when given).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import capi
from .capi import check, lib
from .eit import Context, SolveOptions, Solver

__all__ = ["Mlp", "VoltageFeatures", "generate", "train", "LearnedReconstruction"]


class Mlp:
    def __init__(self, dims, seed: int = 1, ctx: Context | None = None):
        import torch
        self.torch = torch
        self.ctx = ctx or Context()
        self.dims = [int(d) for d in dims]
        arr = (C.c_int32 * len(self.dims))(*self.dims)
        h = C.c_void_p()
        check(lib().mpf_mlp_create(self.ctx.h, len(self.dims) - 1, arr, C.c_uint64(seed), C.byref(h)))
        self.h = h
        self.n_params = int(lib().mpf_mlp_param_count(h))
        self._x = None

    def forward(self, X, out=None):
        """Y = net(X) for X (batch x dims[0]) float32 cuda; the activations are kept for
        ``backward``, which needs X unchanged."""
        torch = self.torch
        X = X.to(torch.float32).contiguous()
        Y = out if out is not None else torch.empty(X.shape[0], self.dims[-1], dtype=torch.float32, device=X.device)
        check(lib().mpf_mlp_fwd(self.h, X.shape[0], C.c_void_p(X.data_ptr()), C.c_void_p(Y.data_ptr())))
        self._x = X
        return Y

    def backward(self, T) -> float:
        T = T.to(self.torch.float32).contiguous()
        loss = C.c_double()
        check(lib().mpf_mlp_bwd(self.h, C.c_void_p(T.data_ptr()), C.byref(loss)))
        return loss.value

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        check(lib().mpf_mlp_adam(self.h, lr, beta1, beta2, eps))

    def train_step(self, X, T, lr) -> float:
        self.forward(X)
        loss = self.backward(T)
        self.adam(lr)
        return loss

    def params(self, grads=False) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        check(lib().mpf_mlp_params(self.h, 1 if grads else 0, 0, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, np.float32)
        assert flat.size == self.n_params
        check(lib().mpf_mlp_params(self.h, 0, 1, flat.ctypes.data_as(C.POINTER(C.c_float))))

    def close(self):
        if getattr(self, "h", None):
            lib().mpf_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class VoltageFeatures:
    """f(U) = sign(U) log1p(|U| / U0), standardised per feature."""

    def fit(self, U: np.ndarray):
        self.U0 = float(np.median(np.abs(U[U != 0]))) if np.any(U != 0) else 1.0
        f = self._raw(U)
        self.mean = f.mean(axis=0)
        self.std = f.std(axis=0) + 1e-12
        return self

    def _raw(self, U):
        return np.sign(U) * np.log1p(np.abs(U) / self.U0)

    def __call__(self, U: np.ndarray) -> np.ndarray:
        return ((self._raw(U) - self.mean) / self.std).astype(np.float32)


def generate(eit, n: int, seed0: int, phantom, noise_db: float = 60.0, options: SolveOptions | None = None,
             rng: np.random.Generator | None = None, log=None):
    """n (voltage, ln gamma) pairs from the device forward solver: ``phantom(seed)`` gives
    gamma per element; ``eit.frames`` phantoms are solved per launch.  Noise: U + N(0, s^2)
    with s = rms(U) 10^(-dB/20), drawn from ``rng`` (default: seeded by seed0)."""
    import torch
    rng = rng or np.random.default_rng(seed0)
    F = eit.frames
    Us, Gs = [], []
    opts = options or SolveOptions(stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-6)
    for b0 in range(0, n, F):
        k = min(F, n - b0)
        g = np.stack([phantom(seed0 + b0 + j) for j in range(k)] +
                     [np.ones(eit.E)] * (F - k))
        eit.solve(torch.tensor(g, dtype=torch.float32, device=eit._dev()), opts)
        U = eit.result64.cpu().numpy()[:k]
        s = np.sqrt(np.mean(U ** 2, axis=1, keepdims=True)) * 10.0 ** (-noise_db / 20.0)
        Us.append(U + s * rng.standard_normal(U.shape))
        Gs.append(np.log(g[:k]).astype(np.float32))
        if log and (b0 // F) % 8 == 7:
            log("generated %d / %d" % (b0 + k, n))
    return np.concatenate(Us), np.concatenate(Gs)


def train(net: Mlp, X: np.ndarray, T: np.ndarray, epochs: int, batch: int = 256, lr: float = 1e-3,
          seed: int = 0, lr_decay: float = 1.0, log=None):
    """Adam on mini-batches of (X, T) (already features / targets); returns the loss per epoch."""
    import torch
    dev = torch.device("cuda", net.ctx.device)
    Xd = torch.tensor(X, dtype=torch.float32, device=dev)
    Td = torch.tensor(T, dtype=torch.float32, device=dev)
    rng = np.random.default_rng(seed)
    n = X.shape[0]
    hist = []
    for ep in range(epochs):
        perm = torch.tensor(rng.permutation(n), device=dev)
        tot, cnt = 0.0, 0
        for b0 in range(0, n - batch + 1, batch):
            idx = perm[b0:b0 + batch]
            tot += net.train_step(Xd[idx].contiguous(), Td[idx].contiguous(), lr)
            cnt += 1
        hist.append(tot / max(cnt, 1))
        lr *= lr_decay
        if log:
            log("epoch %d loss %.5f" % (ep, hist[-1]))
    return hist


class LearnedReconstruction:
    """Network start + Gauss-Newton, with plain Gauss-Newton as fallback."""

    def __init__(self, eit, net: Mlp, features: VoltageFeatures, lam: float, options: SolveOptions,
                 gn_steps: int = 1, plain_steps: int = 3, tau: float = 2.0, gn_refine: int = 4,
                 check_tol: float = 1e-4, fallback_eit=None):
        self.eit, self.net, self.features = eit, net, features
        self.check_tol = check_tol
        self.solver = Solver(eit, lam, gn_refine=gn_refine, options=options)
        self.gn_steps, self.plain_steps, self.tau = gn_steps, plain_steps, tau
        # frames whose network start is rejected are redone one at a time on a one-frame
        # model of the same mesh (else the whole batch is redone)
        self.fallback = Solver(fallback_eit, lam, gn_refine=gn_refine, options=options) if fallback_eit else None

    def _residual(self, Um):
        return float(np.linalg.norm(Um - self.solver.calculation.cpu().numpy(), axis=1).max())

    def _gn(self, Um_t, steps, start=None):
        s = self.solver
        if start is None:
            # Solver's homogeneous start (solver.cpp:55-62): 1, or 0 when logarithmic
            s.material_distribution.fill_(0.0 if self.eit.logarithmic else 1.0)
        else:
            s.material_distribution.copy_(start)
        s.measurement.copy_(Um_t)
        for k in range(steps):
            s.solve_absolute(warm_start=k > 0)
        # residual of the final gamma: one more forward solve, started from the fields of
        # the last step's forward solve (mixed precision; a cold solve otherwise), to
        # 1e-4 relative residual: the test compares residual norms at the noise level
        # (60 dB: 1e-3 of the voltages)
        self.eit.solve(s.material_distribution, s.options.copy(warm_start=1, tol=self.check_tol, refine_steps=0))
        return s.material_distribution.clone()

    def reconstruct(self, Um: np.ndarray, noise_norm: float | None = None, compare_plain: bool = False):
        """Um: frames x m measured voltages (the model's frames).  Returns (gamma, info)."""
        import torch
        dev = self.eit._dev()
        Um_t = torch.tensor(Um, dtype=torch.float64, device=dev)
        X = torch.tensor(self.features(Um), dtype=torch.float32, device=dev)
        ln_sigma = self.net.forward(X)
        # the network predicts ln(sigma / reference); gamma = sigma / reference, or
        # 10 log10(sigma / reference) for a logarithmic model (equation_kernel.cu:35-43)
        if self.eit.logarithmic:
            start = (ln_sigma * (10.0 / math.log(10.0))).to(torch.float32)
        else:
            start = torch.exp(ln_sigma).to(torch.float32)
        g = self._gn(Um_t, self.gn_steps, start)
        r_learned = np.linalg.norm(Um - self.eit.result64.cpu().numpy(), axis=1)
        info = {"residual_learned": r_learned}
        ok = np.ones(Um.shape[0], bool)
        if noise_norm is not None:
            ok = r_learned <= self.tau * noise_norm
        if compare_plain or (not ok.all() and self.fallback is None):
            gp = self._gn(Um_t, self.plain_steps)
            r_plain = np.linalg.norm(Um - self.eit.result64.cpu().numpy(), axis=1)
            info["residual_plain"] = r_plain
            ok = ok & (r_learned <= r_plain) if compare_plain else ok
            for f in np.nonzero(~ok)[0]:
                g[f] = gp[f]
        elif not ok.all():
            fb = self.fallback
            for f in np.nonzero(~ok)[0]:
                fb.material_distribution.fill_(0.0 if fb.eit.logarithmic else 1.0)
                fb.measurement[0].copy_(Um_t[f])
                for k in range(self.plain_steps):
                    fb.solve_absolute(warm_start=k > 0)
                g[f] = fb.material_distribution[0]
        info["accepted"] = ok
        return g, info

    def reconstruct_many(self, Um: np.ndarray, noise_norm: np.ndarray):
        """All frames of ``Um`` (a multiple of the model's frame count), ``noise_norm`` per
        frame: network start + Gauss-Newton per batch of the model's frames; the frames
        whose start is rejected are collected and redone together afterwards, plain
        Gauss-Newton from the homogeneous start in batches of the same model (the last
        batch padded with copies of its first frame).  Returns (gamma, accepted, residual
        of the learned result) for every frame."""
        import torch
        F = self.eit.frames
        n = Um.shape[0]
        if n % F:
            raise ValueError("reconstruct_many: the frame count must be a multiple of the model's frames")
        gam, acc, res = [], [], []
        for b0 in range(0, n, F):
            sl = slice(b0, b0 + F)
            nn = np.asarray(noise_norm[sl], np.float64)
            g, info = self._learned_batch(Um[sl], nn)
            gam.append(g)
            acc.append(info["accepted"])
            res.append(info["residual_learned"])
        gam = torch.cat(gam)
        acc = np.concatenate(acc)
        res = np.concatenate(res)
        rej = np.nonzero(~acc)[0]
        dev = self.eit._dev()
        for b0 in range(0, rej.size, F):
            idx = rej[b0:b0 + F]
            rows = np.concatenate([idx, np.full(F - idx.size, idx[0])])
            gp = self._gn(torch.tensor(Um[rows], dtype=torch.float64, device=dev), self.plain_steps)
            gam[torch.tensor(idx, device=gam.device)] = gp[:idx.size]
        return gam, acc, res

    def _learned_batch(self, Um, noise_norm):
        """One batch of the model's frames: network start, gn_steps Gauss-Newton steps and
        the per-frame acceptance test r_learned <= tau * noise_norm (no fallback)."""
        import torch
        dev = self.eit._dev()
        Um_t = torch.tensor(Um, dtype=torch.float64, device=dev)
        X = torch.tensor(self.features(Um), dtype=torch.float32, device=dev)
        ln_sigma = self.net.forward(X)
        if self.eit.logarithmic:
            start = (ln_sigma * (10.0 / math.log(10.0))).to(torch.float32)
        else:
            start = torch.exp(ln_sigma).to(torch.float32)
        g = self._gn(Um_t, self.gn_steps, start)
        r = np.linalg.norm(Um - self.eit.result64.cpu().numpy(), axis=1)
        return g, {"residual_learned": r, "accepted": r <= self.tau * noise_norm}
