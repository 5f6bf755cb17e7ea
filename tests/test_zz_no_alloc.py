"""mpf_normal_solve_noalloc (the allocation-free Tikhonov normal-equation solve) against
mpf_normal_solve on the same Jacobian and residual: two calls in a row, the second with a
smaller problem (the fixed work space keeps the first call's contents), and the size
limits refused before any device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(rows, cols, seed):
    rng = np.random.default_rng(seed)
    # a Jacobian-like matrix: smooth decay over the columns, rows of mixed scale
    J = rng.standard_normal((rows, cols)) * np.exp(-np.linspace(0, 4, cols))[None, :] * 1e-3
    r = rng.standard_normal(rows) * 1e-4
    return J, r


def _solve(ctx, entry, J, r, shift, refine):
    import torch
    from mpflow_amd import capi
    from mpflow_amd.capi import lib
    rows, cols = J.shape
    Jd = torch.tensor(np.asfortranarray(J).ravel(order="F"), dtype=torch.float32, device="cuda")
    rd = torch.tensor(r, dtype=torch.float64, device="cuda")
    dd = torch.zeros(cols, dtype=torch.float64, device="cuda")
    f = getattr(lib(), entry)
    capi.check(f(ctx.h, rows, cols, C.c_void_p(Jd.data_ptr()), rows, shift, C.c_void_p(rd.data_ptr()), refine,
                 C.c_void_p(dd.data_ptr())))
    return dd.cpu().numpy()


def test_noalloc_matches_allocating_solve(gpu_ctx):
    from oracle.eit_oracle import gn_step_exact
    for rows, cols, seed in [(1024, 40344, 1), (256, 9600, 2)]:
        print("case: mpf_normal_solve_noalloc rows=%d cols=%d" % (rows, cols), flush=True)
        J, r = _problem(rows, cols, seed)
        # a shift well above the BF16 Gram's rounding, so that the FP64-residual passes
        # converge to the FP64 solution of the FP32 Jacobian's system
        shift = 1e-2 * float(np.mean(np.sum(J * J, axis=1)))
        d_alloc = _solve(gpu_ctx, "mpf_normal_solve", J, r, shift, 4)
        d_static = _solve(gpu_ctx, "mpf_normal_solve_noalloc", J, r, shift, 4)
        # the FP64 oracle of the same system with the FP32-rounded Jacobian
        d64 = gn_step_exact(J.astype(np.float32).astype(np.float64), r, shift)
        e_pair = np.linalg.norm(d_static - d_alloc) / np.linalg.norm(d_alloc)
        e_64 = np.linalg.norm(d_static - d64) / np.linalg.norm(d64)
        print("noalloc vs allocating %.3e, vs FP64 %.3e" % (e_pair, e_64))
        assert np.all(np.isfinite(d_static))
        assert e_pair < 1e-9
        assert e_64 < 1e-6


def test_noalloc_refuses_oversized(gpu_ctx):
    from mpflow_amd import capi
    from mpflow_amd.capi import lib
    L = lib()
    dummy = C.c_void_p(8)   # never dereferenced: the shape check comes first
    for rows, cols, ldj in [(1040, 100, 1040), (16, 70000, 16), (18, 100, 18)]:
        rc = L.mpf_normal_solve_noalloc(gpu_ctx.h, rows, cols, dummy, ldj, 1e-6, dummy, 1, dummy)
        assert rc == capi.MPF_EINVAL
        assert b"mpf_normal_solve_noalloc" in L.mpf_last_error()
