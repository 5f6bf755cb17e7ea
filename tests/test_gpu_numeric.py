"""mpf_pcg_solve / mpf_bicgstab_solve on a matrix handed over in the reference's
SparseMatrix layout (32 slots per row, 0xFFFFFFFF for empty slots) against the FP64
direct solve and the oracle's Krylov restatements."""
import ctypes as C

import numpy as np
import pytest

import oracle.eit_oracle as eo
from mpflow_amd.phantom import inclusion_phantom, frame_seed
from tests.problems import make_problem, oracle_of

pytestmark = pytest.mark.gpu


def _sparse_matrix_layout(A):
    """CSR -> (column ids uint32 [n*32], values float32 [n*32], density) in row order."""
    A = A.tocsr()
    n = A.shape[0]
    cid = np.full((n, 32), 0xFFFFFFFF, np.uint32)
    val = np.zeros((n, 32), np.float32)
    density = 0
    for r in range(n):
        cols = A.indices[A.indptr[r]:A.indptr[r + 1]]
        vals = A.data[A.indptr[r]:A.indptr[r + 1]]
        cid[r, :len(cols)] = cols
        val[r, :len(cols)] = vals
        density = max(density, len(cols))
    return cid.ravel(), val.ravel(), density


def _solve(gpu_ctx, entry, A, B, dc_free, **opt):
    import torch
    from mpflow_amd import capi
    from mpflow_amd.capi import lib
    from mpflow_amd.eit import SolveOptions
    n, R = B.shape
    ld = (n + 15) // 16 * 16
    cid, val, density = _sparse_matrix_layout(A)
    dc = torch.tensor(cid.view(np.int32), device="cuda")
    dv = torch.tensor(val, device="cuda")
    Bd = torch.zeros(R, ld, dtype=torch.float32, device="cuda")
    Bd[:, :n] = torch.tensor(B.T, dtype=torch.float32)
    Xd = torch.zeros(R, ld, dtype=torch.float32, device="cuda")
    st = capi.SolveStats()
    o = SolveOptions(**opt)
    f = getattr(lib(), entry)
    rc = f(gpu_ctx.h, n, density, C.c_void_p(dc.data_ptr()), C.c_void_p(dv.data_ptr()), R,
           C.c_void_p(Bd.data_ptr()), ld, C.c_void_p(Xd.data_ptr()), ld, int(dc_free), C.byref(o.s), C.byref(st))
    capi.check(rc, allow=(capi.MPF_OK, capi.MPF_ENOCONV))
    return Xd[:, :n].cpu().numpy().T.astype(np.float64), st


@pytest.fixture(scope="module")
def sysm():
    p = make_problem(16, 16, 0.1, C=2, height=0.4, port_height=0.1)
    o = oracle_of(p)
    g = inclusion_phantom(p["nodes"], p["elements"], frame_seed(1, 8))
    return o, g


def _rel(X, Y):
    return float(np.max(np.abs(X - Y)) / np.max(np.abs(Y)))


@pytest.mark.parametrize("precision", [0, 1])
def test_pcg_solve_dcfree(gpu_ctx, sysm, precision):
    from mpflow_amd import capi
    o, g = sysm
    A = o.system(g, 0)
    B = o.B @ o.Pi
    X = eo.solve_direct(A, B, True)
    print("case: mpf_pcg_solve dcFree precision=%d" % precision, flush=True)
    Xg, st = _solve(gpu_ctx, "mpf_pcg_solve", A, B, True, stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-6,
                    precision=precision, refine_steps=2, refine_tol=1e-3)
    err = _rel(Xg, X)
    e32 = _rel(Xg, X.astype(np.float32).astype(np.float64))
    print("rel err", err, "vs the FP32-rounded solution", e32, "steps", st.steps)
    assert st.not_converged == 0
    # X is FP32: with refinement it is the FP64 solution rounded (a few ulps of the max)
    assert err < (1e-5 if precision == 0 else 3e-7)


def test_pcg_solve_reference_rule_steps(gpu_ctx, sysm):
    """The reference stop rule ends within a few polls of the restatement's step count."""
    from mpflow_amd import capi
    o, g = sysm
    A = o.system(g, 1)
    B = o.B @ o.Pi
    _, so = eo.pcg_reference(A, B, 1.0 / A.diagonal(), False)
    Xg, st = _solve(gpu_ctx, "mpf_pcg_solve", A, B, False, stop_rule=capi.MPF_STOP_REFERENCE)
    print("steps", st.steps, "oracle", so)
    assert _rel(Xg, eo.solve_direct(A, B, False)) < 1e-5
    assert st.steps <= 2 * so + 16


@pytest.mark.parametrize("ignore_dc", [1, 0])
def test_bicgstab_solve(gpu_ctx, sysm, ignore_dc):
    from mpflow_amd import capi
    o, g = sysm
    B = o.B @ o.Pi
    print("case: mpf_bicgstab_solve ignore_dc=%d" % ignore_dc, flush=True)
    # SPD harmonic: the direct solution itself
    A1 = o.system(g, 1)
    X1, st = _solve(gpu_ctx, "mpf_bicgstab_solve", A1, B, False, stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-6,
                    bicgstab_ignore_dc=ignore_dc)
    e1 = _rel(X1, eo.solve_direct(A1, B, False))
    # singular harmonic: the dcFree solution up to a constant per column
    A0 = o.system(g, 0)
    X0, st0 = _solve(gpu_ctx, "mpf_bicgstab_solve", A0, B, True, stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-6,
                     bicgstab_ignore_dc=ignore_dc)
    Xd = eo.solve_direct(A0, B, True)
    D = X0 - Xd
    e0 = float(np.max(np.abs(D - D.mean(axis=0))) / np.max(np.abs(Xd)))
    print("SPD rel err", e1, "steps", st.steps, "singular rel err (mod constants)", e0, "steps", st0.steps,
          "not converged", st0.not_converged)
    assert e1 < 1e-5
    if ignore_dc:
        # the reference's FP32 iteration on the singular matrix may break down; it must then
        # say so instead of returning a wrong solution
        assert st0.not_converged > 0 or e0 < 1e-5
    else:
        assert st0.not_converged == 0 and e0 < 1e-5
