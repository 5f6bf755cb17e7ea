"""Device parity of the dense kernels (Gram on MFMA, FP64 GEMV, Cholesky) and of the
Gauss-Newton step (Inverse::updateJacobian / calcExcitation / solve,
src/solver/inverse.cpp:60-238) against FP64 references."""
import ctypes as C

import numpy as np
import pytest

from mpflow_amd.capi import check, lib
from mpflow_amd.phantom import add_noise, ellipse_phantom, frame_seed, inclusion_phantom
from oracle.eit_oracle import gn_step_exact
from tests.problems import eit_of, make_problem, oracle_of

pytestmark = pytest.mark.gpu


def _bf16(a):
    import torch
    return torch.tensor(a, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()


@pytest.mark.parametrize("trans,rows,k", [(0, 32, 16), (0, 256, 1000), (0, 100, 333), (0, 102, 77),
                                          (0, 1024, 4003), (1, 200, 130)])
def test_gram_bf16_mfma(gpu_ctx, trans, rows, k):
    """MFMA lane layout and tiling: G equals the FP64 Gram of the BF16-rounded operand."""
    import torch
    rng = np.random.default_rng(rows + k)
    A = rng.standard_normal((rows, k)).astype(np.float32)
    At = torch.tensor(A.T.copy()).cuda()          # column-major rows x k == row-major k x rows
    n = k if trans else rows
    G = torch.empty(n, n, dtype=torch.float32, device="cuda")
    print("case: gram trans=%d rows=%d k=%d" % (trans, rows, k), flush=True)
    check(lib().mpf_gram_bf16(gpu_ctx.h, trans, rows, k, C.c_void_p(At.data_ptr()), rows,
                              C.c_void_p(G.data_ptr())))
    Ab = _bf16(A)
    ref = Ab.T @ Ab if trans else Ab @ Ab.T
    Gh = G.cpu().numpy().astype(np.float64)        # column-major n x n, symmetric
    err = np.max(np.abs(Gh - ref)) / np.max(np.abs(ref))
    assert err < 1e-5, err
    assert np.array_equal(Gh, Gh.T)


def test_spd_solve(gpu_ctx):
    import torch
    rng = np.random.default_rng(7)
    n = 300
    X = rng.standard_normal((n, 2 * n))
    G = (X @ X.T / n).astype(np.float32)
    b = rng.standard_normal(n)
    Gd = torch.tensor(G.T.copy()).cuda()
    bd = torch.tensor(b).cuda()
    yd = torch.empty(n, dtype=torch.float64, device="cuda")
    check(lib().mpf_spd_solve(gpu_ctx.h, n, C.c_void_p(Gd.data_ptr()), 0.5, C.c_void_p(bd.data_ptr()),
                              C.c_void_p(yd.data_ptr())))
    ref = np.linalg.solve(G.astype(np.float64) + 0.5 * np.eye(n), b)
    assert np.max(np.abs(yd.cpu().numpy() - ref)) / np.max(np.abs(ref)) < 1e-10


def test_gemv_t(gpu_ctx):
    import torch
    rng = np.random.default_rng(3)
    rows, cols = 1024, 777
    A = rng.standard_normal((rows, cols)).astype(np.float32)
    x = rng.standard_normal(rows)
    Ad = torch.tensor(A.T.copy()).cuda()
    xd = torch.tensor(x).cuda()
    yd = torch.empty(cols, dtype=torch.float64, device="cuda")
    check(lib().mpf_gemv_t(gpu_ctx.h, rows, cols, C.c_void_p(Ad.data_ptr()), rows, C.c_void_p(xd.data_ptr()),
                           C.c_void_p(yd.data_ptr())))
    ref = A.astype(np.float64).T @ x
    assert np.max(np.abs(yd.cpu().numpy() - ref)) < 1e-9 * np.max(np.abs(ref)) * 10


def test_gn_step_matches_oracle(gpu_ctx):
    """One Solver::solveAbsolute from gamma = 1: the device step (BF16 MFMA Gram + FP64
    Cholesky + FP64 refinement) equals the FP64 data-space step within 1e-3 rel L2."""
    import torch
    from mpflow_amd.eit import Solver
    p = make_problem(16, 16, 0.1, C=2, height=0.4, port_height=0.1)
    o = oracle_of(p)
    gt = inclusion_phantom(p["nodes"], p["elements"], frame_seed(1, 9))
    Um, _, _ = o.solve(gt)
    Um = add_noise(Um, 60.0, frame_seed(1, 9))
    g0 = np.ones(p["elements"].shape[0])
    U0, phis, _ = o.solve(g0)
    J = o.jacobian(g0, phis)
    lam = 1e-3
    d_ref = gn_step_exact(J, o.flatten_voltages(Um) - o.flatten_voltages(U0), lam * lam)
    print("case: gn step", flush=True)
    eit = eit_of(p, gpu_ctx)
    s = Solver(eit, lam, gn_refine=3)
    s.measurement[0] = torch.tensor(o.flatten_voltages(Um))
    s.solve_absolute()
    d_gpu = s.material_distribution.cpu().numpy()[0].astype(np.float64) - 1.0
    err = np.linalg.norm(d_gpu - d_ref) / np.linalg.norm(d_ref)
    print("gn step rel err", err)
    assert err < 1e-3


@pytest.mark.parametrize("m,E,ld", [(96, 700, 96), (130, 400, 131)])
def test_normal_solve(gpu_ctx, m, E, ld):
    """mpf_normal_solve: the Tikhonov step J^T (J J^T + shift I)^-1 r with FP64 refinement
    reaches the FP64 solution for the FP32 Jacobian (the BF16 Gram only preconditions)."""
    import torch
    rng = np.random.default_rng(m + E)
    J = rng.standard_normal((m, E)).astype(np.float32)
    Jd = torch.zeros(E, ld, dtype=torch.float32, device="cuda")      # column-major m x E, ld
    Jd[:, :m] = torch.tensor(J.T)
    r = rng.standard_normal(m)
    rd = torch.tensor(r, device="cuda")
    out = torch.empty(E, dtype=torch.float64, device="cuda")
    shift = 0.5
    print("case: normal solve m=%d E=%d ld=%d" % (m, E, ld), flush=True)
    check(lib().mpf_normal_solve(gpu_ctx.h, m, E, C.c_void_p(Jd.data_ptr()), ld, shift, C.c_void_p(rd.data_ptr()),
                                 4, C.c_void_p(out.data_ptr())))
    J64 = J.astype(np.float64)
    want = J64.T @ np.linalg.solve(J64 @ J64.T + shift * np.eye(m), r)
    err = np.linalg.norm(out.cpu().numpy() - want) / np.linalg.norm(want)
    print("rel err", err)
    assert err < 1e-9
