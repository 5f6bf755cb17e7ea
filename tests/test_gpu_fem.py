"""The standalone FEM seams in the reference's layouts, through the C ABI:
mpf_fem_assemble (Equation::update, src/fem/equation.cpp:261-274), mpf_eit_measure
(EIT::applyMeasurementPattern, src/models/eit.cpp:220-245) and mpf_jacobian
(Equation::calcJacobian, src/fem/equation.cpp:282-310), each against the FP64 oracle, and
EIT::solve recomposed from them and mpf_pcg_solve against mpf_eit_solve."""
import ctypes as C

import numpy as np
import pytest

from mpflow_amd.phantom import frame_seed, inclusion_phantom
from oracle.eit_oracle import mat_param_deriv
from tests.problems import eit_of, make_problem, oracle_of

pytestmark = pytest.mark.gpu

INVALID = 0xFFFFFFFF


@pytest.fixture(scope="module")
def prob():
    p = make_problem(8, 16, 0.1, C=2, height=0.4, port_height=0.1)
    o = oracle_of(p)
    g = inclusion_phantom(p["nodes"], p["elements"], frame_seed(3, 2))
    return p, o, g


def _insertion_order(elements, N):
    """Column order of each row's first insertions over the element loop
    (equation.cpp:131-144), written out independently of the product."""
    rows = [[] for _ in range(N)]
    for el in elements:
        for i in el:
            for j in el:
                if j not in rows[i]:
                    rows[i].append(int(j))
    return rows


@pytest.mark.parametrize("keep", [False, True])
def test_fem_assemble(gpu_ctx, prob, keep):
    import torch
    p, o, g = prob
    print("case: mpf_fem_assemble keep_node_order=%d" % keep, flush=True)
    eit = eit_of(p, gpu_ctx, keep_node_order=keep)
    order = _insertion_order(p["elements"], eit.N)
    gt = torch.tensor(g, dtype=torch.float32)
    for c in range(2):
        alpha, _ = o.alpha_beta(c)
        ids, vals = eit.update(gt, alpha)
        ids = ids.cpu().numpy().view(np.uint32)
        vals = vals.cpu().numpy()
        A = o.system(g.astype(np.float32).astype(np.float64), c).tocsr()
        worst = 0.0
        for r in range(eit.N):
            want = order[r]
            assert list(ids[r, :len(want)]) == want, "row %d slot order" % r
            assert np.all(ids[r, len(want):] == INVALID) and np.all(vals[r, len(want):] == 0.0)
            ref = np.array([A[r, j] for j in want])
            worst = max(worst, float(np.max(np.abs(vals[r, :len(want)] - ref) / np.abs(ref).max())))
        print("harmonic", c, "max row-relative error", worst)
        # FP64 sums rounded once to FP32
        assert worst < 2e-7


def test_eit_measure(gpu_ctx, prob):
    import torch
    p, o, _ = prob
    print("case: mpf_eit_measure", flush=True)
    eit = eit_of(p, gpu_ctx)
    D, M, Mp, N = eit.D, eit.M, eit.Mp, eit.N
    ld = (N + 15) // 16 * 16 + 16
    rng = np.random.default_rng(7)
    phi = rng.standard_normal((D + M, ld)).astype(np.float32)
    U0 = rng.standard_normal((D, Mp)).astype(np.float32)
    dphi = torch.tensor(phi, device="cuda")
    dU = torch.tensor(U0, device="cuda")
    want = (o.W @ phi[:D, :N].T.astype(np.float64)).T            # D x M
    eit.apply_measurement_pattern(dphi, dU, additive=False)
    got = dU.cpu().numpy()
    assert np.all(got[:, M:] == 0.0)
    err = np.max(np.abs(got[:, :M] - want)) / np.max(np.abs(want))
    eit.apply_measurement_pattern(dphi, dU, additive=True)
    err2 = np.max(np.abs(dU.cpu().numpy()[:, :M] - 2 * want)) / np.max(np.abs(want))
    print("rel err", err, "additive", err2)
    assert err < 1e-6 and err2 < 1e-6


def test_jacobian_seam(gpu_ctx, prob):
    import torch
    p, o, g = prob
    print("case: mpf_jacobian", flush=True)
    eit = eit_of(p, gpu_ctx)
    D, M, N, E = eit.D, eit.M, eit.N, eit.E
    ld = (N + 15) // 16 * 16
    rng = np.random.default_rng(11)
    phis = [rng.standard_normal((D + M, ld)).astype(np.float32) for _ in range(2)]
    gt = torch.tensor(g, dtype=torch.float32, device="cuda")
    Jd = torch.full((E, eit.m), 7.0, dtype=torch.float32, device="cuda")
    eit.calc_jacobian(torch.tensor(phis[0], device="cuda"), gt, Jd, additive=False)
    J1 = Jd.cpu().numpy().T
    eit.calc_jacobian(torch.tensor(phis[1], device="cuda"), gt, Jd, additive=True)
    J2 = Jd.cpu().numpy().T
    # independent restatement: sigma'_e sum_ij Phi[n_i, d] K_e[i][j] Phi[n_j, D + m]
    ds = mat_param_deriv(g.astype(np.float32).astype(np.float64), 1.0)
    el = o.elements

    def ref(ph):
        X = ph[:, :N].T.astype(np.float64)
        Pd, Pm = X[el][:, :, :D], X[el][:, :, D:]
        T = np.einsum("eij,ejm->eim", o.Ke, Pm)
        J = np.zeros((eit.Dp, eit.Mp, E))
        J[:D, :M, :] = np.einsum("eid,eim->dme", Pd, T) * ds[None, None, :]
        return J.reshape(eit.Dp * eit.Mp, E)

    R1 = ref(phis[0])
    R2 = R1 + ref(phis[1])
    e1 = np.linalg.norm(J1 - R1) / np.linalg.norm(R1)
    e2 = np.linalg.norm(J2 - R2) / np.linalg.norm(R2)
    # the oracle's Jacobian is this seam summed over harmonics and scaled by -height
    Jo = o.jacobian(g.astype(np.float32).astype(np.float64), [phis[0][:, :N].T.astype(np.float64)], quirks=True)
    e3 = np.linalg.norm(J1 * (-o.h) - Jo) / np.linalg.norm(Jo)
    print("rel err", e1, "additive", e2, "vs oracle Jacobian", e3)
    assert e1 < 1e-6 and e2 < 1e-6 and e3 < 1e-6


def test_solve_recomposed_from_seams(gpu_ctx, prob):
    """EIT::solve (eit.cpp:139-213) written out with the seams: per harmonic, update,
    excitation beta_c B Pi, mpf_pcg_solve (dcFree on harmonic 0), measurement and Jacobian
    accumulated; then J *= -height.  Must agree with mpf_eit_solve and the oracle."""
    import torch
    from mpflow_amd import capi
    from mpflow_amd.capi import lib
    from mpflow_amd.eit import SolveOptions
    p, o, g = prob
    print("case: EIT::solve recomposed", flush=True)
    eit = eit_of(p, gpu_ctx)
    D, M, N, E, R = eit.D, eit.M, eit.N, eit.E, eit.R
    ld = (N + 15) // 16 * 16
    gt = torch.tensor(g, dtype=torch.float32, device="cuda")
    opts = SolveOptions(stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-7)
    U = torch.zeros(D, eit.Mp, dtype=torch.float32, device="cuda")
    J = torch.zeros(E, eit.m, dtype=torch.float32, device="cuda")
    for c in range(eit.C):
        alpha, beta = o.alpha_beta(c)
        ids, vals = eit.update(gt, alpha)
        F = np.zeros((R, ld), np.float32)
        F[:, :N] = (beta * (o.B @ o.Pi)).T
        Fd = torch.tensor(F, device="cuda")
        X = torch.zeros(R, ld, dtype=torch.float32, device="cuda")
        st = capi.SolveStats()
        rc = lib().mpf_pcg_solve(gpu_ctx.h, N, eit.width, C.c_void_p(ids.data_ptr()), C.c_void_p(vals.data_ptr()),
                                 R, C.c_void_p(Fd.data_ptr()), ld, C.c_void_p(X.data_ptr()), ld, int(c == 0),
                                 C.byref(opts.s), C.byref(st))
        capi.check(rc)
        eit.apply_measurement_pattern(X, U, additive=c > 0)
        eit.calc_jacobian(X, gt, J, additive=c > 0)
    J *= -o.h
    Us = U.cpu().numpy()[:, :M].T
    eit.solve(gt, SolveOptions(stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-7), jacobian=True)
    Ue = eit.voltages().cpu().numpy()[0]
    Je = eit.jacobian_matrix().cpu().numpy()[0]
    Uo, phis, _ = o.solve(g.astype(np.float32).astype(np.float64))
    Jo = o.jacobian(g.astype(np.float32).astype(np.float64), phis)
    Jc = J.cpu().numpy().T
    eu = np.max(np.abs(Us - Ue)) / np.max(np.abs(Ue))
    euo = np.max(np.abs(Us - Uo)) / np.max(np.abs(Uo))
    ej = np.linalg.norm(Jc - Je) / np.linalg.norm(Je)
    ejo = np.linalg.norm(Jc - Jo) / np.linalg.norm(Jo)
    print("U vs mpf_eit_solve", eu, "vs oracle", euo, "J vs mpf_eit_solve", ej, "vs oracle", ejo)
    assert eu < 1e-5 and euo < 1e-5 and ej < 1e-4 and ejo < 1e-4
