"""Complete electrode model (contact impedance on every port; an extension without a
reference counterpart) on the device against the FP64 oracle's CEM restatement, whose
blocks are pinned by the electrode-current, power-balance, reciprocity,
finite-difference and shunt-limit checks of the oracle tests."""
import numpy as np
import pytest

from mpflow_amd.phantom import add_noise, ellipse_phantom, frame_seed, inclusion_phantom
from tests.problems import eit_of, make_problem, oracle_of

pytestmark = pytest.mark.gpu

Z = 0.01


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("p2", [False, True])
def test_cem_c1_fp32(gpu_ctx, p2):
    """16 electrodes, 2D, z_c = 0.01: FP32 PCG voltages within 1e-5 and the Jacobian
    within 1e-4 (relative L2) of the oracle; U = U^T."""
    import torch
    from mpflow_amd import capi
    from mpflow_amd.eit import SolveOptions
    p = make_problem(16 if not p2 else 8, 16, 0.1, p2=p2)
    o = oracle_of(p, contact_impedance=Z)
    g = inclusion_phantom(p["nodes"], p["elements"], frame_seed(1, 0))
    U, phis, _ = o.solve(g)
    Jo = o.jacobian(g, phis)
    print("case: cem c1 p2=%s" % p2, flush=True)
    eit = eit_of(p, gpu_ctx, contact_impedance=Z)
    eit.solve(torch.tensor(g, dtype=torch.float32), SolveOptions(stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-6),
              jacobian=True)
    Ug = eit.voltages().cpu().numpy()[0]
    Jg = eit.jacobian_matrix().cpu().numpy()[0]
    eu = _rel(Ug, U)
    ej = float(np.linalg.norm(Jg - Jo) / np.linalg.norm(Jo))
    print("rel err", eu, "jacobian", ej, "steps", eit.last_stats.steps)
    assert eu < 1e-5 and ej < 1e-4
    assert _rel(Ug, Ug.T) < 1e-5


def test_cem_c2_2p5d_mixed(gpu_ctx):
    """32 electrodes, 5 harmonics, h = 0.4, h_p = 0.1, z_c = 0.01 (config 2's CEM mode):
    mixed precision within 1e-7 of the oracle."""
    import torch
    from mpflow_amd import capi
    from mpflow_amd.eit import SolveOptions
    p = make_problem(40, 32, 0.05, C=5, height=0.4, port_height=0.1)
    o = oracle_of(p, contact_impedance=Z)
    g = ellipse_phantom(p["nodes"], p["elements"], frame_seed(2, 0))
    U, _, _ = o.solve(g)
    print("case: cem c2 mixed", flush=True)
    eit = eit_of(p, gpu_ctx, contact_impedance=Z)
    eit.solve(torch.tensor(g, dtype=torch.float32),
              SolveOptions(stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-6, precision=capi.MPF_PREC_MIXED,
                           refine_steps=2, refine_tol=1e-3))
    err = _rel(eit.voltages().cpu().numpy()[0], U)
    print("rel err", err, "steps", eit.last_stats.steps, "refine residual", eit.last_stats.refine_residual)
    assert err < 1e-7


def test_cem_gn_step(gpu_ctx):
    """One Gauss-Newton step with the electrode model against the FP64 data-space step."""
    import torch
    from mpflow_amd import capi
    from mpflow_amd.eit import SolveOptions, Solver
    from oracle.eit_oracle import gn_step_exact
    p = make_problem(16, 16, 0.1)
    o = oracle_of(p, contact_impedance=Z)
    g = inclusion_phantom(p["nodes"], p["elements"], frame_seed(1, 1))
    U, _, _ = o.solve(g)
    Um = add_noise(U, 60.0, frame_seed(1, 1))
    g0 = np.ones(p["elements"].shape[0])
    U0, phis, _ = o.solve(g0)
    J = o.jacobian(g0, phis, quirks=False)
    d_ref = gn_step_exact(J, o.flatten_voltages(Um) - o.flatten_voltages(U0), 1e-6)
    print("case: cem gn step", flush=True)
    eit = eit_of(p, gpu_ctx, jacobian_exact=True, contact_impedance=Z)
    s = Solver(eit, 1e-3, gn_refine=4,
               options=SolveOptions(stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-6, precision=capi.MPF_PREC_MIXED,
                                    refine_steps=2, refine_tol=1e-3))
    s.measurement[0] = torch.tensor(o.flatten_voltages(Um))
    s.solve_absolute()
    g1 = s.material_distribution.cpu().numpy()[0].astype(np.float64)
    d = g1 - 1.0
    err = float(np.linalg.norm(d - d_ref) / np.linalg.norm(d_ref))
    err_g = float(np.linalg.norm(g1 - (1.0 + d_ref)) / np.linalg.norm(1.0 + d_ref))
    print("GN update rel err", err, "gamma rel err", err_g)
    # the small update carries the FP32 Jacobian's error amplified by the conditioning of
    # J J^T + lambda^2 I; the reconstruction (gamma) is the quantity held to 1e-3
    assert err < 3e-3
    assert err_g < 1e-5


def _coupled_pair(p, gpu_ctx, frames=1):
    import oracle.eit_oracle as eo
    from mpflow_amd.eit import EIT
    o = eo.CoupledCEMOracle(p["nodes"], p["elements"], p["boundary"], p["ports"], p["drive"], p["meas"],
                            p["current"], 1.0, p["C"], p["height"], p["port_height"], contact_impedance=Z)
    eit = EIT(p["nodes"], p["elements"], p["boundary"], p["ports"], p["drive"], p["meas"], p["current"], 1.0,
              p["C"], p["height"], p["port_height"], jacobian_exact=True, frames=frames, contact_impedance=Z,
              cem="coupled", ctx=gpu_ctx)
    return o, eit


@pytest.mark.parametrize("C,h,hp,mode", [(1, 1.0, 1.0, "fp32"), (3, 0.4, 0.1, "fp32"), (5, 0.4, 0.1, "mixed"),
                                         (3, 2.0, 0.5, "mixed")])
def test_cem_coupled_forward_and_jacobian(gpu_ctx, C, h, hp, mode):
    """Electrodes at one potential for all harmonics (Schur complement on V): voltages and
    Jacobian of the device against the oracle's direct solve of the whole block system."""
    import torch
    from mpflow_amd import capi
    from mpflow_amd.eit import SolveOptions
    p = make_problem(16, 16, 0.1, C=C, height=h, port_height=hp)
    o, eit = _coupled_pair(p, gpu_ctx)
    g = inclusion_phantom(p["nodes"], p["elements"], frame_seed(1, 2))
    U, phis, _ = o.solve(g)
    Jo = o.jacobian(g, phis)
    opts = SolveOptions(stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-6)
    if mode == "mixed":
        opts = SolveOptions(stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-6, precision=capi.MPF_PREC_MIXED,
                            refine_steps=2, refine_tol=1e-3)
    print("case: coupled cem C=%d h=%g hp=%g %s" % (C, h, hp, mode), flush=True)
    eit.solve(torch.tensor(g, dtype=torch.float32), opts, jacobian=True)
    Ug = eit.voltages().cpu().numpy()[0]
    Jg = eit.jacobian_matrix().cpu().numpy()[0]
    eu = _rel(Ug, U)
    ej = float(np.linalg.norm(Jg - Jo) / np.linalg.norm(Jo))
    print("voltages rel err %.2e, jacobian %.2e, steps %d, not converged %d"
          % (eu, ej, eit.last_stats.steps, eit.last_stats.not_converged), flush=True)
    assert eit.last_stats.not_converged == 0
    assert eu < (1e-5 if mode == "fp32" else 1e-7)
    assert ej < 1e-4
    assert _rel(Ug, Ug.T) < 1e-5


def test_cem_coupled_c2_reconstruction(gpu_ctx):
    """Config 2 with the coupled electrode model (40-ring mesh, 32 electrodes, 5 harmonics,
    z_c = 0.01): 3 Gauss-Newton steps, two frames per launch, against the oracle's FP64
    reconstruction."""
    import torch
    from mpflow_amd import capi
    from mpflow_amd.eit import SolveOptions, Solver
    from oracle.eit_oracle import reconstruct
    p = make_problem(40, 32, 0.05, C=5, height=0.4, port_height=0.1)
    o, eit = _coupled_pair(p, gpu_ctx, frames=2)
    Ums, refs = [], []
    for f in range(2):
        g = ellipse_phantom(p["nodes"], p["elements"], frame_seed(2, 10 + f))
        U, _, _ = o.solve(g)
        Um = add_noise(U, 60.0, frame_seed(2, 10 + f))
        Ums.append(o.flatten_voltages(Um))
        print("case: coupled cem c2 oracle reconstruction %d" % f, flush=True)
        refs.append(reconstruct(o, Um, np.ones(o.E), 1e-3, steps=3))
    opts = SolveOptions(stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-6, precision=capi.MPF_PREC_MIXED,
                        refine_steps=1, refine_tol=1e-2, preconditioner=capi.MPF_PRECOND_TWO_LEVEL)
    s = Solver(eit, 1e-3, gn_refine=4, options=opts)
    s.measurement.copy_(torch.tensor(np.stack(Ums)))
    print("case: coupled cem c2 device reconstruction", flush=True)
    for k in range(3):
        s.solve_absolute(warm_start=k > 0)
        assert s.last_stats.not_converged == 0
    gd = s.material_distribution.cpu().numpy().astype(np.float64)
    for f in range(2):
        err = np.linalg.norm(gd[f] - refs[f]) / np.linalg.norm(refs[f])
        err_d = np.linalg.norm(gd[f] - refs[f]) / np.linalg.norm(refs[f] - 1)
        print("frame %d: gamma %.2e, update %.2e" % (f, err, err_d), flush=True)
        assert err < 1e-3
        assert err_d < 1e-2
