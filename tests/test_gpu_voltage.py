"""Fixed (voltage) sources (Sources::Type::Fixed; applyMixedBoundaryCondition,
src/models/eit.cpp:89-91,159-205, src/models/eit_kernel.cu:34-54) through the C ABI
against the FP64 oracle: electrode currents, Jacobian and one Gauss-Newton step."""
import numpy as np
import pytest

import oracle.eit_oracle as eo
from mpflow_amd.phantom import ellipse_phantom, frame_seed
from tests.problems import make_problem

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _pair(p, ctx, voltage_value=2.0, frames=1):
    from mpflow_amd.eit import EIT
    o = eo.EITOracle(p["nodes"], p["elements"], p["boundary"], p["ports"], p["drive"], p["meas"], voltage_value,
                     1.0, p["C"], p["height"], p["port_height"], voltage=True)
    eit = EIT(p["nodes"], p["elements"], p["boundary"], p["ports"], p["drive"], p["meas"], voltage_value, 1.0,
              p["C"], p["height"], p["port_height"], frames=frames, ctx=ctx, source="voltage")
    return o, eit


def _opts(**kw):
    from mpflow_amd import capi
    from mpflow_amd.eit import SolveOptions
    base = dict(stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-6, precision=capi.MPF_PREC_MIXED, refine_steps=2,
                refine_tol=1e-3)
    base.update(kw)
    return SolveOptions(**base)


CASES = [("p1-2d", dict(n_rings=12, L=16, width=0.1, C=1)),
         ("p1-2p5d", dict(n_rings=16, L=16, width=0.1, C=3, height=0.4, port_height=0.1)),
         ("p2-2p5d", dict(n_rings=8, L=16, width=0.1, C=2, height=0.4, port_height=0.1, p2=True))]


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("variant", ["cg", "two-level", "bicgstab", "fp32"])
def test_voltage_forward_and_jacobian(gpu_ctx, name, kw, variant):
    import torch
    from mpflow_amd import capi
    p = make_problem(**kw)
    o, eit = _pair(p, gpu_ctx)
    g = ellipse_phantom(p["nodes"], p["elements"], frame_seed(6, 2))
    print("case: voltage sources %s %s" % (name, variant), flush=True)
    U, phis, _ = o.solve(g.astype(np.float32).astype(np.float64))
    Jo = o.jacobian(g.astype(np.float32).astype(np.float64), phis)
    opts = {"cg": _opts(), "two-level": _opts(preconditioner=capi.MPF_PRECOND_TWO_LEVEL),
            "bicgstab": _opts(krylov=capi.MPF_KRYLOV_BICGSTAB),
            "fp32": _opts(precision=capi.MPF_PREC_FP32)}[variant]
    eit.solve(torch.tensor(g, dtype=torch.float32), opts, jacobian=True)
    assert eit.last_stats.not_converged == 0
    Ug = eit.voltages().cpu().numpy()[0]
    Jg = eit.jacobian_matrix().cpu().numpy()[0].astype(np.float64)
    eu, ej = _rel(Ug, U), _rel(Jg, Jo)
    print("currents rel err", eu, "jacobian rel err", ej, "steps", eit.last_stats.steps)
    assert eu < (1e-5 if variant == "fp32" else 1e-7)
    assert ej < 1e-4


def test_voltage_gn_step(gpu_ctx):
    """One Solver::solveAbsolute with voltage sources equals the oracle's FP64 data-space
    step from the same start (Jacobian x height, no sign flip)."""
    import torch
    from mpflow_amd.eit import Solver
    p = make_problem(12, 16, 0.1, C=2, height=0.4, port_height=0.1)
    o, eit = _pair(p, gpu_ctx)
    gt = ellipse_phantom(p["nodes"], p["elements"], frame_seed(6, 5))
    Um, _, _ = o.solve(gt)
    ones = np.ones(o.E)
    U0, phis0, _ = o.solve(ones)
    lam = 1e-2
    want = 1.0 + eo.gn_step_exact(o.jacobian(ones, phis0), o.flatten_voltages(Um) - o.flatten_voltages(U0),
                                  lam * lam)
    print("case: voltage GN step", flush=True)
    s = Solver(eit, lam, gn_refine=4, options=_opts())
    s.measurement[0] = torch.tensor(o.flatten_voltages(Um), dtype=torch.float64)
    s.solve_absolute()
    got = s.material_distribution[0].cpu().numpy().astype(np.float64)
    err = _rel(got, want)
    print("gamma rel err", err)
    assert err < 1e-4
