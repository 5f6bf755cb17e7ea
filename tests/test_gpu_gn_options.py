"""Gauss-Newton options through the C ABI against the FP64 oracle: noser regularisation
(Inverse::calcRegularizationMatrix, src/solver/inverse.cpp:83-135) and the differential
solve (Solver::preSolve + solveDifferential, src/solver/solver.cpp:136-188) for several
frames at once, with either regularisation."""
import numpy as np
import pytest

import oracle.eit_oracle as eo
from mpflow_amd.phantom import frame_seed, inclusion_phantom
from tests.problems import eit_of, make_problem, oracle_of

pytestmark = pytest.mark.gpu

LAM = 1e-2


@pytest.fixture(scope="module")
def gnprob():
    p = make_problem(10, 16, 0.1, C=2, height=0.4, port_height=0.1)
    o = oracle_of(p)
    ones = np.ones(o.E)
    U1, phis1, _ = o.solve(ones)
    J1 = o.jacobian(ones, phis1)
    gs = [inclusion_phantom(p["nodes"], p["elements"], frame_seed(7, f)) for f in range(3)]
    Um = [o.flatten_voltages(o.solve(g)[0]) for g in gs]
    return p, o, o.flatten_voltages(U1), J1, Um


def _opts():
    from mpflow_amd import capi
    from mpflow_amd.eit import SolveOptions
    return SolveOptions(stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-6, precision=capi.MPF_PREC_MIXED,
                        refine_steps=2, refine_tol=1e-3)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("reg", ["identity", "noser"])
def test_gn_step_regularization(gpu_ctx, gnprob, reg):
    import torch
    from mpflow_amd.eit import Solver
    p, o, U1, J1, Um = gnprob
    print("case: GN step with %s regularisation" % reg, flush=True)
    eit = eit_of(p, gpu_ctx)
    s = Solver(eit, LAM, gn_refine=4, options=_opts())
    s.set_regularization(reg)
    s.measurement[0] = torch.tensor(Um[0], dtype=torch.float64)
    s.solve_absolute()
    g = s.material_distribution[0].cpu().numpy().astype(np.float64)
    R = eo.noser_diag(J1) if reg == "noser" else np.ones(o.E)
    want = 1.0 + eo.gn_step_regularized(J1, Um[0] - U1, LAM ** 2, R)
    err = _rel(g - 1.0, want - 1.0)
    errg = _rel(g, want)
    print("delta rel err", err, "gamma rel err", errg)
    # the step amplifies the FP32 Jacobian's rounding by the normal matrix's conditioning
    # (lambda = 1e-2): the step agrees to ~1e-3, gamma (the reconstruction) far better
    assert err < 1e-2 and errg < 1e-4


@pytest.mark.parametrize("reg", ["identity", "noser"])
def test_differential_solve_frames(gpu_ctx, gnprob, reg):
    """Three frames against one preparation equal three one-step reconstructions from the
    reference distribution (oracle) and the device's own absolute step from it."""
    import torch
    from mpflow_amd.eit import Solver
    p, o, U1, J1, Um = gnprob
    print("case: differential solve, %s" % reg, flush=True)
    eit = eit_of(p, gpu_ctx, frames=3)
    s = Solver(eit, LAM, gn_refine=4, options=_opts())
    s.set_regularization(reg)
    s.pre_solve()
    s.measurement.copy_(torch.tensor(np.stack(Um), dtype=torch.float64))
    s.solve_differential()
    g = s.material_distribution.cpu().numpy().astype(np.float64)
    R = eo.noser_diag(J1) if reg == "noser" else np.ones(o.E)
    errs, errg = [], []
    for f in range(3):
        want = 1.0 + eo.gn_step_regularized(J1, Um[f] - U1, LAM ** 2, R)
        errs.append(_rel(g[f] - 1.0, want - 1.0))
        errg.append(_rel(g[f], want))
    # the absolute step from the reference distribution is the same linear step
    s.material_distribution.fill_(1.0)
    s.solve_absolute()
    ga = s.material_distribution.cpu().numpy().astype(np.float64)
    dab = _rel(g - 1.0, ga - 1.0)
    print("delta rel err per frame", errs, "gamma", errg, "vs absolute step", dab)
    assert max(errs) < 1e-2 and max(errg) < 1e-4 and dab < 1e-5


@pytest.mark.parametrize("differential", [False, True])
def test_laplacian_regularization(gpu_ctx, gnprob, differential):
    """RegularizationType::laplacian (inverse.cpp:96-117): R = L^T L from the mesh; the
    device's E-space CG (preconditioned by the data-space solve with R's diagonal) reaches
    the oracle's dense solve of (J^T J + lam^2 R) delta = J^T r."""
    import torch
    from mpflow_amd.eit import Solver
    p, o, U1, J1, Um = gnprob
    print("case: laplacian regularisation, differential=%d" % differential, flush=True)
    R = eo.laplacian_matrix(p["nodes"], p["elements"])
    eit = eit_of(p, gpu_ctx, frames=3 if differential else 1)
    s = Solver(eit, LAM, gn_refine=4, options=_opts())
    s.set_regularization("laplacian")
    if differential:
        s.pre_solve()
        s.measurement.copy_(torch.tensor(np.stack(Um), dtype=torch.float64))
        s.solve_differential()
        frames = range(3)
    else:
        s.measurement[0] = torch.tensor(Um[0], dtype=torch.float64)
        s.solve_absolute()
        frames = range(1)
    g = s.material_distribution.cpu().numpy().astype(np.float64)
    errs, errg = [], []
    for f in frames:
        want = 1.0 + eo.gn_step_general(J1, Um[f] - U1, LAM ** 2, R)
        errs.append(_rel(g[f] - 1.0, want - 1.0))
        errg.append(_rel(g[f], want))
    print("delta rel err", errs, "gamma rel err", errg)
    assert max(errs) < 1e-2 and max(errg) < 1e-4


def _masked_step(J, r, lam2, reg, mask):
    """The differential step with the Jacobian multiplied by ``mask``: elements whose
    masked column is zero keep delta = 0 (the reference's CG from 0 never leaves their
    null space), the others solve the reduced regularised system."""
    Jm = J * mask
    keep = np.abs(Jm).sum(axis=0) > 0
    Jk = Jm[:, keep]
    R = eo.noser_diag(Jk) if reg == "noser" else np.ones(Jk.shape[1])
    d = np.zeros(J.shape[1])
    d[keep] = eo.gn_step_regularized(Jk, r, lam2, R)
    return d


def _mask(shape, seed):
    rng = np.random.default_rng(seed)
    m = rng.uniform(0.5, 1.5, shape)
    m[:, rng.random(shape[1]) < 0.2] = 0.0          # elements without sensitivity
    m[rng.random(shape[0]) < 0.1, :] = 0.0          # and some measurements dropped
    return m


@pytest.mark.parametrize("reg", ["identity", "noser"])
def test_differential_sensitivity_mask(gpu_ctx, gnprob, reg):
    """noSensitivity (Solver::fromConfig, solver.cpp:107-113): the mask multiplies the
    preSolve Jacobian (and so the noser weights); masked elements do not move."""
    import torch
    from mpflow_amd.eit import Solver
    p, o, U1, J1, Um = gnprob
    print("case: differential solve with a sensitivity mask, %s" % reg, flush=True)
    eit = eit_of(p, gpu_ctx, frames=3)
    mask = _mask(J1.shape, 11)
    eit.set_sensitivity_mask(mask.astype(np.float32))
    s = Solver(eit, LAM, gn_refine=4, options=_opts())
    s.set_regularization(reg)
    s.pre_solve()
    s.measurement.copy_(torch.tensor(np.stack(Um), dtype=torch.float64))
    s.solve_differential()
    g = s.material_distribution.cpu().numpy().astype(np.float64)
    m32 = mask.astype(np.float32).astype(np.float64)
    errs = []
    for f in range(3):
        want = 1.0 + _masked_step(J1, Um[f] - U1, LAM ** 2, reg, m32)
        errs.append(_rel(g[f], want))
        dead = np.abs(J1 * m32).sum(axis=0) == 0
        assert np.all(g[f][dead] == 1.0)
    print("gamma rel err per frame", errs)
    assert max(errs) < 1e-4
    # clearing the mask restores the unmasked preparation
    eit.set_sensitivity_mask(None)
    s.set_regularization(reg)
    s.pre_solve()
    s.solve_differential()
    g0 = s.material_distribution.cpu().numpy().astype(np.float64)
    R = eo.noser_diag(J1) if reg == "noser" else np.ones(o.E)
    assert _rel(g0[0], 1.0 + eo.gn_step_regularized(J1, Um[0] - U1, LAM ** 2, R)) < 1e-4


def test_solver_from_config_differential(gpu_ctx, tmp_path):
    """Solver::fromConfig (solver.cpp:70-135) from a JSON file and mesh text files:
    preSolve at gamma = 1 with the configured regularisation, then one differential solve
    per frame equals the oracle's linear step."""
    import json

    import torch
    from mpflow_amd import config as cfg
    from mpflow_amd.mesh import ring_mesh
    nodes, el = ring_mesh(8, 16, 0.1)
    d = tmp_path / "mesh"
    d.mkdir()
    cfg.savetxt(str(d / "nodes.txt"), nodes)
    cfg.savetxt(str(d / "elements.txt"), el)
    conf = {"model": {"mesh": "mesh", "ports": {"count": 16, "width": 0.1, "height": 0.1},
                      "source": {"type": "current", "value": 1e-3},
                      "2.5D": {"components": 2, "height": 0.4}, "material": 1.0},
            "solver": {"parallelImages": 2, "regularizationFactor": LAM, "regularizationType": "noser"}}
    json.dump(conf, open(tmp_path / "config.json", "w"))
    print("case: solver from config", flush=True)
    s = cfg.solver_from_config(str(tmp_path / "config.json"), ctx=gpu_ctx, options=_opts())
    spec, _ = cfg.parse_config(str(tmp_path / "config.json"))
    o = eo.EITOracle(spec.nodes, spec.elements, spec.boundary, spec.ports, spec.drive, spec.meas, spec.current,
                     1.0, 2, 0.4, 0.1)
    ones = np.ones(o.E)
    U1, phis1, _ = o.solve(ones)
    J1 = o.jacobian(ones, phis1)
    gs = [inclusion_phantom(spec.nodes, spec.elements, frame_seed(8, f)) for f in range(2)]
    Um = np.stack([o.flatten_voltages(o.solve(g)[0]) for g in gs])
    s.measurement.copy_(torch.tensor(Um, dtype=torch.float64))
    s.solve_differential()
    g = s.material_distribution.cpu().numpy().astype(np.float64)
    errs = [_rel(g[f], 1.0 + eo.gn_step_regularized(J1, Um[f] - o.flatten_voltages(U1), LAM ** 2,
                                                      eo.noser_diag(J1))) for f in range(2)]
    print("gamma rel err", errs)
    assert max(errs) < 1e-4


def test_logarithmic_parametrisation(gpu_ctx):
    """Equation<..., logarithmic=true> (equation_kernel.cu:35-53): sigma = ref 10^(gamma/10);
    forward voltages and one GN step from gamma = 0 against the oracle."""
    import torch
    from mpflow_amd.eit import EIT, Solver
    p = make_problem(10, 16, 0.1, C=2, height=0.4, port_height=0.1)
    o = eo.EITOracle(p["nodes"], p["elements"], p["boundary"], p["ports"], p["drive"], p["meas"], p["current"],
                     2.0, 2, 0.4, 0.1, logarithmic=True)
    g = 10.0 * np.log10(inclusion_phantom(p["nodes"], p["elements"], frame_seed(9, 0)))
    g32 = g.astype(np.float32).astype(np.float64)
    print("case: logarithmic", flush=True)
    eit = EIT(p["nodes"], p["elements"], p["boundary"], p["ports"], p["drive"], p["meas"], p["current"], 2.0, 2,
              0.4, 0.1, logarithmic=True, ctx=gpu_ctx)
    eit.solve(torch.tensor(g, dtype=torch.float32), _opts())
    U, _, _ = o.solve(g32)
    ef = float(np.max(np.abs(eit.voltages().cpu().numpy()[0] - U)) / np.max(np.abs(U)))
    s = Solver(eit, LAM, options=_opts())
    assert float(s.material_distribution.abs().max()) == 0.0      # log reference distribution
    Um = o.flatten_voltages(U)
    s.measurement[0] = torch.tensor(Um, dtype=torch.float64)
    s.solve_absolute()
    z = np.zeros(o.E)
    U0, phis0, _ = o.solve(z)
    want = eo.gn_step_exact(o.jacobian(z, phis0), Um - o.flatten_voltages(U0), (2.0 * LAM) ** 2)
    got = s.material_distribution[0].cpu().numpy().astype(np.float64)
    es = _rel(got, want)
    print("forward rel err", ef, "GN step rel err", es)
    assert ef < 1e-6 and es < 1e-2
