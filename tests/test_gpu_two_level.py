"""The two-level preconditioner (Jacobi + coarse correction over the 64-row quarters of the 256-row tiles; an
extension, the reference has only the diagonal preconditioner of
src/numeric/preconditioner.cpp:59-76) on the device, through the C ABI: step counts
against the oracle's restatement, voltages against the FP64 direct solve, and the same
answers as Jacobi in every mode (mixed precision, batches with graphs, CEM, GN)."""
import numpy as np
import pytest

import oracle.eit_oracle as eo
from mpflow_amd.phantom import ellipse_phantom, frame_seed, inclusion_phantom
from tests.problems import eit_of, make_problem, oracle_of, rcb_renumbered

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _opts(pre, **kw):
    from mpflow_amd import capi
    from mpflow_amd.eit import SolveOptions
    kw.setdefault("stop_rule", capi.MPF_STOP_TOLERANCE)
    kw.setdefault("tol", 1e-6)
    return SolveOptions(preconditioner=pre, **kw)


def test_two_level_steps_match_oracle(gpu_ctx):
    """Caller order kept: the device tiles are the caller's consecutive 256 rows, here RCB
    leaves; the device step count follows the oracle's two-level PCG."""
    import torch
    p = rcb_renumbered(make_problem(24, 16, 0.1, C=2, height=0.4, port_height=0.1), 256)
    o = oracle_of(p)
    g = inclusion_phantom(p["nodes"], p["elements"], frame_seed(1, 6))
    g32 = g.astype(np.float32).astype(np.float64)
    print("case: two-level steps vs oracle", flush=True)
    ones = np.ones(o.E)
    Sd, S1d = o.system(g32, 0).diagonal(), o.system(ones, 0).diagonal()
    steps_o, steps_j = 0, 0
    for c in range(o.C):
        A = o.system(g32, c)
        F = o.alpha_beta(c)[1] * (o.B @ o.Pi)
        prec, _ = eo.two_level_preconditioner(A, o.system(ones, c), Sd, S1d, c == 0, tile=256)
        _, s = eo.pcg_reference(A, F, None, c == 0, rule="tolerance", tol=1e-6, prec=prec)
        _, sj = eo.pcg_reference(A, F, 1.0 / A.diagonal(), c == 0, rule="tolerance", tol=1e-6)
        steps_o += s
        steps_j += sj
    eit = eit_of(p, gpu_ctx, keep_node_order=True)
    gt = torch.tensor(g, dtype=torch.float32)
    eit.solve(gt, _opts(1))
    st2 = eit.last_stats
    U2 = eit.voltages().cpu().numpy()[0]
    eit.solve(gt, _opts(0))
    st1 = eit.last_stats
    U, _, _ = o.solve(g32)
    print("device two-level steps", st2.steps, "oracle", steps_o, "| device jacobi", st1.steps, "oracle", steps_j,
          "| rel err", _rel(U2, U))
    assert st2.not_converged == 0
    # FP32 iterations to 1e-6 run longer than FP64 ones (15-35%, the more so the
    # stronger the preconditioner); the two-level / Jacobi step ratio is the
    # preconditioner's own signature
    assert abs(st2.steps - steps_o) <= 0.4 * steps_o
    assert abs(st1.steps - steps_j) <= 0.25 * steps_j
    assert abs(st2.steps / st1.steps - steps_o / steps_j) < 0.2 * steps_o / steps_j
    assert st2.steps < 0.85 * st1.steps
    assert _rel(U2, U) < 1e-5


@pytest.mark.parametrize("mixed", [False, True])
def test_two_level_renumbered_2p5d(gpu_ctx, mixed):
    """Device renumbering (RCB tiles), 32 electrodes, 5 harmonics: the same voltages and
    Jacobian as Jacobi, fewer steps."""
    import torch
    from mpflow_amd import capi
    p = make_problem(40, 32, 0.05, C=5, height=0.4, port_height=0.1)
    o = oracle_of(p)
    g = ellipse_phantom(p["nodes"], p["elements"], frame_seed(2, 3))
    print("case: two-level 2.5D mixed=%d" % mixed, flush=True)
    U, _, _ = o.solve(g.astype(np.float32).astype(np.float64))
    eit = eit_of(p, gpu_ctx)
    gt = torch.tensor(g, dtype=torch.float32)
    kw = dict(precision=capi.MPF_PREC_MIXED, refine_steps=2, refine_tol=1e-3, tol=1e-5) if mixed else {}
    res = {}
    for pre in (0, 1):
        eit.solve(gt, _opts(pre, **kw), jacobian=True)
        res[pre] = (eit.voltages().cpu().numpy()[0], eit.jacobian_matrix().cpu().numpy()[0], eit.last_stats.steps)
    e0, e1 = _rel(res[0][0], U), _rel(res[1][0], U)
    ej = float(np.linalg.norm(res[1][1] - res[0][1]) / np.linalg.norm(res[0][1]))
    print("jacobi steps", res[0][2], "two-level steps", res[1][2], "rel err", e0, e1, "J diff", ej)
    bar = 1e-7 if mixed else 1e-5
    assert e0 < bar and e1 < bar
    assert ej < (1e-6 if mixed else 1e-4)
    assert res[1][2] < 0.7 * res[0][2]


def test_two_level_batch_graphs_cem_gn(gpu_ctx):
    """Two models of different sizes (one with the complete electrode model) in one batch
    with HIP-graph replay: each model's voltages equal its own solve's; a Gauss-Newton step
    with the two-level preconditioner lands where the Jacobi one does."""
    import torch
    from mpflow_amd import capi
    from mpflow_amd.eit import Batch, Solver
    pa = make_problem(20, 16, 0.1, C=3, height=0.4, port_height=0.1)
    pb = make_problem(28, 32, 0.05, C=2, height=0.4, port_height=0.1)
    ga = inclusion_phantom(pa["nodes"], pa["elements"], frame_seed(3, 1))
    gb = ellipse_phantom(pb["nodes"], pb["elements"], frame_seed(3, 2))
    print("case: two-level batch + graphs + CEM", flush=True)
    ea = eit_of(pa, gpu_ctx)
    eb = eit_of(pb, gpu_ctx, contact_impedance=0.01)
    ta, tb = torch.tensor(ga, dtype=torch.float32), torch.tensor(gb, dtype=torch.float32)
    o2 = _opts(1, graphs=1)
    ea.solve(ta, o2)
    Ua = ea.result.clone()
    eb.solve(tb, o2)
    Ub = eb.result.clone()
    b = Batch([ea, eb])
    b.solve([ta, tb], o2)
    da = (ea.result - Ua).abs().max().item() / Ua.abs().max().item()
    db = (eb.result - Ub).abs().max().item() / Ub.abs().max().item()
    Uo, _, _ = oracle_of(pb, contact_impedance=0.01).solve(gb.astype(np.float32).astype(np.float64))
    ecem = _rel(eb.voltages().cpu().numpy()[0], Uo)
    b.close()
    print("batch vs single", da, db, "CEM vs oracle", ecem)
    assert da < 1e-6 and db < 1e-6 and ecem < 1e-5
    # one GN step from gamma = 1 towards the measured voltages of ga
    oa = oracle_of(pa)
    Um = oa.flatten_voltages(oa.solve(ga)[0])
    out = {}
    for pre in (0, 1):
        s = Solver(ea, 1e-3, options=_opts(pre, precision=capi.MPF_PREC_MIXED, refine_steps=2, refine_tol=1e-3))
        s.measurement[0] = torch.tensor(Um, dtype=torch.float64)
        s.solve_absolute()
        out[pre] = s.material_distribution[0].cpu().numpy().astype(np.float64)
    dg = float(np.linalg.norm(out[1] - out[0]) / np.linalg.norm(out[0]))
    print("GN step two-level vs Jacobi", dg)
    assert dg < 1e-5
