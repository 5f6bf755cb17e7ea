"""Device parity of the forward path against the FP64 oracle (models::EIT::solve,
src/models/eit.cpp:139-245), through the C ABI."""
import numpy as np
import pytest

from mpflow_amd.phantom import ellipse_phantom, frame_seed, inclusion_phantom
from tests.problems import eit_of, make_problem, oracle_of

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _torch():
    import torch
    return torch


@pytest.fixture(scope="module")
def c1():
    p = make_problem(16, 16, 0.1)
    o = oracle_of(p)
    g = inclusion_phantom(p["nodes"], p["elements"], frame_seed(1, 0))
    U, phis, _ = o.solve(g)
    return p, o, g, U, phis


def test_forward_c1_fp32_reference_rule(gpu_ctx, c1):
    """16 electrodes, 2D: FP32 PCG with the column-0 FLT_EPSILON stop rule; voltages
    within 1e-5 (max-norm relative) of the FP64 direct solve."""
    torch = _torch()
    p, o, g, U, _ = c1
    print("case: c1 fp32 ref rule", flush=True)
    eit = eit_of(p, gpu_ctx)
    eit.solve(torch.tensor(g, dtype=torch.float32))
    Ug = eit.voltages().cpu().numpy()[0]
    err = _rel(Ug, U)
    print("rel err", err, "steps", eit.last_stats.steps)
    assert eit.last_stats.not_converged == 0
    assert err < 1e-5
    # reciprocity of the device result (U = U^T for identical drive/measurement patterns)
    assert _rel(Ug, Ug.T) < 1e-5


def test_forward_c1_mixed_refinement(gpu_ctx, c1):
    torch = _torch()
    from mpflow_amd.eit import SolveOptions
    from mpflow_amd import capi
    p, o, g, U, _ = c1
    print("case: c1 mixed", flush=True)
    eit = eit_of(p, gpu_ctx)
    opts = SolveOptions(stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-5, precision=capi.MPF_PREC_MIXED,
                        refine_steps=2, refine_tol=1e-3)
    eit.solve(torch.tensor(g, dtype=torch.float32), opts)
    Ug = eit.voltages().cpu().numpy()[0]
    err = _rel(Ug, U)
    print("rel err", err, "refine residual", eit.last_stats.refine_residual)
    assert err < 1e-7


def test_forward_c1_warm_start(gpu_ctx, c1):
    """A warm-started mixed solve at a new gamma (correction solve of the FP64 residual of
    the previous fields) reaches the same voltages as the FP64 oracle at that gamma."""
    torch = _torch()
    from mpflow_amd.eit import SolveOptions
    from mpflow_amd import capi
    p, o, g, U, _ = c1
    print("case: c1 warm start", flush=True)
    g1 = g * (1.0 + 0.05 * np.sin(np.arange(g.size)))
    U1, _, _ = o.solve(g1)
    eit = eit_of(p, gpu_ctx)
    kw = dict(stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-5, precision=capi.MPF_PREC_MIXED,
              refine_steps=2, refine_tol=1e-3)
    eit.solve(torch.tensor(g1, dtype=torch.float32), SolveOptions(**kw))
    cold = eit.last_stats.steps
    eit.solve(torch.tensor(g, dtype=torch.float32), SolveOptions(**kw))
    eit.solve(torch.tensor(g1, dtype=torch.float32), SolveOptions(warm_start=1, **kw))
    warm = eit.last_stats.steps
    Ug = eit.voltages().cpu().numpy()[0]
    err = _rel(Ug, U1)
    print("rel err", err, "steps cold", cold, "warm", warm)
    assert err < 1e-7
    assert warm <= cold


@pytest.mark.parametrize("rule,ignore_dc", [("reference", 1), ("tolerance", 1), ("reference", 0)])
def test_forward_c1_bicgstab(gpu_ctx, c1, rule, ignore_dc):
    """BiCGSTAB (FP32; the reference's iteration, or with the harmonic-0 Krylov vectors kept
    orthogonal to constants) reaches the FP64 direct solution's voltages and Jacobian: the
    constant mode reaches neither.  The step count follows the FP64 restatement's."""
    torch = _torch()
    from mpflow_amd.eit import SolveOptions
    from mpflow_amd import capi
    import oracle.eit_oracle as eo
    p, o, g, U, phis = c1
    print("case: c1 bicgstab %s ignore_dc %d" % (rule, ignore_dc), flush=True)
    eit = eit_of(p, gpu_ctx)
    stop = capi.MPF_STOP_REFERENCE if rule == "reference" else capi.MPF_STOP_TOLERANCE
    eit.solve(torch.tensor(g, dtype=torch.float32),
              SolveOptions(stop_rule=stop, tol=1e-6, krylov=capi.MPF_KRYLOV_BICGSTAB,
                           bicgstab_ignore_dc=ignore_dc), jacobian=True)
    Ug = eit.voltages().cpu().numpy()[0]
    Jg = eit.jacobian_matrix().cpu().numpy()[0]
    Jo = o.jacobian(g, phis, quirks=True)
    eu = _rel(Ug, U)
    ej = float(np.linalg.norm(Jg - Jo) / np.linalg.norm(Jo))
    # the restatement's step count for the same system (FP64)
    A = o.system(g, 0)
    F = o.alpha_beta(0)[1] * (o.B @ o.Pi)
    _, so = eo.bicgstab_reference(A, F, 1.0 / A.diagonal(), rule=rule, tol=1e-6, project=not ignore_dc)
    print("rel err", eu, "jacobian", ej, "steps", eit.last_stats.steps, "oracle steps", so)
    assert eit.last_stats.not_converged == 0
    assert eu < 1e-5 and ej < 1e-4
    assert eit.last_stats.steps <= 2 * so + 16


def test_forward_2p5d_bicgstab_mixed(gpu_ctx):
    """BiCGSTAB with FP64 residual refinement on the 2.5D 32-electrode mesh.  FP32
    BiCGSTAB on the singular harmonic bottoms out near a 1e-6 relative residual, so the
    FP32 solves stop at 1e-5 and the refinement supplies the rest."""
    torch = _torch()
    from mpflow_amd.eit import SolveOptions
    from mpflow_amd import capi
    p = make_problem(40, 32, 0.05, C=5, height=0.4, port_height=0.1)
    o = oracle_of(p)
    g = ellipse_phantom(p["nodes"], p["elements"], frame_seed(2, 2))
    U, _, _ = o.solve(g)
    print("case: c2 bicgstab mixed", flush=True)
    eit = eit_of(p, gpu_ctx)
    eit.solve(torch.tensor(g, dtype=torch.float32),
              SolveOptions(stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-5, precision=capi.MPF_PREC_MIXED,
                           refine_steps=3, refine_tol=1e-3, krylov=capi.MPF_KRYLOV_BICGSTAB))
    Ug = eit.voltages().cpu().numpy()[0]
    err = _rel(Ug, U)
    print("rel err", err, "steps", eit.last_stats.steps, "refine residual", eit.last_stats.refine_residual)
    assert eit.last_stats.not_converged == 0
    assert err < 1e-7


@pytest.mark.parametrize("keep", [True, False])
def test_forward_node_order(gpu_ctx, keep):
    """The caller's node numbering and the internal locality numbering give the same
    voltages and Jacobian (2.5D, P2 mesh with midpoints numbered after the vertices)."""
    torch = _torch()
    p = make_problem(12, 16, 0.1, C=3, height=0.4, port_height=0.1, p2=True)
    o = oracle_of(p)
    g = ellipse_phantom(p["nodes"], p["elements"], frame_seed(2, 1))
    U, phis, _ = o.solve(g)
    Jo = o.jacobian(g, phis, quirks=True)
    print("case: node order keep=%s" % keep, flush=True)
    eit = eit_of(p, gpu_ctx, keep_node_order=keep)
    eit.solve(torch.tensor(g, dtype=torch.float32), jacobian=True)
    Ug = eit.voltages().cpu().numpy()[0]
    Jg = eit.jacobian_matrix().cpu().numpy()[0]
    eu = _rel(Ug, U)
    ej = float(np.linalg.norm(Jg - Jo) / np.linalg.norm(Jo))
    print("voltage rel err", eu, "jacobian rel err", ej)
    assert eu < 1e-5 and ej < 1e-4


def test_forward_batched_frames(gpu_ctx):
    """4 frames with different phantoms in one batched solve equal 4 oracle solves."""
    torch = _torch()
    p = make_problem(16, 16, 0.1)
    o = oracle_of(p)
    gs = [inclusion_phantom(p["nodes"], p["elements"], frame_seed(1, f)) for f in range(4)]
    print("case: batched frames", flush=True)
    eit = eit_of(p, gpu_ctx, frames=4)
    eit.solve(torch.tensor(np.stack(gs), dtype=torch.float32))
    Ug = eit.voltages().cpu().numpy()
    for f in range(4):
        U, _, _ = o.solve(gs[f])
        assert _rel(Ug[f], U) < 1e-5


def test_forward_2p5d_c2(gpu_ctx):
    """32 electrodes, 5 harmonics, h = 0.4, h_p = 0.1 on the 9600-element ring mesh."""
    torch = _torch()
    p = make_problem(40, 32, 0.05, C=5, height=0.4, port_height=0.1)
    o = oracle_of(p)
    g = ellipse_phantom(p["nodes"], p["elements"], frame_seed(2, 0))
    U, _, _ = o.solve(g)
    print("case: c2 2.5D", flush=True)
    eit = eit_of(p, gpu_ctx)
    eit.solve(torch.tensor(g, dtype=torch.float32))
    Ug = eit.voltages().cpu().numpy()[0]
    err = _rel(Ug, U)
    print("rel err", err, "steps", eit.last_stats.steps)
    assert err < 1e-5


def test_forward_p2_quirk(gpu_ctx):
    """P2 (quadratic) elements with the reference's boundary-integral quirk."""
    torch = _torch()
    p = make_problem(8, 16, 0.1, p2=True)
    o = oracle_of(p, quirks=True)
    g = inclusion_phantom(p["nodes"], p["elements"], frame_seed(1, 3))
    U, _, _ = o.solve(g)
    print("case: p2 quirk", flush=True)
    eit = eit_of(p, gpu_ctx, quirks=True)
    eit.solve(torch.tensor(g, dtype=torch.float32))
    Ug = eit.voltages().cpu().numpy()[0]
    assert _rel(Ug, U) < 1e-5


@pytest.mark.parametrize("exact", [False, True])
def test_jacobian_2p5d(gpu_ctx, exact):
    torch = _torch()
    p = make_problem(16, 16, 0.1, C=3, height=0.4, port_height=0.1)
    o = oracle_of(p)
    g = inclusion_phantom(p["nodes"], p["elements"], frame_seed(1, 5))
    U, phis, _ = o.solve(g)
    Jo = o.jacobian(g, phis, quirks=not exact)
    print("case: jacobian exact=%s" % exact, flush=True)
    eit = eit_of(p, gpu_ctx, jacobian_exact=exact)
    eit.solve(torch.tensor(g, dtype=torch.float32), jacobian=True)
    Jg = eit.jacobian_matrix().cpu().numpy()[0]
    err = float(np.linalg.norm(Jg - Jo) / np.linalg.norm(Jo))
    print("jacobian rel err", err)
    assert err < 1e-4
