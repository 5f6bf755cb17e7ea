"""Batched solves of several models of different sizes, electrode and harmonic counts
(the uneven multi-job workload, config 4) through mpf_batch_*: parity with the FP64
oracle, with the same models solved one at a time, and between HIP-graph replay and
plain launches."""
import numpy as np
import pytest

from mpflow_amd.phantom import add_noise, ellipse_phantom, frame_seed
from oracle.eit_oracle import EITOracle

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def c4():
    from bench import build_problem, c4_job
    jobs = []
    for j in range(40):
        w = c4_job(j)
        if w["rings"] <= 45:
            jobs.append((j, w))
        if len(jobs) == 5:
            break
    out = []
    for j, w in jobs:
        nodes, el, bd, ports, pat = build_problem(w)
        o = EITOracle(nodes, el, bd, ports, pat, pat, 1e-3, 1.0, w["C"], w["height"], w["port_height"])
        g = ellipse_phantom(nodes, el, frame_seed(4, j))
        U, _, _ = o.solve(g)
        out.append(dict(j=j, w=w, nodes=nodes, el=el, bd=bd, ports=ports, pat=pat, o=o, g=g, U=U))
    print("case: c4 jobs", [(d["j"], d["w"]["rings"], d["w"]["L"], d["w"]["C"]) for d in out], flush=True)
    return out


def _models(c4, ctx):
    from mpflow_amd.eit import EIT
    return [EIT(d["nodes"], d["el"], d["bd"], d["ports"], d["pat"], d["pat"], 1e-3, 1.0, d["w"]["C"],
                d["w"]["height"], d["w"]["port_height"], ctx=ctx) for d in c4]


def _opts(**kw):
    from mpflow_amd import capi
    from mpflow_amd.eit import SolveOptions
    base = dict(stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-6, precision=capi.MPF_PREC_MIXED,
                refine_steps=2, refine_tol=1e-3)
    base.update(kw)
    return SolveOptions(**base)


@pytest.mark.parametrize("graphs", [0, 1])
def test_batch_forward_matches_oracle_and_single(gpu_ctx, c4, graphs):
    import torch
    from mpflow_amd.eit import Batch
    print("case: c4 batch forward graphs=%d" % graphs, flush=True)
    models = _models(c4, gpu_ctx)
    batch = Batch(models)
    gs = [torch.tensor(d["g"], dtype=torch.float32) for d in c4]
    batch.solve(gs, _opts(graphs=graphs))
    Ub = [m.voltages().cpu().numpy()[0] for m in models]
    print("batch steps", batch.last_stats.steps, "refine residual", batch.last_stats.refine_residual)
    assert batch.last_stats.not_converged == 0
    for d, U in zip(c4, Ub):
        err = _rel(U, d["U"])
        print("job", d["j"], "rel err vs FP64", err)
        assert err < 1e-7
    # the same models one at a time
    for m, g, U in zip(models, gs, Ub):
        m.solve(g, _opts(graphs=graphs))
        assert _rel(m.voltages().cpu().numpy()[0], U) < 1e-9
    batch.close()


def test_batch_graphs_replay_identical(gpu_ctx, c4):
    """Graph replay runs the same kernels on the same data: FP32 results are identical."""
    import torch
    from mpflow_amd import capi
    from mpflow_amd.eit import Batch
    print("case: c4 graphs vs launches", flush=True)
    models = _models(c4, gpu_ctx)
    batch = Batch(models)
    gs = [torch.tensor(d["g"], dtype=torch.float32) for d in c4]
    out = []
    for graphs in (0, 1, 1):
        batch.solve(gs, _opts(precision=capi.MPF_PREC_FP32, graphs=graphs))
        out.append([m.result.clone() for m in models])
    for k in (1, 2):
        for a, b in zip(out[0], out[k]):
            assert torch.equal(a, b)
    batch.close()


def test_batch_gn_step(gpu_ctx, c4):
    """One batched Gauss-Newton step equals the per-model steps."""
    import torch
    from mpflow_amd.eit import Batch, BatchSolver, Solver
    print("case: c4 batch gn step", flush=True)
    models = _models(c4, gpu_ctx)
    solvers = [Solver(m, 1e-3, gn_refine=4, options=_opts(refine_steps=1, refine_tol=1e-2)) for m in models]
    for d, s in zip(c4, solvers):
        s.measurement[0] = torch.tensor(d["o"].flatten_voltages(add_noise(d["U"], 60.0, frame_seed(4, d["j"]))))
    batch = Batch(models)
    bs = BatchSolver(batch, solvers)
    bs.solve_absolute()
    gb = [s.material_distribution.clone() for s in solvers]
    for s in solvers:
        s.material_distribution.fill_(1.0)
        s.solve_absolute()
    for a, s in zip(gb, solvers):
        d = float((a - s.material_distribution).abs().max() / (s.material_distribution - 1.0).abs().max())
        print("batched vs single GN update rel diff", d)
        assert d < 1e-5
    batch.close()
