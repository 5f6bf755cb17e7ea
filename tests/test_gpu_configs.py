"""The benchmark configurations that the other GPU tests cover only in part, at their
stated shapes, against the FP64 oracle:
  * config 2: 3 Gauss-Newton steps on the 40-ring P1 mesh (9600 elements), 32 electrodes,
    5 harmonics, complete electrode model with z_c = 0.01, 60 dB noise;
  * config 5: the 1024 -> 2048 -> 2048 -> 2048 -> E network on the config-2 mesh --
    gradients against the FP64 network oracle, and the learned start's gain;
  * config 4: all 256 uneven jobs in one batch -- every (job, harmonic) converged and a
    sample of jobs against the oracle."""
import numpy as np
import pytest

from mpflow_amd.phantom import add_noise, ellipse_phantom, frame_seed
from oracle.eit_oracle import EITOracle, reconstruct

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]


@pytest.fixture(scope="module")
def c2():
    from bench import WORKLOADS, build_problem
    w = WORKLOADS["c2"]
    nodes, el, bd, ports, pat = build_problem(w)
    return dict(w=w, nodes=nodes, el=el, bd=bd, ports=ports, pat=pat)


def test_c2_cem_reconstruction_3_steps(gpu_ctx, c2):
    import torch
    from mpflow_amd import capi
    from mpflow_amd.eit import EIT, SolveOptions, Solver
    w = c2["w"]
    z = 0.01
    print("case: c2 CEM oracle 3-step reconstruction", flush=True)
    o = EITOracle(c2["nodes"], c2["el"], c2["bd"], c2["ports"], c2["pat"], c2["pat"], 1e-3, 1.0, w["C"],
                  w["height"], w["port_height"], contact_impedance=z)
    g = ellipse_phantom(c2["nodes"], c2["el"], frame_seed(2, 0))
    U, _, _ = o.solve(g)
    Um = add_noise(U, 60.0, frame_seed(2, 0))
    g_ref = reconstruct(o, Um, np.ones(c2["el"].shape[0]), 1e-3, steps=3, quirks_jacobian=False)
    print("case: c2 CEM device 3-step reconstruction", flush=True)
    eit = EIT(c2["nodes"], c2["el"], c2["bd"], c2["ports"], c2["pat"], c2["pat"], 1e-3, 1.0, w["C"],
              w["height"], w["port_height"], jacobian_exact=True, contact_impedance=z, ctx=gpu_ctx)
    for name, opts in [("fp32", SolveOptions(stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-6)),
                       ("mixed", SolveOptions(stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-6,
                                              precision=capi.MPF_PREC_MIXED, refine_steps=1, refine_tol=1e-2))]:
        s = Solver(eit, 1e-3, gn_refine=4, options=opts)
        s.measurement[0] = torch.tensor(o.flatten_voltages(Um))
        for k in range(3):
            s.solve_absolute()
            assert s.last_stats.not_converged == 0
        gd = s.material_distribution.cpu().numpy()[0].astype(np.float64)
        err = np.linalg.norm(gd - g_ref) / np.linalg.norm(g_ref)
        err_d = np.linalg.norm((gd - 1) - (g_ref - 1)) / np.linalg.norm(g_ref - 1)
        print("c2 CEM %s: gamma rel L2 error %.3e, update part %.3e" % (name, err, err_d), flush=True)
        assert err < 1e-3
        assert err_d < 1e-2


def test_c5_network_shape_gradients(gpu_ctx, c2):
    """The config-5 network (1024 -> 2048 x 3 GELU -> E = 9600) forward and backward
    against the FP64 oracle with the device's BF16 operand rounding."""
    import torch
    from mpflow_amd.learned import Mlp
    from oracle.mlp_oracle import MlpOracle
    E = c2["el"].shape[0]
    dims = [1024, 2048, 2048, 2048, E]
    print("case: c5 network gradients dims=%s" % dims, flush=True)
    net = Mlp(dims, seed=5, ctx=gpu_ctx)
    rng = np.random.default_rng(21)
    B = 16
    X = rng.standard_normal((B, 1024)).astype(np.float32)
    T = (0.3 * rng.standard_normal((B, E))).astype(np.float32)
    ora = MlpOracle(dims, net.params())
    Y = net.forward(torch.tensor(X, device="cuda")).cpu().numpy()
    Yo, _ = ora.forward(X, bf16=True)
    loss = net.backward(torch.tensor(T, device="cuda"))
    g = net.params(grads=True)
    lo, go = ora.grads_flat(X, T, bf16=True)
    ey = float(np.linalg.norm(Y - Yo) / np.linalg.norm(Yo))
    eg = float(np.linalg.norm(g - go) / np.linalg.norm(go))
    lx, gx = ora.grads_flat(X, T, bf16=False)
    ex = float(np.linalg.norm(g - gx) / np.linalg.norm(gx))
    print("forward %.2e, loss %.6f vs %.6f, grads %.2e (vs the unrounded FP64 network %.2e)"
          % (ey, loss, lo, eg, ex))
    # where the FP32 accumulation order moves a sum across a BF16 rounding boundary, the
    # device and the oracle round an activation to neighbouring BF16 values (2^-8 apart);
    # over four layers of width 2048 such differences feed each other, so the bound is
    # looser than for the two-hidden-layer networks of test_gpu_mlp.py (measured: forward
    # 5.7e-4, gradients 1.2e-3), and the device stays several times closer to the BF16
    # oracle than to the network without BF16 rounding (5.1e-3)
    assert ey < 2e-3
    assert abs(loss - lo) < 1e-4 * lo
    assert eg < 3e-3
    assert ex > 3 * eg


def test_c5_learned_start_gain_c2_mesh(gpu_ctx, c2):
    """Config 5 at its network shape on the config-2 mesh: trained on device-generated
    pairs (1-4 ellipses, 60 dB), the network start followed by one Gauss-Newton step
    reconstructs held-out frames better than one plain step from gamma = 1, and the
    accepted frames' final voltage residual is no worse than 3-step plain GN's."""
    import torch
    from mpflow_amd import capi
    from mpflow_amd.eit import EIT, SolveOptions
    from mpflow_amd.learned import LearnedReconstruction, Mlp, VoltageFeatures, generate, train
    w = c2["w"]
    F = 64
    eit = EIT(c2["nodes"], c2["el"], c2["bd"], c2["ports"], c2["pat"], c2["pat"], 1e-3, 1.0, w["C"], w["height"],
              w["port_height"], jacobian_exact=True, frames=F, ctx=gpu_ctx)
    phantom = lambda s: ellipse_phantom(c2["nodes"], c2["el"], s, n_max=4)
    two = SolveOptions(stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-5, preconditioner=capi.MPF_PRECOND_TWO_LEVEL)
    print("case: c5 data on the c2 mesh", flush=True)
    U, G = generate(eit, 8192, frame_seed(5, 0), phantom, options=two)
    Uh, Gh = generate(eit, F, frame_seed(5, 1000000), phantom, options=two)
    feats = VoltageFeatures().fit(U)
    net = Mlp([eit.m, 2048, 2048, 2048, eit.E], seed=5, ctx=gpu_ctx)
    print("case: c5 training", flush=True)
    hist = train(net, feats(U), G, epochs=8, batch=256, lr=1e-3, lr_decay=0.85)
    print("train loss %.4f -> %.4f" % (hist[0], hist[-1]))
    opts = SolveOptions(stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-6, precision=capi.MPF_PREC_MIXED,
                        refine_steps=1, refine_tol=1e-2, preconditioner=capi.MPF_PRECOND_TWO_LEVEL)
    rec = LearnedReconstruction(eit, net, feats, 1e-3, opts, gn_steps=1, plain_steps=3)
    gl, info = rec.reconstruct(Uh, compare_plain=True)
    gp1 = rec._gn(torch.tensor(Uh, device="cuda"), 1)
    gtrue = np.exp(Gh)
    el_ = np.linalg.norm(gl.cpu().numpy() - gtrue, axis=1) / np.linalg.norm(gtrue, axis=1)
    ep_ = np.linalg.norm(gp1.cpu().numpy() - gtrue, axis=1) / np.linalg.norm(gtrue, axis=1)
    acc = info["accepted"]
    print("learned+1GN error %.4f, plain 1GN error %.4f, accepted %d/%d"
          % (el_.mean(), ep_.mean(), acc.sum(), F))
    assert el_.mean() < ep_.mean()
    assert acc.mean() >= 0.5
    assert np.all(info["residual_learned"][acc] <= info["residual_plain"][acc])


def test_c4_all_256_jobs(gpu_ctx):
    """Config 4 at its stated size: 256 jobs (2k-40k P1 elements, 16 or 32 electrodes,
    1-5 harmonics, 5.3M elements) solved together, two-level PCG with HIP-graph replay and
    FP64 refinement: every (job, harmonic) converges, and the four smallest and the largest
    job match the FP64 oracle."""
    import torch
    from bench import build_problem, c4_job
    from mpflow_amd import capi
    from mpflow_amd.eit import EIT, Batch, SolveOptions
    print("case: c4 256-job set-up", flush=True)
    models, probs = [], []
    for j in range(256):
        w = c4_job(j)
        nodes, el, bd, ports, pat = build_problem(w)
        models.append(EIT(nodes, el, bd, ports, pat, pat, 1e-3, 1.0, w["C"], w["height"], w["port_height"],
                          jacobian_exact=True, ctx=gpu_ctx))
        probs.append((j, w, nodes, el, bd, ports, pat))
    assert sum(m.E for m in models) > 5_000_000
    batch = Batch(models)
    opts = SolveOptions(stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-6, precision=capi.MPF_PREC_MIXED,
                        refine_steps=2, refine_tol=1e-3, graphs=1, preconditioner=capi.MPF_PRECOND_TWO_LEVEL)
    gs = [torch.tensor(ellipse_phantom(n, e, frame_seed(4, j)), dtype=torch.float32) for j, _, n, e, _, _, _ in probs]
    print("case: c4 256-job batched solve", flush=True)
    batch.solve(gs, opts)
    st = batch.last_stats
    print("steps %d, max steps %d, not converged %d, refine residual %.2e"
          % (st.steps, st.max_steps, st.not_converged, st.refine_residual))
    assert st.not_converged == 0
    assert st.refine_residual < 1e-6
    order = np.argsort([m.E for m in models])
    for i in list(order[:4]) + [order[-1]]:
        j, w, nodes, el, bd, ports, pat = probs[i]
        print("case: c4 job %d oracle (%d elements, %d electrodes, %d harmonics)" % (j, el.shape[0], w["L"], w["C"]),
              flush=True)
        o = EITOracle(nodes, el, bd, ports, pat, pat, 1e-3, 1.0, w["C"], w["height"], w["port_height"])
        U, _, _ = o.solve(gs[i].numpy().astype(np.float64))
        Ud = models[i].result64.cpu().numpy()[0].reshape(models[i].Dp, models[i].Mp)[:models[i].D, :models[i].M].T
        err = float(np.max(np.abs(Ud - U)) / np.max(np.abs(U)))
        print("job %d voltage rel err %.2e" % (j, err))
        assert err < 1e-5
    batch.close()
