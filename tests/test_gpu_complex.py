"""Complex conductivity on the device (models::EIT over Equation<thrust::complex<float>>,
src/models/eit.cpp:289-304; ConjugateGradient<thrust::complex<float>>,
src/numeric/conjugate_gradient.cpp:231-232; Inverse<thrust::complex<float>>,
src/solver/inverse.cpp:243-248) against the FP64 complex oracle: voltages, Jacobian
(reference formula and exact 2.5D), both CG variants, a Gauss-Newton step, several
frames, and the entry points that refuse a complex model."""
import ctypes as C

import numpy as np
import pytest

from tests.problems import make_problem

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _problem(C=1, h=1.0, hp=1.0, p2=False, rings=10, L=16, w=0.1):
    from oracle.eit_oracle import ComplexEITOracle
    p = make_problem(rings, L, w, C=C, height=h, port_height=hp, p2=p2)
    o = ComplexEITOracle(p["nodes"], p["elements"], p["boundary"], p["ports"], p["drive"], p["meas"],
                         p["current"], 1.0, C, h, hp)
    return p, o


def _eit(p, ctx, frames=1, exact=False):
    from mpflow_amd.eit import EIT
    return EIT(p["nodes"], p["elements"], p["boundary"], p["ports"], p["drive"], p["meas"], p["current"], 1.0,
               p["C"], p["height"], p["port_height"], jacobian_exact=exact, frames=frames, ctx=ctx,
               complex_gamma=True)


def _gamma(E, seed, imag):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.5, 2.0, E) + 1j * imag * rng.uniform(-1.0, 1.0, E)


def _opts(**kw):
    from mpflow_amd import capi
    from mpflow_amd.eit import SolveOptions
    base = dict(stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-7, max_iter=4000)
    base.update(kw)
    return SolveOptions(**base)


@pytest.mark.parametrize("C,h,hp,p2,imag", [(1, 1.0, 1.0, False, 0.3), (3, 0.4, 0.1, False, 1.0),
                                             (2, 0.5, 0.2, True, 0.3)])
def test_complex_forward_and_jacobian(gpu_ctx, C, h, hp, p2, imag):
    import torch
    print("case: complex forward C=%d p2=%d imag=%g" % (C, p2, imag), flush=True)
    p, o = _problem(C, h, hp, p2)
    g = _gamma(o.E, 11, imag)
    U0, ph, _ = o.solve(g)
    for exact in (False, True):
        eit = _eit(p, gpu_ctx, exact=exact)
        eit.solve(torch.tensor(g, dtype=torch.complex64), _opts(), jacobian=True)
        st = eit.last_stats
        U = eit.voltages(eit.result64)[0].cpu().numpy()
        err = _rel(U, U0)
        J = eit.jacobian_matrix()[0].cpu().numpy().astype(np.complex128)
        J0 = o.jacobian(g, ph, quirks=not exact)
        errj = _rel(J, J0)
        print("exact", exact, "steps", st.max_steps, "U rel err", err, "J rel err", errj)
        assert st.not_converged == 0
        assert err < 2e-5
        assert errj < 2e-4
        eit.close()


def test_complex_reference_iteration_matches_restatement(gpu_ctx):
    """complex_hermitian=1 runs the reference's conjugated iteration: with a small
    imaginary part it converges, and its reference-rule step count follows the FP64
    restatement's."""
    import torch
    from mpflow_amd import capi
    from oracle.eit_oracle import pcg_reference
    p, o = _problem(1)
    g = _gamma(o.E, 3, 0.05)
    U0, _, _ = o.solve(g)
    A = o.system(g, 0)
    F = (o.B @ o.Pi).astype(complex)
    _, s_ref = pcg_reference(A, F, 1.0 / A.diagonal(), True)
    eit = _eit(p, gpu_ctx)
    eit.solve(torch.tensor(g, dtype=torch.complex64), _opts(stop_rule=capi.MPF_STOP_REFERENCE, complex_hermitian=1))
    st = eit.last_stats
    err = _rel(eit.voltages(eit.result64)[0].cpu().numpy(), U0)
    print("hermitian: device steps", st.max_steps, "restatement", s_ref, "U rel err", err)
    assert st.not_converged == 0
    assert abs(st.max_steps - s_ref) <= 0.15 * s_ref + 8
    assert err < 1e-4
    # the default (conjugate orthogonal) iteration needs fewer steps on the same system
    eit.solve(torch.tensor(g, dtype=torch.complex64), _opts(stop_rule=capi.MPF_STOP_REFERENCE))
    print("cocg steps", eit.last_stats.max_steps)
    assert eit.last_stats.max_steps < st.max_steps
    eit.close()


def test_complex_gn_step_and_frames(gpu_ctx):
    """Two frames with different phantoms, one Gauss-Newton step each from gamma = 1:
    the step equals the oracle's data-space step (plain transposes) from the same start."""
    import torch
    from mpflow_amd.eit import Solver
    from oracle.eit_oracle import gn_step_exact
    p, o = _problem(2, 0.5, 0.2)
    lam = 1e-2
    eit = _eit(p, gpu_ctx, frames=2, exact=True)
    s = Solver(eit, lam, options=_opts())
    gt = [1.0 + _gamma(o.E, 20 + f, 0.4) * 0.2 for f in range(2)]
    um = np.stack([o.flatten_voltages(o.solve(g)[0]) for g in gt])
    s.measurement.copy_(torch.tensor(um, dtype=torch.complex128))
    s.solve_absolute()
    assert s.last_stats.not_converged == 0
    out = s.material_distribution.cpu().numpy().astype(np.complex128)
    g1 = np.ones(o.E, complex)
    U1, ph1, _ = o.solve(g1)
    J1 = o.jacobian(g1, ph1, quirks=False)
    for f in range(2):
        d0 = gn_step_exact(J1, um[f] - o.flatten_voltages(U1), (1.0 * lam) ** 2)
        errd = float(np.linalg.norm(out[f] - 1.0 - d0) / np.linalg.norm(d0))
        print("frame", f, "delta rel err", errd)
        assert errd < 1e-3
    np.testing.assert_allclose(s.calculation.cpu().numpy()[0], o.flatten_voltages(U1), rtol=0,
                               atol=2e-5 * np.abs(U1).max())
    eit.close()


def test_complex_model_refusals(gpu_ctx):
    import torch
    from mpflow_amd import capi
    from mpflow_amd.capi import lib
    p, o = _problem(1, rings=6, L=8, w=0.3)
    eit = _eit(p, gpu_ctx)
    g = torch.ones(o.E, dtype=torch.complex64, device="cuda")
    with pytest.raises(capi.MpfError) as e:
        eit.solve(g, _opts(krylov=capi.MPF_KRYLOV_BICGSTAB))
    assert e.value.code == capi.MPF_EINVAL
    with pytest.raises(capi.MpfError):
        eit.solve(g, _opts(precision=capi.MPF_PREC_MIXED))
    vals = torch.empty(eit.N, 32, dtype=torch.float32, device="cuda")
    ids = torch.empty(eit.N, 32, dtype=torch.int32, device="cuda")
    rc = lib().mpf_fem_assemble(eit.h, C.c_void_p(g.data_ptr()), 0.0, C.c_void_p(ids.data_ptr()),
                                C.c_void_p(vals.data_ptr()))
    assert rc == capi.MPF_EINVAL
    eit.close()
