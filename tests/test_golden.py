"""Golden fixtures (tests/golden/*): the oracle reproduces them on the CPU, and the
device path matches them on the GPU (voltages 1e-5 max-norm relative in FP32 PCG,
Jacobian 1e-4 relative L2)."""
import os

import numpy as np
import pytest

from tests.golden.make_golden import CASES, case_gamma, case_problem
from tests.problems import eit_of, oracle_of

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name, f):
    return np.loadtxt(os.path.join(HERE, name, f), ndmin=2)


def _cols(o):
    rows = [d * o.Mp + m for d in range(o.D) for m in range(o.M)]
    return rows, np.arange(0, o.E, 4)


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_golden(name):
    p, q = case_problem(name)
    np.testing.assert_array_equal(p["elements"], _load(name, "elements.txt").astype(np.int32))
    np.testing.assert_allclose(p["nodes"], _load(name, "nodes.txt"), rtol=0, atol=1e-15)
    np.testing.assert_array_equal(p["ports"], _load(name, "ports.txt").astype(np.int32))
    o = oracle_of(p, quirks=q)
    g = case_gamma(name, p)
    np.testing.assert_allclose(g, _load(name, "gamma.txt")[:, 0] if _load(name, "gamma.txt").shape[1] == 1
                               else _load(name, "gamma.txt").ravel())
    U, phis, _ = o.solve(g)
    np.testing.assert_allclose(U, _load(name, "U.txt"), rtol=1e-10, atol=1e-16)
    rows, cols = _cols(o)
    Jr = o.jacobian(g, phis, quirks=True)[np.ix_(rows, cols)]
    np.testing.assert_allclose(Jr, _load(name, "J_reference.txt"), rtol=1e-8, atol=1e-18)
    if o.C > 1:
        Jx = o.jacobian(g, phis, quirks=False)[np.ix_(rows, cols)]
        np.testing.assert_allclose(Jx, _load(name, "J_exact.txt"), rtol=1e-8, atol=1e-18)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_device_matches_golden(gpu_ctx, name):
    import torch
    p, q = case_problem(name)
    o = oracle_of(p, quirks=q)
    g = _load(name, "gamma.txt").ravel()
    print("case: golden %s" % name, flush=True)
    for exact in ((False, True) if o.C > 1 else (False,)):
        eit = eit_of(p, gpu_ctx, quirks=q, jacobian_exact=exact)
        eit.solve(torch.tensor(g, dtype=torch.float32), jacobian=True)
        U = eit.voltages().cpu().numpy()[0]
        Uref = _load(name, "U.txt")
        assert np.max(np.abs(U - Uref)) / np.max(np.abs(Uref)) < 1e-5
        rows, cols = _cols(o)
        J = eit.jacobian_matrix().cpu().numpy()[0][np.ix_(rows, cols)]
        Jref = _load(name, "J_exact.txt" if exact else "J_reference.txt")
        assert np.linalg.norm(J - Jref) / np.linalg.norm(Jref) < 1e-4
