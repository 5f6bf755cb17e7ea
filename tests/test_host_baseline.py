"""The host baseline of whole reconstruction frames (oracle/gn_host.cpp, the timed
cpu_baseline of bench.py) against the FP64 oracle's Gauss-Newton reconstruction: the
same algorithm in FP32 with FP64 scalars and refinement must land on the oracle's result."""
import numpy as np
import pytest

from mpflow_amd.phantom import add_noise, ellipse_phantom, frame_seed
from oracle.eit_oracle import reconstruct
from oracle.gn_host import GnHost
from tests.problems import make_problem, oracle_of


@pytest.mark.parametrize("z,two_level,cem", [(0.0, True, "harmonic"), (0.0, False, "harmonic"),
                                            (0.01, True, "harmonic"), (0.01, True, "coupled"),
                                            (0.01, False, "coupled")])
def test_host_reconstruction_matches_oracle(z, two_level, cem):
    from oracle.eit_oracle import CoupledCEMOracle
    p = make_problem(10, 16, 0.1, C=3, height=0.4, port_height=0.1)
    if cem == "coupled":
        o = CoupledCEMOracle(p["nodes"], p["elements"], p["boundary"], p["ports"], p["drive"], p["meas"],
                             p["current"], 1.0, p["C"], p["height"], p["port_height"], contact_impedance=z)
    else:
        o = oracle_of(p, quirks=False, contact_impedance=z)
    frames = []
    for f in range(2):
        g = ellipse_phantom(p["nodes"], p["elements"], frame_seed(2, f))
        U, _, _ = o.solve(g)
        frames.append(add_noise(U, 60.0, frame_seed(2, f)))
    h = GnHost(p["nodes"], p["elements"], p["boundary"], p["ports"], p["drive"], p["meas"], p["current"], 1.0,
               p["C"], p["height"], p["port_height"], contact_impedance=z, two_level=two_level, tol=1e-6,
               refine=1, refine_tol=1e-2, cem=cem)
    Um = np.stack([o.flatten_voltages(U) for U in frames])
    G, secs = h.reconstruct(Um, steps=3, threads=2)
    for f, U in enumerate(frames):
        g_ref = reconstruct(o, U, np.ones(o.E), 1e-3, steps=3, quirks_jacobian=False)
        err = np.linalg.norm(G[f] - g_ref) / np.linalg.norm(g_ref)
        err_d = np.linalg.norm(G[f] - 1 - (g_ref - 1)) / np.linalg.norm(g_ref - 1)
        print("frame %d: gamma %.2e, update %.2e (%.2f s)" % (f, err, err_d, secs))
        assert err < 1e-4
        assert err_d < 1e-2
