"""Device forward solves on meshes from the force-equilibrium mesher (irregular node
placement, graded edge lengths, electrode corners fixed): voltages and Jacobian against
the FP64 oracle on the same mesh, P1 and promoted P2, 2.5D."""
import math

import numpy as np
import pytest

from mpflow_amd.mesh import boundary_edges, circular_points, distmesh_disc, promote_p2
from mpflow_amd.phantom import ellipse_phantom, frame_seed
from oracle import eit_oracle as eo

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("p2", [False, True])
def test_distmesh_forward_matches_oracle(gpu_ctx, p2):
    import torch
    from mpflow_amd import capi
    from mpflow_amd.eit import EIT, SolveOptions
    L, w = 16, 0.1
    d = 2 * math.pi / L
    corners = np.concatenate([circular_points(1.0, d, 0.0, False), circular_points(1.0, d, w, False)])
    nodes, el = distmesh_disc(1.0, 0.06, 0.1, corners)
    if p2:
        nodes, el = promote_p2(nodes, el)
    bd = boundary_edges(el)
    ports = eo.circular_boundary_ports(nodes, el, bd, L, w).astype(np.int32)
    pat = eo.adjacent_pattern(L)
    g = ellipse_phantom(nodes, el, frame_seed(1, 77))
    o = eo.EITOracle(nodes, el, bd, ports, pat, pat, 1e-3, 1.0, 3, 0.4, 0.1, quirks=False)
    U, phis, _ = o.solve(g)
    Jo = o.jacobian(g, phis, quirks=False)
    print("case: distmesh mesh p2=%s (%d nodes, %d elements)" % (p2, nodes.shape[0], el.shape[0]), flush=True)
    eit = EIT(nodes, el, bd, ports, pat, pat, 1e-3, 1.0, 3, 0.4, 0.1, quirks=False, jacobian_exact=True,
              ctx=gpu_ctx)
    eit.solve(torch.tensor(g, dtype=torch.float32),
              SolveOptions(stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-6, precision=capi.MPF_PREC_MIXED,
                           refine_steps=2, refine_tol=1e-3), jacobian=True)
    Ug = eit.voltages().cpu().numpy()[0]
    Jg = eit.jacobian_matrix().cpu().numpy()[0]
    eu = float(np.max(np.abs(Ug - U)) / np.max(np.abs(U)))
    ej = float(np.linalg.norm(Jg - Jo) / np.linalg.norm(Jo))
    print("voltages %.2e, jacobian %.2e" % (eu, ej))
    assert eit.last_stats.not_converged == 0
    assert eu < 1e-7 and ej < 1e-5
