"""Shared problem definitions for the tests: identical meshes, ports and patterns are
handed to the FP64 oracle and to the device path."""
import numpy as np

from mpflow_amd.mesh import boundary_edges, promote_p2, ring_mesh
from oracle.eit_oracle import EITOracle, adjacent_pattern, circular_boundary_ports


def make_problem(n_rings=16, L=16, width=0.1, C=1, height=1.0, port_height=1.0, p2=False,
                 current=1e-3):
    nodes, el = ring_mesh(n_rings, L, width)
    if p2:
        nodes, el = promote_p2(nodes, el)
    bd = boundary_edges(el)
    ports = circular_boundary_ports(nodes, el, bd, L, width)
    pat = adjacent_pattern(L)
    return dict(nodes=nodes, elements=el, boundary=bd, ports=ports.astype(np.int32), drive=pat,
                meas=pat, current=current, C=C, height=height, port_height=port_height)


def oracle_of(p, quirks=True, contact_impedance=0.0):
    return EITOracle(p["nodes"], p["elements"], p["boundary"], p["ports"], p["drive"], p["meas"],
                     p["current"], 1.0, p["C"], p["height"], p["port_height"], quirks=quirks,
                     contact_impedance=contact_impedance)


def eit_of(p, ctx=None, frames=1, quirks=True, jacobian_exact=False, keep_node_order=False,
           contact_impedance=0.0):
    from mpflow_amd.eit import EIT
    return EIT(p["nodes"], p["elements"], p["boundary"], p["ports"], p["drive"], p["meas"],
               p["current"], 1.0, p["C"], p["height"], p["port_height"], quirks=quirks,
               jacobian_exact=jacobian_exact, frames=frames, ctx=ctx, keep_node_order=keep_node_order,
               contact_impedance=contact_impedance)


def rcb_renumbered(p, leaf=256):
    """The problem with its nodes renumbered by recursive coordinate bisection into
    leaves of ``leaf`` consecutive nodes, each bisected into four runs of leaf / 4 (compact
    aggregates for the two-level preconditioner when the device keeps the caller's node
    order)."""
    from oracle.eit_oracle import rcb_order
    order = rcb_order(p["nodes"], leaf, leaf // 4)
    perm = np.empty_like(order)
    perm[order] = np.arange(order.size)
    q = dict(p)
    q["nodes"] = p["nodes"][order]
    q["elements"] = perm[p["elements"]].astype(p["elements"].dtype)
    q["boundary"] = perm[p["boundary"]].astype(p["boundary"].dtype)
    return q
