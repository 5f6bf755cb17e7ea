"""Mesh inputs for the EIT path: concentric-ring disc meshes, force-equilibrium disc
meshes with fixed electrode corners, P1->P2 promotion, boundary edges and the text
formats of ``IrregularMesh``.

The reference meshes with libdistmesh (``src/numeric/irregular_mesh.cpp:87-122``), which
is not available (an unpinned third-party library).  ``distmesh_disc`` restates the
published algorithm it implements (Persson & Strang, "A Simple Mesh Generator in
MATLAB", SIAM Review 46(2), 2004) for the reference's call: a disc, graded edge length,
electrode corners as fixed points.  The benchmark configurations use the deterministic
ring generator below so that the GPU path, the CPU baseline and the FP64 oracle all see
byte-identical ``nodes``/``elements`` arrays.

Ring generator: ring ``j`` (1..n_r) of radius ``j/n_r`` holds ``6j`` equally spaced
nodes; the centre is node 0.  Neighbouring rings are stitched by an angular merge,
giving ``V = 1 + 3 n_r (n_r + 1)`` nodes and ``E = 6 n_r^2`` counter-clockwise
triangles.  Before stitching, the boundary node nearest to each electrode corner is
moved onto that corner, so electrode arcs start and end on mesh nodes.
"""
from __future__ import annotations

import math
import os

import numpy as np

__all__ = [
    "ring_mesh", "promote_p2", "boundary_edges", "unique_edges",
    "port_corner_angles", "save_mesh", "load_mesh", "rings_for_elements", "distmesh_disc",
    "circular_points",
]


def port_corner_angles(count: int, width: float, offset: float = 0.0):
    """Start/end angles of ``count`` electrodes of angular ``width`` on the unit circle.

    Electrode ``l`` spans ``[offset + 2 pi l / count, offset + 2 pi l / count + width]``,
    counter-clockwise, as ``math::circularPoints`` lays them out for
    ``Ports::circularBoundary`` (``src/fem/ports.cpp:43-46``, ``src/mathematics.cpp:91-99``).
    """
    base = offset + 2.0 * math.pi * np.arange(count) / count
    return base, base + width


def rings_for_elements(n_elements: int) -> int:
    """Ring count whose mesh has about ``n_elements`` triangles (E = 6 n_r^2)."""
    return max(2, int(round(math.sqrt(n_elements / 6.0))))


def _snap_boundary(theta: np.ndarray, corners: np.ndarray) -> np.ndarray:
    """Move the boundary node nearest each corner angle onto that corner.

    ``theta`` holds the sorted boundary angles in [0, 2 pi).  Each corner takes a
    distinct node; the angular order of the boundary is kept (checked).
    """
    n = theta.size
    theta = theta.copy()
    taken = np.zeros(n, dtype=bool)
    order = np.argsort(np.mod(corners, 2 * math.pi), kind="stable")
    for c in np.mod(corners, 2 * math.pi)[order]:
        d = np.abs(np.angle(np.exp(1j * (theta - c))))
        d[taken] = np.inf
        i = int(np.argmin(d))
        taken[i] = True
        theta[i] = c
    # node 0 may have been moved to an angle near 2 pi; keep the ring sorted
    # starting from the node closest to angle 0
    t = np.mod(theta, 2 * math.pi)
    if np.any(np.diff(t) <= 0.0):
        raise ValueError("electrode corner snapping broke the boundary node order; "
                         "use more rings or narrower electrodes")
    return t


def ring_mesh(n_rings: int, port_count: int | None = None, port_width: float | None = None,
              port_offset: float = 0.0, radius: float = 1.0):
    """Concentric-ring triangulation of a disc.

    Returns ``(nodes[V,2] float64, elements[E,3] int32)``; triangles are
    counter-clockwise.  If ``port_count``/``port_width`` are given, boundary nodes are
    snapped to the electrode corners (nearest node per corner).
    """
    if n_rings < 1:
        raise ValueError("n_rings < 1")
    rings_theta = [np.zeros(1)]
    for j in range(1, n_rings + 1):
        rings_theta.append(2.0 * math.pi * np.arange(6 * j) / (6 * j))
    if port_count:
        s, e = port_corner_angles(port_count, port_width, port_offset)
        rings_theta[-1] = _snap_boundary(rings_theta[-1], np.concatenate([s, e]))

    nodes = [np.zeros((1, 2))]
    start = [0]
    off = 1
    for j in range(1, n_rings + 1):
        r = radius * j / n_rings
        th = rings_theta[j]
        nodes.append(np.stack([r * np.cos(th), r * np.sin(th)], axis=1))
        start.append(off)
        off += th.size
    nodes = np.concatenate(nodes, axis=0)

    tris = []
    # centre fan
    n1 = rings_theta[1].size
    for jb in range(n1):
        tris.append((0, start[1] + jb, start[1] + (jb + 1) % n1))
    # merge neighbouring rings by angle
    for j in range(2, n_rings + 1):
        A, B = rings_theta[j - 1], rings_theta[j]
        na, nb = A.size, B.size
        sa, sb = start[j - 1], start[j]
        i = k = 0
        while i < na or k < nb:
            ta = A[i + 1] if i + 1 < na else 2 * math.pi + A[0]
            tb = B[k + 1] if k + 1 < nb else 2 * math.pi + B[0]
            if k < nb and (i >= na or tb <= ta):
                tris.append((sa + i % na, sb + k % nb, sb + (k + 1) % nb))
                k += 1
            else:
                tris.append((sa + i % na, sb + k % nb, sa + (i + 1) % na))
                i += 1
    elements = np.asarray(tris, dtype=np.int32)
    _check_ccw(nodes, elements)
    return nodes, elements


def circular_points(radius: float, distance: float, offset: float, invert_direction: bool):
    """``math::circularPoints`` (src/mathematics.cpp:79-99): points on the circle of
    ``radius`` spaced by the arc length ``distance``, starting at arc length ``offset``
    from angle 0, counter-clockwise (clockwise: y mirrored)."""
    n = int(math.ceil(2.0 * math.pi / (distance / radius)))
    phi = offset / radius + np.arange(n) * (distance / radius)
    sgn = -1.0 if invert_direction else 1.0
    return np.stack([radius * np.cos(phi), sgn * radius * np.sin(phi)], axis=1)


def distmesh_disc(radius: float, outer_edge_length: float, inner_edge_length: float,
                  fixed_points=None, seed: int = 0, max_iter: int = 1000):
    """Disc mesh by force equilibrium (Persson & Strang 2004), as the reference's
    ``IrregularMesh::fromConfig`` asks libdistmesh for it (irregular_mesh.cpp:87-122):

      * signed distance d(p) = |p| - radius;
      * relative edge length h(p) = 1 + (1 - inner / outer) d(p) / radius (outer edge
        length on the boundary, inner at the centre), initial spacing outer_edge_length;
      * bounding box [-1.1 radius, 1.1 radius]^2, ``fixed_points`` kept in place (the
        electrode corners, ``circular_points`` of the port configuration).

    Steps: an equilateral grid in the box, points outside removed, the rest kept with
    probability (min h / h(p))^2 (seeded generator); then repeated Delaunay
    triangulation, bar forces max(L0 - L, 0) with L0 = h(mid) * 1.2 * sqrt(sum L^2 /
    sum h(mid)^2), explicit steps of 0.2, points leaving the disc projected back onto the
    circle, until interior points move less than 1e-3 of the spacing.  Returns
    ``(nodes[V, 2], elements[E, 3] int32)``, triangles counter-clockwise, the fixed points
    first."""
    from scipy.spatial import Delaunay
    h0 = float(outer_edge_length)
    ratio = float(inner_edge_length) / h0
    fd = lambda p: np.hypot(p[:, 0], p[:, 1]) - radius
    fh = lambda p: 1.0 + (1.0 - ratio) * fd(p) / radius
    geps = 1e-3 * h0
    pfix = np.zeros((0, 2)) if fixed_points is None else np.asarray(fixed_points, np.float64).reshape(-1, 2)
    # 1. equilateral grid over the bounding box
    b = 1.1 * radius
    xs = np.arange(-b, b + 1e-12, h0)
    ys = np.arange(-b, b + 1e-12, h0 * math.sqrt(3.0) / 2.0)
    X, Y = np.meshgrid(xs, ys)
    X[1::2, :] += h0 / 2.0
    p = np.stack([X.ravel(), Y.ravel()], axis=1)
    # 2. inside points, thinned to the size function
    p = p[fd(p) < -geps]
    r0 = 1.0 / fh(p) ** 2
    rng = np.random.default_rng(seed)
    p = p[rng.random(p.shape[0]) < r0 / r0.max()]
    # drop grid points on top of fixed points
    if pfix.shape[0]:
        dmin = np.min(np.hypot(p[:, None, 0] - pfix[None, :, 0], p[:, None, 1] - pfix[None, :, 1]), axis=1)
        p = p[dmin > 0.5 * h0 * ratio]
    p = np.concatenate([pfix, p], axis=0)
    nfix = pfix.shape[0]
    pold = np.full_like(p, np.inf)
    bars = tri = None
    for _ in range(max_iter):
        if np.max(np.hypot(*(p - pold).T)) / h0 > 0.1:
            pold = p.copy()
            tri = Delaunay(p).simplices
            tri = tri[fd(p[tri].mean(axis=1)) < -geps]
            e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]], axis=0)
            bars = np.unique(np.sort(e, axis=1), axis=0)
        vec = p[bars[:, 0]] - p[bars[:, 1]]
        L = np.hypot(vec[:, 0], vec[:, 1])
        hb = fh(0.5 * (p[bars[:, 0]] + p[bars[:, 1]]))
        L0 = hb * 1.2 * math.sqrt(np.sum(L ** 2) / np.sum(hb ** 2))
        F = np.maximum(L0 - L, 0.0)
        Fv = (F / L)[:, None] * vec
        Ft = np.zeros_like(p)
        np.add.at(Ft, bars[:, 0], Fv)
        np.add.at(Ft, bars[:, 1], -Fv)
        Ft[:nfix] = 0.0
        p = p + 0.2 * Ft
        # back onto the circle (the exact projection for d = |p| - radius)
        r = np.hypot(p[:, 0], p[:, 1])
        out = r > radius
        p[out] *= (radius / r[out])[:, None]
        inside = fd(p) < -geps
        if np.max(np.hypot(*(0.2 * Ft[inside]).T), initial=0.0) / h0 < 1e-3:
            break
    tri = Delaunay(p).simplices
    tri = tri[fd(p[tri].mean(axis=1)) < -geps]
    # counter-clockwise, unused points removed (fixed points stay first)
    a, bb, c = p[tri[:, 0]], p[tri[:, 1]], p[tri[:, 2]]
    cross = (bb[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (bb[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    tri[cross < 0] = tri[cross < 0][:, [0, 2, 1]]
    used = np.zeros(p.shape[0], bool)
    used[tri.ravel()] = True
    used[:nfix] = True
    new = -np.ones(p.shape[0], np.int64)
    new[used] = np.arange(int(used.sum()))
    nodes = p[used]
    elements = new[tri].astype(np.int32)
    _check_ccw(nodes, elements)
    return nodes, elements


def _check_ccw(nodes, elements):
    p = nodes[elements[:, :3]]
    cross = (p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) - \
            (p[:, 1, 1] - p[:, 0, 1]) * (p[:, 2, 0] - p[:, 0, 0])
    if np.any(cross <= 0.0):
        raise ValueError("mesh has %d non-positive triangles" % int(np.sum(cross <= 0)))


def promote_p2(nodes: np.ndarray, elements: np.ndarray):
    """P1 -> P2 promotion with straight edge midpoints.

    Element rows become ``[v0, v1, v2, m(v0,v1), m(v1,v2), m(v2,v0)]`` and midpoint
    nodes are appended after the vertices in first-encounter order, the layout of the
    reference's (disabled) ``quadraticMeshFromLinear``
    (``src/numeric/irregular_mesh.cpp:229-300``).
    """
    V = nodes.shape[0]
    E = elements.shape[0]
    mid = {}
    new_xy = []
    el6 = np.zeros((E, 6), dtype=np.int32)
    el6[:, :3] = elements[:, :3]
    for e in range(E):
        for t in range(3):
            a, b = int(elements[e, t]), int(elements[e, (t + 1) % 3])
            key = (a, b) if a < b else (b, a)
            idx = mid.get(key)
            if idx is None:
                idx = V + len(new_xy)
                mid[key] = idx
                new_xy.append(0.5 * (nodes[a] + nodes[b]))
            el6[e, 3 + t] = idx
    nodes2 = np.concatenate([nodes, np.asarray(new_xy).reshape(-1, 2)], axis=0)
    return nodes2, el6


def unique_edges(elements: np.ndarray):
    """Unique vertex edges ``[a,b]`` (a<b) and the element->edge index table."""
    tri = elements[:, :3]
    pairs = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]], axis=0)
    key = np.sort(pairs, axis=1)
    uniq, inv = np.unique(key, axis=0, return_inverse=True)
    E = tri.shape[0]
    elem_edges = inv.reshape(3, E).T.astype(np.int32)
    return uniq.astype(np.int32), elem_edges


def boundary_edges(elements: np.ndarray):
    """Counter-clockwise boundary edges.

    P1 rows are ``[start, end]``; P2 rows are ``[start, mid, end]`` so that the edge
    parameter runs from node 0 to the last node, which is the order
    ``basis::Quadratic::boundaryIntegral`` integrates over
    (``src/fem/basis_quadratic.cpp:101-153``).  An edge is on the boundary when exactly
    one triangle has it; orientation follows that triangle (interior on the left).
    """
    k = elements.shape[1]
    tri = elements[:, :3]
    loc = [(0, 1, 3), (1, 2, 4), (2, 0, 5)]
    rows = []
    for (i, j, m) in loc:
        a, b = tri[:, i], tri[:, j]
        mm = elements[:, m] if k == 6 else None
        rows.append((a, b, mm))
    a = np.concatenate([r[0] for r in rows])
    b = np.concatenate([r[1] for r in rows])
    key = np.sort(np.stack([a, b], axis=1), axis=1)
    _, inv, cnt = np.unique(key, axis=0, return_inverse=True, return_counts=True)
    once = cnt[inv] == 1
    if k == 6:
        m = np.concatenate([r[2] for r in rows])
        out = np.stack([a[once], m[once], b[once]], axis=1)
    else:
        out = np.stack([a[once], b[once]], axis=1)
    return out.astype(np.int32)


def save_mesh(path: str, nodes: np.ndarray, elements: np.ndarray, edges: np.ndarray | None = None):
    """Write ``nodes.txt``/``elements.txt`` (+ ``edges.txt``) in ``Matrix::savetxt`` form
    (whitespace-separated rows, 16 significant digits; ``src/numeric/matrix.cpp:684-733``)."""
    os.makedirs(path, exist_ok=True)
    np.savetxt(os.path.join(path, "nodes.txt"), nodes, fmt="%.16g")
    np.savetxt(os.path.join(path, "elements.txt"), elements, fmt="%d")
    if edges is not None:
        np.savetxt(os.path.join(path, "edges.txt"), edges, fmt="%d")


def load_mesh(path: str):
    nodes = np.loadtxt(os.path.join(path, "nodes.txt"), dtype=np.float64, ndmin=2)
    elements = np.loadtxt(os.path.join(path, "elements.txt"), dtype=np.int64, ndmin=2).astype(np.int32)
    return nodes, elements
