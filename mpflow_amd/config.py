"""JSON model / solver configuration and the text-matrix files of the reference:
``Solver::fromConfig`` (src/solver/solver.cpp:70-135), ``EIT::fromConfig``
(src/models/eit.cpp:102-131), ``IrregularMesh::fromConfig``
(src/numeric/irregular_mesh.cpp:48-126), ``Ports::fromConfig`` (src/fem/ports.cpp:79-114),
``Sources::fromConfig`` (src/fem/sources.cpp:92-131) and ``Matrix::loadtxt / savetxt``
(src/numeric/matrix.cpp:684-822).

Schema (JSON integers are accepted wherever the reference reads doubles)::

    {"model": {
        "mesh": "meshdir" | {"path": "meshdir"}                  nodes/elements(/edges) text files
              | {"radius": r, "outerEdgeLength": h, "innerEdgeLength": h_i},
        "ports": "ports.txt" | {"edges": "ports.txt"}            rows of edge indices, -1 unused
               | {"count": P, "width": w, "offset": o, "clockwise": false, "height": hp,
                  "invertDirection": false},       the mesher's corner points read this key
        "source": {"type": "current", "value": I | [I_1..I_D],
                   "drivePattern": [[..]], "measurementPattern": [[..]]},   default eye(P)
        "material": ref | {"referenceValue": ref, "distribution": "gamma.txt"} | "gamma.txt",
        "2.5D": {"components": C, "height": h}},
     "solver": {"parallelImages": F, "regularizationFactor": lam,
                "regularizationType": "identity" | "noser" | "laplacian"}}

Differences from the reference, each raising or documented:
  * the distmesh branch runs this package's force-equilibrium mesher
    (``mesh.distmesh_disc``, Persson & Strang's algorithm with the reference's distance
    and edge-length functions and the electrode corners as fixed points); libdistmesh
    itself is unavailable, so node positions are not the library's.  The mesh is written
    to ``<path>/mesh`` as the reference does.
  * ``"type": "voltage"`` gives Fixed sources (``value``: the applied voltages), any other
    type current sources (sources.cpp:123-125).
  * a ``noSensitivity`` file name (solver.cpp:107-113) is read as an (m, E) text matrix
    and multiplied into the Jacobians of the noser weights and of preSolve.
  * port edge files index the mesh's edge list (``edges.txt``, or the unique vertex edges
    in ``mesh.unique_edges`` order when the file is absent); every port edge must lie on
    the boundary.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from . import mesh as meshmod

__all__ = ["savetxt", "loadtxt", "ModelSpec", "parse_config", "eit_from_config", "solver_from_config"]


def savetxt(filename: str, M, delimiter: str = " "):
    """Matrix::savetxt: one row per line, values separated by ``delimiter``, 16
    significant digits, no newline after the last row (matrix.cpp:684-733)."""
    A = np.asarray(M)
    if A.ndim == 1:
        A = A[:, None]
    integer = np.issubdtype(A.dtype, np.integer)
    fmt = (lambda v: "%d" % v) if integer else (lambda v: "%.16g" % v)
    with open(filename, "w") as f:
        f.write("\n".join(delimiter.join(fmt(v) for v in row) for row in A))


def loadtxt(filename: str, dtype=np.float64, delimiter: str = " ") -> np.ndarray:
    """Matrix::loadtxt: rows split at ``delimiter``; tokens that do not parse are skipped,
    empty lines are dropped, and an empty result is an error (matrix.cpp:735-822)."""
    integer = np.issubdtype(np.dtype(dtype), np.integer)

    def parse(tok):
        # stream extraction into an int stops at the decimal point: "2.5" reads as 2
        return int(tok) if not integer or tok.lstrip("+-").isdigit() else int(float(tok))

    rows = []
    with open(filename) as f:
        for line in f.read().split("\n"):
            row = []
            for tok in line.split(delimiter):
                try:
                    row.append(parse(tok) if integer else float(tok))
                except ValueError:
                    pass
            if row:
                rows.append(row)
    if not rows or not rows[0]:
        raise ValueError("loadtxt: cannot parse file: %s" % filename)
    ncol = len(rows[0])
    return np.array([r[:ncol] + [0] * (ncol - len(r)) for r in rows], dtype=dtype)


class ModelSpec:
    """Everything ``EIT`` needs, parsed from a configuration (host arrays)."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def _num(v, default):
    # jsonHelper::parseNumericValue (json_helper.cpp:26-35), integers accepted
    return float(v) if isinstance(v, (int, float)) and not isinstance(v, bool) else default


def _load_mesh_dir(d):
    nodes = loadtxt(os.path.join(d, "nodes.txt"), np.float64)
    elements = loadtxt(os.path.join(d, "elements.txt"), np.int64).astype(np.int32)
    ef = os.path.join(d, "edges.txt")
    if os.path.exists(ef):
        edges = loadtxt(ef, np.int64).astype(np.int32)
    else:
        edges, elem_edges = meshmod.unique_edges(elements)
        savetxt(ef, edges)
        savetxt(os.path.join(d, "elementEdges.txt"), elem_edges)
    return nodes, elements, edges


def _ports_from_edge_file(fname, edges, boundary):
    """Rows of mesh-edge indices (-1 unused) -> rows of indices into ``boundary``."""
    P = loadtxt(fname, np.int64)
    key = {tuple(sorted((int(b[0]), int(b[-1])))): i for i, b in enumerate(boundary)}
    out = np.full(P.shape, -1, np.int32)
    for p in range(P.shape[0]):
        for q in range(P.shape[1]):
            e = int(P[p, q])
            if e < 0 or e >= edges.shape[0]:
                continue
            k = tuple(sorted((int(edges[e][0]), int(edges[e][-1]))))
            if k not in key:
                raise ValueError("ports: edge %d of port %d is not a boundary edge" % (e, p))
            out[p, q] = key[k]
    return out


def parse_config(config, path: str = "."):
    """(ModelSpec, solver dict) from a configuration dict or a JSON file name.  Relative
    file names are taken relative to ``path`` (for a file name: its directory)."""
    if isinstance(config, str):
        if path == ".":
            path = os.path.dirname(os.path.abspath(config))
        with open(config) as f:
            config = json.load(f)
    model = config.get("model", {})
    solver = config.get("solver", {})
    mcfg, pcfg = model.get("mesh"), model.get("ports", {})
    if isinstance(mcfg, str) or (isinstance(mcfg, dict) and isinstance(mcfg.get("path"), str)):
        mesh_path = mcfg if isinstance(mcfg, str) else mcfg["path"]
        nodes, elements, edges = _load_mesh_dir(os.path.join(path, mesh_path))
    elif isinstance(mcfg, dict) and all(k in mcfg for k in ("radius", "outerEdgeLength", "innerEdgeLength")):
        mesh_path = "mesh"
        radius = float(mcfg["radius"])
        count = int(pcfg.get("count", 0)) if isinstance(pcfg, dict) else 0
        width = float(pcfg.get("width", 0.0)) if isinstance(pcfg, dict) else 0.0
        offset = float(pcfg.get("offset", 0.0)) if isinstance(pcfg, dict) else 0.0
        invert = bool(pcfg.get("invertDirection", False)) if isinstance(pcfg, dict) else False
        # the electrode corners are fixed points of the mesh (irregular_mesh.cpp:92-103):
        # circularPoints at the port offsets and offsets + width (arc lengths)
        fixed = None
        if count and width:
            dist = 2.0 * math.pi * radius / count
            fixed = np.concatenate([meshmod.circular_points(radius, dist, offset, invert),
                                    meshmod.circular_points(radius, dist, offset + width, invert)])
        nodes, elements = meshmod.distmesh_disc(radius, float(mcfg["outerEdgeLength"]), float(mcfg["innerEdgeLength"]),
                                                fixed)
        edges, elem_edges = meshmod.unique_edges(elements)
        d = os.path.join(path, "mesh")
        os.makedirs(d, exist_ok=True)
        savetxt(os.path.join(d, "nodes.txt"), nodes)
        savetxt(os.path.join(d, "elements.txt"), elements)
        savetxt(os.path.join(d, "edges.txt"), edges)
        savetxt(os.path.join(d, "elementEdges.txt"), elem_edges)
    else:
        raise ValueError("config: model.mesh must name a mesh directory or give radius / edge lengths")
    boundary = meshmod.boundary_edges(elements)

    if isinstance(pcfg, str) or (isinstance(pcfg, dict) and isinstance(pcfg.get("edges"), str)):
        fname = pcfg if isinstance(pcfg, str) else pcfg["edges"]
        ports = _ports_from_edge_file(os.path.join(path, fname), edges, boundary)
    elif isinstance(pcfg, dict) and "count" in pcfg and "width" in pcfg:
        from .eit import circular_ports
        ports = circular_ports(nodes, elements, boundary, int(pcfg["count"]), float(pcfg["width"]),
                               float(pcfg.get("offset", 0.0)), bool(pcfg.get("clockwise", False)))
        # the reference writes the generated port edges next to the mesh (ports.cpp:104-111),
        # as rows of indices into the mesh's edge list
        eidx = {tuple(sorted((int(a), int(b)))): i for i, (a, b) in enumerate(edges)}
        rows = np.full(ports.shape, -1, np.int64)
        for p in range(ports.shape[0]):
            for q in range(ports.shape[1]):
                if ports[p, q] >= 0:
                    b = boundary[ports[p, q]]
                    rows[p, q] = eidx[tuple(sorted((int(b[0]), int(b[-1]))))]
        os.makedirs(os.path.join(path, mesh_path), exist_ok=True)
        savetxt(os.path.join(path, mesh_path, "ports.txt"), rows)
    else:
        raise ValueError("config: model.ports must name a port edge file or give count and width")
    P = ports.shape[0]

    scfg = model.get("source", {})
    # sources.cpp:123-125: "voltage" is Fixed, anything else Open (current)
    source = "voltage" if scfg.get("type", "current") == "voltage" else "current"
    drive = np.asarray(scfg["drivePattern"], np.float64) if "drivePattern" in scfg else np.eye(P)
    meas = np.asarray(scfg["measurementPattern"], np.float64) if "measurementPattern" in scfg else np.eye(P)
    if drive.ndim == 1:
        drive = drive[:, None]
    if meas.ndim == 1:
        meas = meas[:, None]
    D = drive.shape[1]
    val = scfg.get("value", 1.0)
    current = np.asarray([float(v) for v in val][:D] if isinstance(val, list) else [_num(val, 1.0)] * D)

    mat = model.get("material", 1.0)
    gamma_ref = None
    if isinstance(mat, dict):
        ref = _num(mat.get("referenceValue"), 1.0)
        if isinstance(mat.get("distribution"), str):
            gamma_ref = loadtxt(os.path.join(path, mat["distribution"])).ravel()
    elif isinstance(mat, str):
        ref = 1.0
        gamma_ref = loadtxt(os.path.join(path, mat)).ravel()
    else:
        ref = _num(mat, 1.0)
    no_sens = None
    if isinstance(model.get("noSensitivity"), str):
        no_sens = loadtxt(os.path.join(path, model["noSensitivity"]))
    c25 = model.get("2.5D", {})
    spec = ModelSpec(nodes=nodes, elements=elements, boundary=boundary, ports=ports, drive=drive, meas=meas,
                     current=current, source=source, reference_value=ref,
                     components=max(1, int(c25.get("components", 1))),
                     height=_num(c25.get("height"), 1.0),
                     port_height=_num(pcfg.get("height"), 1.0) if isinstance(pcfg, dict) else 1.0,
                     gamma_ref=gamma_ref, edges=edges, no_sensitivity=no_sens)
    sol = {"parallel_images": max(1, int(solver.get("parallelImages", 1))),
           "regularization_factor": _num(solver.get("regularizationFactor"), 0.0),
           "regularization_type": solver.get("regularizationType", "identity")}
    if sol["regularization_type"] not in ("identity", "noser", "laplacian"):
        sol["regularization_type"] = "identity"   # solver.cpp:126-128: anything else is identity
    if gamma_ref is not None and gamma_ref.size != elements.shape[0]:
        raise ValueError("config: material distribution has %d entries for %d elements"
                         % (gamma_ref.size, elements.shape[0]))
    return spec, sol


def eit_from_config(config, path: str = ".", ctx=None, frames: int = 1, **eit_kw):
    """EIT::fromConfig: the forward model of a configuration."""
    from .eit import EIT
    spec, _ = parse_config(config, path)
    return EIT(spec.nodes, spec.elements, spec.boundary, spec.ports, spec.drive, spec.meas, spec.current,
               spec.reference_value, spec.components, spec.height, spec.port_height, frames=frames, ctx=ctx,
               source=spec.source, **eit_kw)


def solver_from_config(config, path: str = ".", ctx=None, options=None, **eit_kw):
    """Solver::fromConfig: the forward model for ``parallelImages`` frames, the
    noSensitivity mask, the regularisation, and preSolve at the reference distribution
    (the configured one, else gamma = 1).  Returns the Solver; ``solver.eit`` is the
    model."""
    import torch

    from .eit import EIT, Solver
    spec, sol = parse_config(config, path)
    eit = EIT(spec.nodes, spec.elements, spec.boundary, spec.ports, spec.drive, spec.meas, spec.current,
              spec.reference_value, spec.components, spec.height, spec.port_height,
              frames=sol["parallel_images"], ctx=ctx, source=spec.source, **eit_kw)
    s = Solver(eit, sol["regularization_factor"], options=options)
    if spec.no_sensitivity is not None:
        eit.set_sensitivity_mask(spec.no_sensitivity)
    g = None
    if spec.gamma_ref is not None:
        g = torch.tensor(spec.gamma_ref, dtype=torch.float32, device=eit._dev())
    s.set_regularization(sol["regularization_type"], g)
    s.pre_solve(g)
    if g is not None:
        s.material_distribution.copy_(g.expand_as(s.material_distribution))
    return s
