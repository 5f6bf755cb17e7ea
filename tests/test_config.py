"""Configuration files and text matrices (Solver::fromConfig, src/solver/solver.cpp:70-135;
Matrix::loadtxt / savetxt, src/numeric/matrix.cpp:684-822): host-side parsing only."""
import json
import os

import numpy as np
import pytest

from mpflow_amd import config as cfg
from mpflow_amd.mesh import boundary_edges, ring_mesh, unique_edges


def test_savetxt_format_and_round_trip(tmp_path):
    f = str(tmp_path / "m.txt")
    cfg.savetxt(f, np.array([[1.5, 2.0], [3.0, -0.25]]))
    assert open(f).read() == "1.5 2\n3 -0.25"          # no newline after the last row
    rng = np.random.default_rng(1)
    A = rng.standard_normal((7, 3)) * 10.0 ** rng.integers(-8, 8, (7, 3))
    cfg.savetxt(f, A)
    B = cfg.loadtxt(f)
    np.testing.assert_allclose(B, A, rtol=2e-16 * 10, atol=0)   # 16 significant digits
    cfg.savetxt(f, np.array([[1, -1], [2, 3]], np.int32))
    assert open(f).read() == "1 -1\n2 3"


def test_loadtxt_tolerates_stream_quirks(tmp_path):
    f = str(tmp_path / "m.txt")
    open(f, "w").write("1  2 3\n\n4 5 6\n")              # double blank, empty line, final newline
    np.testing.assert_array_equal(cfg.loadtxt(f), [[1, 2, 3], [4, 5, 6]])
    open(f, "w").write("2.5 -1\n")
    np.testing.assert_array_equal(cfg.loadtxt(f, np.int64), [[2, -1]])   # int extraction stops at '.'
    open(f, "w").write("\n \n")
    with pytest.raises(ValueError):
        cfg.loadtxt(f)


def _mesh_dir(tmp_path, rings=8, L=16, width=0.2):
    nodes, el = ring_mesh(rings, L, width)
    d = tmp_path / "mesh"
    d.mkdir()
    cfg.savetxt(str(d / "nodes.txt"), nodes)
    cfg.savetxt(str(d / "elements.txt"), el)
    return nodes, el


def _config(**model):
    base = {"model": {"mesh": "mesh", "ports": {"count": 16, "width": 0.2, "offset": 0, "clockwise": False},
                      "source": {"type": "current", "value": 1e-3},
                      "2.5D": {"components": 2, "height": 0.4}, "material": 1},
            "solver": {"parallelImages": 2, "regularizationFactor": 0.01, "regularizationType": "noser"}}
    base["model"].update(model)
    return base


def test_parse_config_generated_ports_and_edge_file(tmp_path):
    from mpflow_amd.eit import circular_ports
    nodes, el = _mesh_dir(tmp_path)
    spec, sol = cfg.parse_config(_config(), str(tmp_path))
    bd = boundary_edges(el)
    np.testing.assert_array_equal(spec.ports, circular_ports(nodes, el, bd, 16, 0.2))
    np.testing.assert_array_equal(spec.drive, np.eye(16))            # default patterns
    assert spec.components == 2 and spec.height == 0.4 and spec.reference_value == 1.0
    assert sol == {"parallel_images": 2, "regularization_factor": 0.01, "regularization_type": "noser"}
    # the generated ports were written as mesh-edge indices (ports.cpp:104-111) and the
    # edge list as edges.txt; reading them back gives the same ports
    assert os.path.exists(tmp_path / "mesh" / "ports.txt") and os.path.exists(tmp_path / "mesh" / "edges.txt")
    spec2, _ = cfg.parse_config(_config(ports="mesh/ports.txt"), str(tmp_path))
    np.testing.assert_array_equal(spec2.ports, spec.ports)
    edges, _ = unique_edges(el)
    rows = cfg.loadtxt(str(tmp_path / "mesh" / "ports.txt"), np.int64)
    for p in range(rows.shape[0]):
        for q in range(rows.shape[1]):
            if rows[p, q] >= 0:
                assert sorted(edges[rows[p, q]]) == sorted(bd[spec.ports[p, q]][[0, -1]])


def test_parse_config_patterns_values_material_file(tmp_path):
    nodes, el = _mesh_dir(tmp_path)
    L = 16
    drive = np.zeros((L, L))
    for d in range(L):
        drive[d, d], drive[(d + 1) % L, d] = 1.0, -1.0
    g = np.linspace(0.5, 2.0, el.shape[0])
    cfg.savetxt(str(tmp_path / "gamma.txt"), g[:, None])
    c = _config(source={"type": "current", "value": [float(v) for v in np.arange(1, L + 1)],
                        "drivePattern": drive.tolist(), "measurementPattern": drive.tolist()},
                material={"referenceValue": 2, "distribution": "gamma.txt"})
    path = str(tmp_path / "c.json")
    json.dump(c, open(path, "w"))
    spec, _ = cfg.parse_config(path)                    # file name: paths relative to it
    np.testing.assert_array_equal(spec.drive, drive)
    np.testing.assert_array_equal(spec.current, np.arange(1, L + 1))
    assert spec.reference_value == 2.0
    np.testing.assert_allclose(spec.gamma_ref, g, rtol=1e-15)


def test_parse_config_generated_mesh_and_rejections(tmp_path):
    c = _config(mesh={"radius": 1.0, "outerEdgeLength": 0.1, "innerEdgeLength": 0.1})
    spec, _ = cfg.parse_config(c, str(tmp_path))
    # the mesher keeps the 16 electrodes' corners (arc lengths 0 and 0.2 past 2 pi l / 16)
    # as its first nodes, and every port tiles its arc
    t = 2 * np.pi * np.arange(16) / 16
    corners = np.concatenate([np.stack([np.cos(t), np.sin(t)], 1), np.stack([np.cos(t + 0.2), np.sin(t + 0.2)], 1)])
    np.testing.assert_allclose(spec.nodes[:32], corners, atol=1e-15)
    assert 300 < spec.nodes.shape[0] < 700
    assert (spec.ports >= 0).sum(axis=1).min() >= 1
    assert os.path.exists(tmp_path / "mesh" / "elementEdges.txt")
    assert cfg.parse_config(_config(source={"type": "voltage", "value": 1.0}), str(tmp_path))[0].source == "voltage"
    assert cfg.parse_config(_config(source={"type": "open", "value": 1.0}), str(tmp_path))[0].source == "current"
    lap = _config()
    lap["solver"]["regularizationType"] = "laplacian"
    assert cfg.parse_config(lap, str(tmp_path))[1]["regularization_type"] == "laplacian"
    other = _config()
    other["solver"]["regularizationType"] = "tikhonov"
    assert cfg.parse_config(other, str(tmp_path))[1]["regularization_type"] == "identity"


def test_parse_config_no_sensitivity(tmp_path):
    """Solver::fromConfig reads noSensitivity only when it is a file name (solver.cpp:108)."""
    _mesh_dir(tmp_path)
    mask = np.arange(12.0).reshape(3, 4) / 7.0
    cfg.savetxt(str(tmp_path / "mask.txt"), mask)
    spec, _ = cfg.parse_config(_config(noSensitivity="mask.txt"), str(tmp_path))
    np.testing.assert_allclose(spec.no_sensitivity, mask, rtol=1e-15)
    spec, _ = cfg.parse_config(_config(noSensitivity=1), str(tmp_path))
    assert spec.no_sensitivity is None
    spec, _ = cfg.parse_config(_config(), str(tmp_path))
    assert spec.no_sensitivity is None
