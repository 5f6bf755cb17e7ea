"""CPU tests that pin the FP64 oracle (oracle/eit_oracle.py) to closed forms,
invariants, brute force, finite differences and the analytic disc solution, so that a
plausible restatement mistake (dropped term, wrong sign, transposed operand) fails one
of them.  The reference ships no golden vectors or tests; these checks are the pin."""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from mpflow_amd.mesh import boundary_edges, promote_p2, ring_mesh
from mpflow_amd.phantom import SplitMix64, ellipse_phantom, frame_seed, inclusion_phantom
from oracle import eit_oracle as eo
from tests.problems import make_problem, oracle_of, rcb_renumbered


# ----------------------------------------------------------------------------------
# element integrals
# ----------------------------------------------------------------------------------
def _rand_tri(rng):
    while True:
        p = rng.uniform(-1, 1, (3, 2))
        a = 0.5 * ((p[1, 0] - p[0, 0]) * (p[2, 1] - p[0, 1]) - (p[1, 1] - p[0, 1]) * (p[2, 0] - p[0, 0]))
        if a > 0.05:
            return p, a


def test_p1_integrals_closed_form():
    rng = np.random.default_rng(0)
    for _ in range(20):
        p, area = _rand_tri(rng)
        K, M = eo.element_integrals(p)
        # grad(lambda_i) = (y_j - y_k, x_k - x_j) / (2 area)
        g = np.array([[p[(i + 1) % 3, 1] - p[(i + 2) % 3, 1], p[(i + 2) % 3, 0] - p[(i + 1) % 3, 0]]
                      for i in range(3)]) / (2 * area)
        np.testing.assert_allclose(K, area * g @ g.T, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(M, area / 12.0 * (np.ones((3, 3)) + np.eye(3)), rtol=1e-12)


def _p2_nodes(p):
    return np.vstack([p, 0.5 * (p[0] + p[1]), 0.5 * (p[1] + p[2]), 0.5 * (p[2] + p[0])])


def test_p2_integrals_textbook():
    """P2 mass matrix A/180 * [[6,-1,-1,0,-4,0], ...] (vertex-midpoint ordering
    v0 v1 v2 m01 m12 m20); stiffness rows sum to zero and reproduce quadratics."""
    Mref = np.array([[6, -1, -1, 0, -4, 0], [-1, 6, -1, 0, 0, -4], [-1, -1, 6, -4, 0, 0],
                     [0, 0, -4, 32, 16, 16], [-4, 0, 0, 16, 32, 16], [0, -4, 0, 16, 16, 32]], float) / 180.0
    rng = np.random.default_rng(1)
    for _ in range(10):
        p, area = _rand_tri(rng)
        xy = _p2_nodes(p)
        K, M = eo.element_integrals(xy)
        np.testing.assert_allclose(M, area * Mref, rtol=1e-10, atol=1e-13)
        np.testing.assert_allclose(K.sum(axis=1), 0.0, atol=1e-11)
        np.testing.assert_allclose(K, K.T, atol=1e-12)
        # energy of u = x^2 + xy: int |grad u|^2 = int (2x+y)^2 + x^2 (exact on P2)
        u = xy[:, 0] ** 2 + xy[:, 0] * xy[:, 1]
        lam = eo._Q4[:, :3]
        X = lam @ p[:, 0]
        Y = lam @ p[:, 1]
        exact = area * np.sum(eo._Q4[:, 3] * ((2 * X + Y) ** 2 + X ** 2))
        assert abs(u @ K @ u - exact) < 1e-10 * max(1.0, abs(exact))


def test_product_element_integrals_match_oracle():
    """The product's host set-up (barycentric closed forms) equals the oracle's
    Vandermonde + quadrature restatement, P1 and P2."""
    import ctypes as C
    from mpflow_amd.capi import check, lib
    rng = np.random.default_rng(2)
    for k in (3, 6):
        for _ in range(10):
            p, _ = _rand_tri(rng)
            xy = np.ascontiguousarray(p if k == 3 else _p2_nodes(p))
            K = np.zeros((k, k)); M = np.zeros((k, k))
            check(lib().mpf_element_integrals(k, xy.ctypes.data, K.ctypes.data, M.ctypes.data))
            Ko, Mo = eo.element_integrals(xy)
            np.testing.assert_allclose(K, Ko, rtol=1e-10, atol=1e-12)
            np.testing.assert_allclose(M, Mo, rtol=1e-10, atol=1e-14)


def test_boundary_integrals():
    L = 0.37
    s1 = np.array([0.0, L])
    assert eo.boundary_integral(s1, 0) == pytest.approx(L / 2)
    assert eo.boundary_integral(s1, 1) == pytest.approx(L / 2)
    s2 = np.array([0.0, L / 2, L])
    assert eo.boundary_integral(s2, 0, quirks=False) == pytest.approx(L / 6)
    assert eo.boundary_integral(s2, 1, quirks=False) == pytest.approx(2 * L / 3)
    assert eo.boundary_integral(s2, 2, quirks=False) == pytest.approx(L / 6)
    # reference code evaluates the last P2 edge node with node 1's coefficients
    assert eo.boundary_integral(s2, 2, quirks=True) == pytest.approx(2 * L / 3)


# ----------------------------------------------------------------------------------
# mesh, ports, phantoms
# ----------------------------------------------------------------------------------
@pytest.mark.parametrize("nr", [2, 5, 16, 40])
def test_ring_mesh_counts(nr):
    n, e = ring_mesh(nr)
    assert n.shape[0] == 1 + 3 * nr * (nr + 1)
    assert e.shape[0] == 6 * nr * nr
    b = boundary_edges(e)
    assert b.shape[0] == 6 * nr
    a, b = n[e[:, 1]] - n[e[:, 0]], n[e[:, 2]] - n[e[:, 0]]
    area = 0.5 * np.abs(a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]).sum()
    # inscribed polygon of 6 nr boundary nodes
    assert area == pytest.approx(0.5 * 6 * nr * math.sin(2 * math.pi / (6 * nr)), rel=1e-12)


def test_p2_promotion():
    n, e = ring_mesh(4)
    n2, e2 = promote_p2(n, e)
    V, E = n.shape[0], e.shape[0]
    edges = V + E - 1
    assert n2.shape[0] == V + edges
    for t, (a, b) in enumerate([(0, 1), (1, 2), (2, 0)]):
        np.testing.assert_allclose(n2[e2[:, 3 + t]], 0.5 * (n2[e2[:, a]] + n2[e2[:, b]]))
    bd = boundary_edges(e2)
    assert bd.shape == (6 * 4, 3)
    np.testing.assert_allclose(n2[bd[:, 1]], 0.5 * (n2[bd[:, 0]] + n2[bd[:, 2]]))


def test_electrode_corners_on_nodes():
    L, w = 16, 0.1
    n, e = ring_mesh(16, L, w)
    ang = np.mod(np.arctan2(n[:, 1], n[:, 0]), 2 * math.pi)
    r = np.hypot(n[:, 0], n[:, 1])
    bang = ang[np.abs(r - 1) < 1e-12]
    for l in range(L):
        for c in (2 * math.pi * l / L, 2 * math.pi * l / L + w):
            assert np.min(np.abs(np.angle(np.exp(1j * (bang - c))))) < 1e-12


def test_ports_product_matches_oracle():
    from mpflow_amd.eit import circular_ports
    for p2 in (False, True):
        n, e = ring_mesh(16, 16, 0.1)
        if p2:
            n, e = promote_p2(n, e)
        bd = boundary_edges(e)
        a = eo.circular_boundary_ports(n, e, bd, 16, 0.1)
        b = circular_ports(n, e, bd, 16, 0.1)
        np.testing.assert_array_equal(a, b)
        # every port is covered by edges whose endpoints lie inside its arc
        assert np.all((a >= 0).sum(axis=1) >= 1)


@pytest.mark.parametrize("rings,L,w,p2", [(16, 16, 0.1, False), (16, 16, 0.1, True), (40, 32, 0.05, False),
                                           (12, 8, 0.3, True)])
def test_circular_boundary_tiles_each_arc(rings, L, w, p2):
    """Ports::circularBoundary (ports.cpp:32-77) pinned by geometry: on a ring mesh whose
    boundary nodes include every electrode corner, the edges selected for port l chain
    end to end from the corner at angle 2 pi l / L to the corner at 2 pi l / L + w (they
    tile the arc exactly), and no other boundary edge lies inside the arc."""
    n, e = ring_mesh(rings, L, w)
    if p2:
        n, e = promote_p2(n, e)
    bd = boundary_edges(e)
    ports = eo.circular_boundary_ports(n, e, bd, L, w)
    ang = lambda i: math.atan2(n[i, 1], n[i, 0])
    for l in range(L):
        t0 = 2 * math.pi * l / L
        # angles measured counter-clockwise from the port's first corner, in [0, 2 pi)
        rel = lambda i: (ang(i) - t0) % (2 * math.pi)
        sel = [int(k) for k in ports[l] if k >= 0]
        inside = [k for k in range(bd.shape[0])
                  if rel(bd[k, 0]) <= w + 1e-12 and 0 < rel(bd[k, -1]) <= w + 1e-12
                  and rel(bd[k, 0]) < rel(bd[k, -1])]
        assert sorted(sel) == sorted(inside)
        chain = sorted(sel, key=lambda k: rel(bd[k, 0]))
        assert abs(rel(bd[chain[0], 0])) < 1e-12 or abs(rel(bd[chain[0], 0]) - 2 * math.pi) < 1e-12
        assert abs(rel(bd[chain[-1], -1]) - w) < 1e-12
        for a, b in zip(chain, chain[1:]):
            assert bd[a, -1] == bd[b, 0]


def test_splitmix64_known_answer():
    # reference outputs of SplitMix64 seeded with 1234567
    g = SplitMix64(1234567)
    assert [g.next_u64() for _ in range(5)] == [
        6457827717110365317, 3203168211198807973, 9817491932198370423,
        4593380528125082431, 16408922859458223821]


def test_phantoms_deterministic():
    n, e = ring_mesh(16, 16, 0.1)
    a = ellipse_phantom(n, e, frame_seed(2, 7))
    b = ellipse_phantom(n, e, frame_seed(2, 7))
    np.testing.assert_array_equal(a, b)
    assert a.min() >= 0.2 - 1e-12 and a.max() <= 5.0 + 1e-12
    g = inclusion_phantom(n, e, frame_seed(1, 0))
    assert set(np.unique(g)) <= {1.0, 2.0} and np.any(g == 2.0)


# ----------------------------------------------------------------------------------
# assembly, excitation, solver
# ----------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def small():
    p = make_problem(8, 8, 0.2)
    o = oracle_of(p)
    g = inclusion_phantom(p["nodes"], p["elements"], frame_seed(1, 2))
    return p, o, g


def test_assembly_invariants(small):
    p, o, g = small
    S = eo.assemble(o.elements, o.Ke, o.N, g)
    R = eo.assemble(o.elements, o.Me, o.N, g)
    assert abs(S - S.T).max() < 1e-13
    np.testing.assert_allclose(S @ np.ones(o.N), 0.0, atol=1e-12)
    tri = p["nodes"][p["elements"]]
    a, b = tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]
    area = 0.5 * np.abs(a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0])
    assert np.ones(o.N) @ R @ np.ones(o.N) == pytest.approx(np.sum(g * area), rel=1e-12)


def test_excitation_columns(small):
    p, o, g = small
    np.testing.assert_allclose(o.B.sum(axis=0), 1.0, rtol=1e-12)
    # support of each column = the nodes of the port's edges
    for q in range(o.B.shape[1]):
        nodes = set(p["boundary"][[e for e in p["ports"][q] if e >= 0]].ravel())
        assert set(np.nonzero(o.B[:, q])[0]) == nodes


def test_excitation_p2_quirk_factor():
    p = make_problem(4, 8, 0.3, p2=True)
    oq = oracle_of(p, quirks=True)
    oc = oracle_of(p, quirks=False)
    np.testing.assert_allclose(oc.B.sum(axis=0), 1.0, rtol=1e-12)
    np.testing.assert_allclose(oq.B.sum(axis=0), 1.5, rtol=1e-12)


def test_pcg_restatement_against_direct(small):
    p, o, g = small
    A = o.system(g, 0)
    F = o.alpha_beta(0)[1] * (o.B @ o.Pi)
    X = eo.solve_direct(A, F, True)
    Xc, steps = eo.pcg_reference(A, F, 1.0 / A.diagonal(), True, rule="tolerance", tol=1e-13)
    assert np.max(np.abs(Xc - X)) < 1e-10 * np.max(np.abs(X))
    # the dcFree solution has zero nodal sum for zero-sum excitations
    np.testing.assert_allclose(X.sum(axis=0), 0.0, atol=1e-12)
    # brute force: dense (A + 1 1^T) solve
    Ad = A.toarray() + np.ones((o.N, o.N))
    np.testing.assert_allclose(np.linalg.solve(Ad, F), X, atol=1e-12 * np.max(np.abs(X)))
    # right-hand sides with a nonzero sum (single-electrode drives): the dense brute force
    # of (A + 1 1^T) fixes the constant part as 1^T x = 1^T F / N
    Fe = o.alpha_beta(0)[1] * o.B[:, :3]
    Xe = eo.solve_direct(A, Fe, True)
    np.testing.assert_allclose(Xe, np.linalg.solve(Ad, Fe), atol=1e-11 * np.max(np.abs(Xe)))
    np.testing.assert_allclose(Xe.sum(axis=0), Fe.sum(axis=0) / o.N, rtol=1e-10)
    Xce, _ = eo.pcg_reference(A, Fe, 1.0 / A.diagonal(), True, rule="tolerance", tol=1e-13)
    np.testing.assert_allclose(Xce, Xe, atol=1e-10 * np.max(np.abs(Xe)))
    # reference stop rule: column 0 reaches sqrt(r.z / b.b) < FLT_EPSILON
    Xr, sr = eo.pcg_reference(A, F, 1.0 / A.diagonal(), True)
    assert sr < 10 * o.N
    assert np.max(np.abs(Xr - X)) < 1e-5 * np.max(np.abs(X))


def test_bicgstab_restatement_against_direct(small):
    """BiCGSTAB (no dcFree) on the 2.5D harmonic-1 matrix (SPD, alpha_1 > 0) reaches the
    direct solve; on the singular harmonic-0 matrix with zero-sum excitations it reaches
    a solution that differs from the dcFree one by a constant per column, so electrode
    difference voltages agree."""
    p, o, g = small
    F = o.B @ o.Pi
    for c in (0, 1):
        A = o.system(g, c)   # alpha_c = (2 pi c / h)^2
        X = eo.solve_direct(A, F, c == 0)
        Xb, steps = eo.bicgstab_reference(A, F, 1.0 / A.diagonal(), rule="tolerance", tol=1e-13)
        assert steps < 10 * o.N
        res = np.max(np.abs(A @ Xb - F)) / np.max(np.abs(F))
        assert res < 1e-11
        D = Xb - X
        if c == 0:
            D = D - D.mean(axis=0, keepdims=True)
        assert np.max(np.abs(D)) < 1e-9 * np.max(np.abs(X))
    # the projected variant reaches the same solution up to a constant
    A = o.system(g, 0)
    X = eo.solve_direct(A, F, True)
    Xp, _ = eo.bicgstab_reference(A, F, 1.0 / A.diagonal(), rule="tolerance", tol=1e-13, project=True)
    D = Xp - X
    assert np.max(np.abs(D - D.mean(axis=0, keepdims=True))) < 1e-9 * np.max(np.abs(X))
    # reference stop rule (column 0, sqrt(r.r / b.b) < FLT_EPSILON)
    Xr, sr = eo.bicgstab_reference(A, F, 1.0 / A.diagonal())
    assert sr < 10 * o.N
    r0 = F[:, 0] - A @ Xr[:, 0]
    assert np.sqrt(r0 @ r0 / (F[:, 0] @ F[:, 0])) < eo.FLT_EPSILON


def test_bicgstab_tiny_spd_exact_in_few_steps():
    """On a 6 x 6 SPD matrix, BiCGSTAB reaches the exact solution within 2N steps (a
    Krylov method on an N-dimensional space), for several right-hand sides at once."""
    rng = np.random.default_rng(7)
    M = rng.standard_normal((6, 6))
    A = M @ M.T + 6 * np.eye(6)
    b = rng.standard_normal((6, 3))
    x, steps = eo.bicgstab_reference(A, b, 1.0 / np.diag(A), rule="tolerance", tol=1e-14)
    assert steps <= 12
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)


def test_pcg_start_vector(small):
    """ConjugateGradient::solve starts from the caller's x (conjugate_gradient.cpp:80-88).
    SPD harmonic: any start reaches the direct solution.  dcFree harmonic: the reference
    forms the start residual b - A x0 + 1 (1^T x0) while the operator is A + 1 1^T, so the
    solution moves by exactly 2 (1^T x0) / N per column (A 1 = 0); a start of 0 is the
    default."""
    p, o, g = small
    rng = np.random.default_rng(3)
    F = o.alpha_beta(0)[1] * (o.B @ o.Pi)
    x0 = rng.standard_normal(F.shape)
    A1 = o.system(g, 1)
    X1 = eo.solve_direct(A1, F, False)
    Xs, _ = eo.pcg_reference(A1, F, 1.0 / A1.diagonal(), False, rule="tolerance", tol=1e-13, x0=x0)
    np.testing.assert_allclose(Xs, X1, atol=1e-10 * np.max(np.abs(X1)))
    A0 = o.system(g, 0)
    X0 = eo.solve_direct(A0, F, True)
    Xd, _ = eo.pcg_reference(A0, F, 1.0 / A0.diagonal(), True, rule="tolerance", tol=1e-13, x0=x0)
    shift = 2.0 * x0.sum(axis=0, keepdims=True) / o.N
    np.testing.assert_allclose(Xd, X0 + shift, atol=1e-9 * np.max(np.abs(X0)))
    Xz, sz = eo.pcg_reference(A0, F, 1.0 / A0.diagonal(), True, x0=np.zeros_like(F))
    Xn, sn = eo.pcg_reference(A0, F, 1.0 / A0.diagonal(), True)
    assert sz == sn and np.array_equal(Xz, Xn)


def test_pcg_fp32_arithmetic_mode(small):
    """The float32 mode of the restatement (the reference's arithmetic) reaches the FP64
    solution to float32 accuracy, takes the FP64 step count where rounding does not bind
    (the mass-dominated harmonic 1) and more steps than FP64 on the dcFree harmonic under
    the FLT_EPSILON rule (float32 rounding slows the last digits)."""
    p, o, g = small
    F = o.alpha_beta(0)[1] * (o.B @ o.Pi)
    A1 = o.system(g, 1)
    X1 = eo.solve_direct(A1, F, False)
    d1 = 1.0 / A1.diagonal()
    x32, s32 = eo.pcg_reference(A1, F, d1, False, rule="tolerance", tol=1e-6, fp32=True)
    assert np.max(np.abs(x32 - X1)) < 1e-5 * np.max(np.abs(X1))
    assert x32.astype(np.float32).astype(np.float64).tolist() == x32.tolist()
    assert eo.pcg_reference(A1, F, d1, False, fp32=True)[1] == eo.pcg_reference(A1, F, d1, False)[1]
    A0 = o.system(g, 0)
    d0 = 1.0 / A0.diagonal()
    assert eo.pcg_reference(A0, F, d0, True, fp32=True)[1] > eo.pcg_reference(A0, F, d0, True)[1]


def test_pcg_zero_rhs_stops_after_one_step(small):
    p, o, g = small
    A = o.system(g, 0)
    b = np.zeros((o.N, 2))
    x, steps = eo.pcg_reference(A, b, 1.0 / A.diagonal(), True)
    assert steps == 1 and np.all(x == 0.0)


def _grid_laplacian(m, shift=0.0):
    """5-point Laplacian on an m x m grid with Neumann edges (A 1 = 0 when shift = 0)."""
    n = m * m
    rows, cols, vals = [], [], []
    for i in range(m):
        for j in range(m):
            k = i * m + j
            for di, dj in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                a, b = i + di, j + dj
                if 0 <= a < m and 0 <= b < m:
                    rows += [k, k]
                    cols += [k, a * m + b]
                    vals += [1.0, -1.0]
    return (sp.csr_matrix((vals, (rows, cols)), shape=(n, n)) + shift * sp.eye(n)).tocsr()


def test_block_pcg_exact_when_the_block_spans_the_space():
    """With as many independent columns as unknowns, the first block of directions spans
    the whole space: one block step solves every column exactly."""
    rng = np.random.default_rng(5)
    M = rng.standard_normal((24, 24))
    A = sp.csr_matrix(M @ M.T + 24 * np.eye(24))
    B = rng.standard_normal((24, 32))
    X, steps = eo.block_pcg(A, B, tol=1e-12)
    assert steps == 1
    np.testing.assert_allclose(X, np.linalg.solve(A.toarray(), B), rtol=0, atol=1e-10)


def test_block_pcg_no_more_steps_than_single_columns():
    """Each column's iterate minimises the A-norm error over the sum of every column's
    Krylov space, which contains its own: the block method needs no more steps than
    single-column PCG for the same stop rule (FP64)."""
    A = _grid_laplacian(16, 1e-3)
    B = np.random.default_rng(2).standard_normal((A.shape[0], 8))
    Xs, ss = eo.pcg_reference(A, B, 1.0 / A.diagonal(), False, rule="tolerance", tol=1e-8)
    Xb, sb = eo.block_pcg(A, B, tol=1e-8)
    assert sb <= ss and sb < 0.8 * ss
    Xd = spla.spsolve(A.tocsc(), B)
    assert np.max(np.abs(Xb - Xd)) / np.max(np.abs(Xd)) < 1e-6


def test_block_pcg_rank_deficient_and_dcfree(small):
    """Drive patterns that sum to zero give a rank-deficient block (the dependent
    directions are dropped); the dcFree operator A + 1 1^T of the singular harmonic;
    a zero right-hand side stops before the first step."""
    p, o, g = small
    B = o.B @ o.Pi
    assert np.linalg.matrix_rank(B) < B.shape[1]
    T, k = eo._pivoted_cholesky_orth(B, 1e-10)
    assert k == np.linalg.matrix_rank(B)
    Q = B @ T
    np.testing.assert_allclose(Q[:, :k].T @ Q[:, :k], np.eye(k), atol=1e-8)
    for c, dc in ((1, False), (0, True)):
        A = o.system(g, c)
        Xb, sb = eo.block_pcg(A, B, None, dc, tol=1e-10)
        Xd = eo.solve_direct(A, B, dc)
        assert np.max(np.abs(Xb - Xd)) / np.max(np.abs(Xd)) < 1e-8
        _, ss = eo.pcg_reference(A, B, 1.0 / A.diagonal(), dc, rule="tolerance", tol=1e-10)
        assert sb <= ss
    X0, s0 = eo.block_pcg(o.system(g, 1), np.zeros_like(B))
    assert s0 == 0 and np.all(X0 == 0.0)


def _disc_voltages(L, width, current, nmax=4000):
    """Gap-model voltages of the unit disc (sigma = 1): Fourier series of the Neumann
    problem with current density +-I/w on the driving electrodes, potentials averaged
    over each measuring electrode; adjacent drive and measurement."""
    n = np.arange(1, nmax + 1)[:, None]
    t1 = 2 * np.pi * np.arange(L) / L
    t2 = t1 + width
    # Fourier coefficients of the indicator of each electrode, per unit density
    ca = (np.sin(n * t2) - np.sin(n * t1)) / (n * np.pi)        # [n, L]
    cb = (np.cos(n * t1) - np.cos(n * t2)) / (n * np.pi)
    # electrode averages of cos / sin
    av_c = (np.sin(n * t2) - np.sin(n * t1)) / (n * width)
    av_s = (np.cos(n * t1) - np.cos(n * t2)) / (n * width)
    # potential of unit density on electrode e, averaged over electrode m
    G = (ca.T @ (av_c / n) + cb.T @ (av_s / n))                # [e, m]
    V = np.zeros((L, L))
    for d in range(L):
        j = np.zeros(L)
        j[d], j[(d + 1) % L] = current / width, -current / width
        V[:, d] = j @ G
    P = eo.adjacent_pattern(L)
    return P.T @ V


def test_forward_converges_to_analytic_disc():
    errs = []
    for nr in (8, 16, 32):
        p = make_problem(nr, 8, 0.3)
        o = oracle_of(p)
        U, _, _ = o.solve(np.ones(p["elements"].shape[0]))
        Ua = _disc_voltages(8, 0.3, 1e-3)
        errs.append(np.max(np.abs(U - Ua)) / np.max(np.abs(Ua)))
    assert errs[2] < 0.6 * errs[1] < 0.6 * errs[0]
    assert errs[2] < 2e-2


def test_forward_reciprocity_and_2p5d_reduction(small):
    p, o, g = small
    U, _, _ = o.solve(g)
    np.testing.assert_allclose(U, U.T, atol=1e-12 * np.max(np.abs(U)))
    # C = 1 with h = 1 is the 2D model
    p2 = dict(p, C=1, height=1.0, port_height=0.3)
    U2, _, _ = oracle_of(p2).solve(g)
    np.testing.assert_allclose(U2, U, rtol=1e-12, atol=0)


def test_harmonic_constants():
    p = make_problem(4, 8, 0.3, C=3, height=0.4, port_height=0.1)
    o = oracle_of(p)
    a1, b1 = o.alpha_beta(1)
    assert a1 == pytest.approx((2 * math.pi / 0.4) ** 2)
    assert b1 == pytest.approx(2 * math.sin(math.pi * 0.1 / 0.4) / (math.pi * 0.1))
    assert o.alpha_beta(0) == (0.0, 1 / 0.4)


def _fd_jacobian(o, g, elems, eps=1e-5):
    cols = []
    for e in elems:
        gp, gm = g.copy(), g.copy()
        gp[e] += eps
        gm[e] -= eps
        Up, _, _ = o.solve(gp)
        Um, _, _ = o.solve(gm)
        cols.append(o.flatten_voltages((Up - Um) / (2 * eps)))
    return np.stack(cols, axis=1)


def test_jacobian_matches_finite_differences_2d(small):
    p, o, g = small
    U, phis, _ = o.solve(g)
    J = o.jacobian(g, phis)
    elems = [0, 7, 50, 200, o.E - 1]
    Jfd = _fd_jacobian(o, g, elems)
    np.testing.assert_allclose(J[:, elems], Jfd, atol=1e-7 * np.max(np.abs(Jfd)))


def test_jacobian_2p5d_exact_vs_reference_formula():
    p = make_problem(6, 8, 0.3, C=3, height=0.4, port_height=0.1)
    o = oracle_of(p)
    g = inclusion_phantom(p["nodes"], p["elements"], frame_seed(1, 4))
    U, phis, _ = o.solve(g)
    elems = [3, 40, 111]
    Jfd = _fd_jacobian(o, g, elems)
    Jx = o.jacobian(g, phis, quirks=False)
    Jr = o.jacobian(g, phis, quirks=True)
    np.testing.assert_allclose(Jx[:, elems], Jfd, atol=1e-6 * np.max(np.abs(Jfd)))
    # the reference formula (K_e only, factor height) is not the derivative for C >= 2
    assert np.max(np.abs(Jr[:, elems] - Jfd)) > 1e-2 * np.max(np.abs(Jfd))


def test_jacobian_reference_formula_is_frozen_mass_derivative():
    """The reference Jacobian for C >= 2 (calcJacobianKernel with K_e only, summed over
    harmonics with the factor height; equation_kernel.cu:197-245, eit.cpp:180-182,
    202-205) pinned by finite differences: per harmonic c, U_c solved with
    A_c = S(gamma) + alpha_c R(gamma_0) -- the mass matrix frozen at the expansion point so
    only K_e carries sigma -- differentiated centrally and weighted by height * beta_c
    (the reciprocity factor 1/beta_c of the derivative replaced by height), summed over c.
    Height 2 and port height 0.5 keep alpha_c small enough that harmonics 1 and 2 carry a
    large share of every column (weights 1, h or 1/beta_c miss by 5e-2 or more)."""
    p = make_problem(6, 8, 0.3, C=3, height=2.0, port_height=0.5)
    o = oracle_of(p)
    g = inclusion_phantom(p["nodes"], p["elements"], frame_seed(1, 4))
    U, phis, _ = o.solve(g)
    Jr = o.jacobian(g, phis, quirks=True)
    sig0 = eo.mat_param(g, o.ref, o.log)
    R0 = eo.assemble(o.elements, o.Me, o.N, sig0)

    def u_harmonic(gam, c):
        S = eo.assemble(o.elements, o.Ke, o.N, eo.mat_param(gam, o.ref, o.log))
        alpha, beta = o.alpha_beta(c)
        X = eo.solve_direct((S + alpha * R0).tocsr(), beta * (o.B @ o.Pi), c == 0)
        return o.flatten_voltages(o.W @ X[:, :o.D])

    elems, eps = [3, 40, 158], 1e-5
    for e in elems:
        col = np.zeros(Jr.shape[0])
        for c in range(o.C):
            gp, gm = g.copy(), g.copy()
            gp[e] += eps
            gm[e] -= eps
            beta = o.alpha_beta(c)[1]
            col += o.h * beta * (u_harmonic(gp, c) - u_harmonic(gm, c)) / (2 * eps)
        np.testing.assert_allclose(Jr[:, e], col, atol=1e-6 * np.max(np.abs(col)))


def _coupled(C, h, hp, z=0.01, rings=6, L=8, w=0.3):
    p = make_problem(rings, L, w, C=C, height=h, port_height=hp)
    o = eo.CoupledCEMOracle(p["nodes"], p["elements"], p["boundary"], p["ports"], p["drive"], p["meas"],
                            p["current"], 1.0, C, h, hp, contact_impedance=z)
    g = inclusion_phantom(p["nodes"], p["elements"], frame_seed(1, 4))
    return p, o, g


def test_coupled_cem_one_harmonic_is_harmonic_wise():
    """One harmonic with the electrode over the full height: the coupled model is the
    harmonic-wise complete electrode model with current I / h (beta_0 = 1 / h)."""
    p, o, g = _coupled(1, 1.0, 1.0)
    U, _, _ = o.solve(g)
    oh = eo.EITOracle(p["nodes"], p["elements"], p["boundary"], p["ports"], p["drive"], p["meas"], p["current"],
                      1.0, 1, 1.0, 1.0, quirks=False, contact_impedance=0.01)
    Uh, _, _ = oh.solve(g)
    np.testing.assert_allclose(U, Uh, atol=1e-12 * np.max(np.abs(Uh)))


@pytest.mark.parametrize("C,h,hp", [(3, 0.4, 0.1), (5, 0.4, 0.1), (3, 2.0, 0.5)])
def test_coupled_cem_energy_reciprocity_definiteness(C, h, hp):
    """The block system is symmetric positive definite for C >= 2; U = U^T; and for
    every drive column the input power I^T V equals the energy
    sum_c n_c [u_c^T A_c u_c + (1/z') int_e (u_c - h beta_c V)^2], evaluated from the
    fields with the edge integrals of the contact term."""
    p, o, g = _coupled(C, h, hp)
    H = o.coupled_system(g).toarray()
    assert np.linalg.eigvalsh(H)[0] > 0.0
    U, phis, _ = o.solve(g)
    assert np.max(np.abs(U - U.T)) < 1e-12 * np.max(np.abs(U))
    sig = eo.mat_param(g, 1.0)
    for d in range(o.D):
        V = o.V[:, d]
        power = o.Pi[:, d] @ V
        energy = 0.0
        for c, X in enumerate(phis):
            alpha, beta = o.alpha_beta(c)
            A = eo.assemble(o.elements, o.Ke, o.N, sig) + alpha * eo.assemble(o.elements, o.Me, o.N, sig)
            u = X[:, d]
            hb = o.h * beta
            # int_e (u - hb V_l)^2 over the electrode edges: u^T Zb u - 2 hb u^T Bc V + hb^2 |e| V^2
            contact = u @ (o.Zb @ u) - 2.0 * hb * (u @ (o.Bc @ V)) + hb * hb * (o.elen @ (V * V))
            energy += o.n_c[c] * (u @ (A @ u) + contact / o.zp)
        assert abs(power - energy) < 1e-10 * abs(power)


def test_coupled_cem_jacobian_matches_finite_differences():
    p, o, g = _coupled(3, 0.4, 0.1)
    U, phis, _ = o.solve(g)
    J = o.jacobian(g, phis)
    elems = [3, 40, 158]
    Jfd = _fd_jacobian(o, g, elems)
    np.testing.assert_allclose(J[:, elems], Jfd, atol=1e-6 * np.max(np.abs(Jfd)))


def test_gn_data_space_identity():
    rng = np.random.default_rng(5)
    J = rng.standard_normal((40, 120)) * 1e-3
    r = rng.standard_normal(40) * 1e-4
    lam2 = 1e-6
    d1 = eo.gn_step_exact(J, r, lam2)
    d2 = eo.gn_step_normal(J, r, lam2)
    np.testing.assert_allclose(d1, d2, rtol=1e-9, atol=1e-15)
    d3, _ = eo.gn_step_cg_reference(J, r, lam2, eps=1e-30, max_iter=2000)
    np.testing.assert_allclose(d3, d1, rtol=1e-6, atol=1e-12)


def test_reconstruction_improves_fit():
    p = make_problem(8, 8, 0.2)
    o = oracle_of(p)
    gt = inclusion_phantom(p["nodes"], p["elements"], frame_seed(1, 11))
    Um, _, _ = o.solve(gt)
    U0, _, _ = o.solve(np.ones(o.E))
    res = []
    for steps in (1, 3):
        g = eo.reconstruct(o, Um, np.ones(o.E), 1e-4, steps=steps)
        U1, _, _ = o.solve(g)
        res.append(np.linalg.norm(U1 - Um) / np.linalg.norm(U0 - Um))
    assert res[1] < res[0] < 0.8 and res[1] < 0.6


def test_cpu_baseline_matches_direct(small):
    from oracle import baseline
    p, o, g = small
    A = o.system(g, 0)
    F = o.alpha_beta(0)[1] * (o.B @ o.Pi)
    X, steps, _ = baseline.pcg(A, F, 1, tol=1e-6)
    Xd = eo.solve_direct(A, F, True)
    assert np.max(np.abs(X - Xd)) < 1e-5 * np.max(np.abs(Xd))


# ----------------------------------------------------------------------------------
# complete electrode model (an extension without a reference counterpart)
# ----------------------------------------------------------------------------------
def _cem_oracle(p, z, **kw):
    return eo.EITOracle(p["nodes"], p["elements"], p["boundary"], p["ports"], p["drive"], p["meas"],
                        p["current"], 1.0, p["C"], p["height"], p["port_height"], contact_impedance=z, **kw)


@pytest.mark.parametrize("p2,C", [(False, 1), (True, 1), (False, 3)])
def test_cem_currents_and_power_balance(p2, C):
    """Per drive pattern and harmonic: the electrode currents (1/z) int_e (V_l - phi) ds,
    integrated by 5-point Gauss-Legendre along each electrode edge, equal the injected
    pattern currents; and the input power V^T I equals phi^T (S + alpha R) phi plus the
    contact losses (1/z) int_e (phi - V_l)^2 ds, integrated the same way."""
    p = make_problem(8, 8, 0.3, C=C, height=0.4 if C > 1 else 1.0, port_height=0.1 if C > 1 else 1.0, p2=p2)
    z = 0.01
    o = _cem_oracle(p, z, quirks=False)
    g = inclusion_phantom(p["nodes"], p["elements"], frame_seed(1, 6))
    sig = eo.mat_param(g, 1.0, False)
    S = eo.assemble(o.elements, o.Ke, o.N, sig)
    R = eo.assemble(o.elements, o.Me, o.N, sig)
    xg, wg = np.polynomial.legendre.leggauss(5)
    for c in range(C):
        alpha, beta = o.alpha_beta(c)
        X = eo.solve_direct(o.system(g, c), beta * (o.B @ o.Pi), c == 0)
        for d in range(o.D):
            phi, V = X[:o.N, d], X[o.N:, d]
            I = beta * o.Pi[:, d]
            cur = np.zeros(o.L)
            loss = 0.0
            for l in range(o.L):
                for e in p["ports"][l]:
                    if e < 0:
                        continue
                    ed = p["boundary"][e]
                    pts = p["nodes"][ed]
                    s = np.sqrt(np.sum((pts - pts[0]) ** 2, axis=1))
                    n = len(ed)
                    coef = np.linalg.solve(np.vander(s, n, increasing=True), np.eye(n))
                    t = 0.5 * s[-1] * (xg + 1.0)
                    u = np.vander(t, n, increasing=True) @ coef @ phi[ed]
                    cur[l] += np.sum(0.5 * s[-1] * wg * (V[l] - u)) / z
                    loss += np.sum(0.5 * s[-1] * wg * (u - V[l]) ** 2) / z
            np.testing.assert_allclose(cur, I, atol=1e-9 * np.max(np.abs(I)))
            power = V @ I
            assert abs(power - (phi @ (S @ phi) + alpha * phi @ (R @ phi) + loss)) < 1e-9 * abs(power)


def test_cem_reciprocity_and_jacobian_fd():
    """2D CEM: U = U^T for identical drive and measurement patterns, and the Jacobian
    (node part of the fields) equals central finite differences of U."""
    p = make_problem(8, 8, 0.3)
    o = _cem_oracle(p, 0.01)
    g = inclusion_phantom(p["nodes"], p["elements"], frame_seed(1, 7))
    U, phis, _ = o.solve(g)
    np.testing.assert_allclose(U, U.T, atol=1e-12 * np.max(np.abs(U)))
    J = o.jacobian(g, phis)
    elems = [0, 9, 60, o.E - 1]
    Jfd = _fd_jacobian(o, g, elems)
    np.testing.assert_allclose(J[:, elems], Jfd, atol=1e-7 * np.max(np.abs(Jfd)))


def test_cem_approaches_shunt_limit():
    """As z -> 0 the electrode potentials become the nodal potentials on the electrode
    (shunt model): the spread of phi over each electrode shrinks in proportion to z."""
    p = make_problem(8, 8, 0.3)
    g = np.ones(p["elements"].shape[0])
    spreads = []
    for z in (1e-2, 1e-4):
        o = _cem_oracle(p, z)
        X = eo.solve_direct(o.system(g, 0), o.B @ o.Pi, True)
        s = 0.0
        for l in range(o.L):
            nodes_l = np.unique(p["boundary"][[e for e in p["ports"][l] if e >= 0]])
            s = max(s, np.max(np.abs(X[nodes_l, 0] - X[o.N + l, 0])))
        spreads.append(s)
    assert spreads[1] < 0.05 * spreads[0]


# ----------------------------------------------------------------------------------
# two-level preconditioner (product extension) restated in the oracle
# ----------------------------------------------------------------------------------
def _two_level_brute(A, A1, Sd, S1d, dc, tile):
    """Dense, loop-by-loop construction of D^-1 + W C E1^-1 C W^T over the aggregates
    'rows [q tile / 4, (q + 1) tile / 4) of each tile', q = 0 .. 3."""
    N = A.shape[0]
    Ad, A1d = A.toarray(), A1.toarray()
    groups = []
    for start in range(0, N, tile):
        rows = list(range(start, min(N, start + tile)))
        for q in range(4):
            part = rows[q * tile // 4:(q + 1) * tile // 4]
            if part:
                groups.append(part)
    na = len(groups)
    a_of = [0] * N
    for a, rows in enumerate(groups):
        for i in rows:
            a_of[i] = a
    E1 = np.zeros((na, na))
    for i in range(N):
        for j in range(N):
            E1[a_of[i], a_of[j]] += A1d[i, j] + (1.0 if dc else 0.0)
    c = np.zeros(na)
    for a, rows in enumerate(groups):
        c[a] = (sum(Sd[i] for i in rows) / sum(S1d[i] for i in rows)) ** -0.5
    Ei = np.linalg.inv(E1)
    M = np.diag(1.0 / np.diag(Ad))
    for i in range(N):
        for j in range(N):
            M[i, j] += c[a_of[i]] * Ei[a_of[i], a_of[j]] * c[a_of[j]]
    return M


@pytest.mark.parametrize("c", [0, 1])
def test_two_level_preconditioner_spd_and_brute_force(c):
    p = make_problem(6, 8, 0.2, C=2, height=0.5, port_height=0.2)
    o = oracle_of(p)
    g = inclusion_phantom(p["nodes"], p["elements"], frame_seed(1, 4))
    A, A1 = o.system(g, c), o.system(np.ones(o.E), c)
    Sd, S1d = o.system(g, 0).diagonal(), o.system(np.ones(o.E), 0).diagonal()
    prec, Minv = eo.two_level_preconditioner(A, A1, Sd, S1d, c == 0, tile=16)
    assert Minv is not None
    np.testing.assert_allclose(Minv, Minv.T, atol=1e-14 * np.abs(Minv).max())
    assert np.linalg.eigvalsh(Minv).min() > 0
    np.testing.assert_allclose(Minv, _two_level_brute(A, A1, Sd, S1d, c == 0, 16), rtol=1e-10,
                               atol=1e-12 * np.abs(Minv).max())
    r = np.random.default_rng(3).standard_normal((o.N, 3))
    np.testing.assert_allclose(prec(r), Minv @ r, rtol=1e-10, atol=1e-12 * np.abs(Minv @ r).max())


def test_two_level_single_aggregate_closed_form(small):
    """One aggregate on the singular harmonic: A1 1 = 0, so E1 = N^2 and the coarse term is
    c^2 1 (1^T r) / N^2 with c^2 = sum S1_ii / sum S_ii."""
    p, o, g = small
    A, A1 = o.system(g, 0), o.system(np.ones(o.E), 0)
    Sd, S1d = A.diagonal(), A1.diagonal()
    prec, _ = eo.two_level_preconditioner(A, A1, Sd, S1d, True, tile=o.N, parts=1)
    r = np.random.default_rng(5).standard_normal((o.N, 2))
    c2 = S1d.sum() / Sd.sum()
    want = r / A.diagonal()[:, None] + c2 * np.ones((o.N, 1)) * r.sum(axis=0, keepdims=True) / o.N ** 2
    np.testing.assert_allclose(prec(r), want, rtol=1e-10)


def test_two_level_pcg_converges_and_saves_iterations():
    # aggregates of consecutive rows are compact patches after the RCB renumbering
    p = rcb_renumbered(make_problem(24, 16, 0.1, C=2, height=0.4, port_height=0.1), 64)
    o = oracle_of(p)
    g = inclusion_phantom(p["nodes"], p["elements"], frame_seed(1, 6))
    ones = np.ones(o.E)
    Sd, S1d = o.system(g, 0).diagonal(), o.system(ones, 0).diagonal()
    for c in range(2):
        A = o.system(g, c)
        F = o.alpha_beta(c)[1] * (o.B @ o.Pi)[:, :4]
        X = eo.solve_direct(A, F, c == 0)
        prec, _ = eo.two_level_preconditioner(A, o.system(ones, c), Sd, S1d, c == 0, tile=64)
        Xt, st = eo.pcg_reference(A, F, None, c == 0, rule="tolerance", tol=1e-10, prec=prec)
        _, sj = eo.pcg_reference(A, F, 1.0 / A.diagonal(), c == 0, rule="tolerance", tol=1e-10)
        assert np.max(np.abs(Xt - X)) < 1e-8 * np.max(np.abs(X))
        # the stiffness-dominated harmonic 0 is where the coarse level pays; on the
        # mass-weighted harmonic 1 of a small mesh Jacobi alone is already good
        if c == 0:
            assert st < 0.7 * sj
        else:
            assert st < 1.2 * sj


def test_cpu_baseline_two_level_matches_oracle():
    """The host baseline's coarse term reproduces the oracle's two-level PCG."""
    from oracle import baseline
    p = rcb_renumbered(make_problem(24, 16, 0.1, C=1), 64)
    o = oracle_of(p)
    g = inclusion_phantom(p["nodes"], p["elements"], frame_seed(1, 6))
    A, A1 = o.system(g, 0), o.system(np.ones(o.E), 0)
    prec, _ = eo.two_level_preconditioner(A, A1, A.diagonal(), A1.diagonal(), True, tile=64)
    F = o.alpha_beta(0)[1] * (o.B @ o.Pi)[:, :8]
    X = eo.solve_direct(A, F, True)
    _, so = eo.pcg_reference(A, F, None, True, rule="tolerance", tol=1e-6, prec=prec)
    Xb, sb, _ = baseline.pcg(A, F, 1, tol=1e-6, einv=prec.einv, tile=64)
    _, sj, _ = baseline.pcg(A, F, 1, tol=1e-6)
    assert np.max(np.abs(Xb - X)) < 1e-5 * np.max(np.abs(X))
    # FP32 CG takes more steps than the FP64 oracle, the more so the stronger the
    # preconditioner (here ~35% on 1801 nodes)
    assert abs(sb - so) <= 0.4 * so and sb < 0.7 * sj


def test_regularized_gn_step_pins():
    """R = I reduces to the data-space step; for a diagonal R the E-space solution equals
    the column-scaled data-space form delta = W Js^T (Js Js^T + lam I)^-1 r, Js = J W,
    W = R^-1/2 (the form the device uses)."""
    rng = np.random.default_rng(9)
    J = rng.standard_normal((12, 30))
    r = rng.standard_normal(12)
    lam = 0.3
    np.testing.assert_allclose(eo.gn_step_regularized(J, r, lam, np.ones(30)), eo.gn_step_exact(J, r, lam),
                               rtol=1e-10, atol=1e-12)
    d = eo.noser_diag(J)
    np.testing.assert_allclose(d, np.diag(J.T @ J), rtol=1e-12)
    w = 1.0 / np.sqrt(d)
    Js = J * w[None, :]
    want = w * eo.gn_step_exact(Js, r, lam)
    np.testing.assert_allclose(eo.gn_step_regularized(J, r, lam, d), want, rtol=1e-9, atol=1e-12)


def test_laplacian_matrix_pins():
    """R = L^T L of inverse.cpp:96-117 against its quadratic form written from the edge
    list: x^T R x = sum over interior edges |e|^2 (x_a - x_b)^2 + sum over boundary edges
    |e|^2 x_a^2, so R 1 carries the boundary-edge lengths only; also P2 meshes use their
    vertex edges."""
    nodes, el = ring_mesh(4, 8, 0.2)
    R = eo.laplacian_matrix(nodes, el)
    E = el.shape[0]
    assert np.allclose(R, R.T)
    users = {}
    for e in range(E):
        for i in range(3):
            a, b = int(el[e, i]), int(el[e, (i + 1) % 3])
            users.setdefault((min(a, b), max(a, b)), []).append(e)
    rng = np.random.default_rng(5)
    for _ in range(3):
        x = rng.standard_normal(E)
        q = 0.0
        for (a, b), es in users.items():
            l2 = float(np.sum((nodes[a] - nodes[b]) ** 2))
            q += l2 * ((x[es[0]] - x[es[1]]) ** 2 if len(es) == 2 else x[es[0]] ** 2)
        assert math.isclose(float(x @ R @ x), q, rel_tol=1e-12)
    bd = np.zeros(E)
    for (a, b), es in users.items():
        if len(es) == 1:
            bd[es[0]] += float(np.sum((nodes[a] - nodes[b]) ** 2))
    np.testing.assert_allclose(R @ np.ones(E), bd, atol=1e-14)
    assert np.all(np.linalg.eigvalsh(R) > 0.0)
    n2, e2 = promote_p2(nodes, el)
    np.testing.assert_allclose(eo.laplacian_matrix(n2, e2), R, rtol=1e-12, atol=1e-15)
    # the general step with R = I is the identity-regularised step
    J = rng.standard_normal((10, E))
    r = rng.standard_normal(10)
    np.testing.assert_allclose(eo.gn_step_general(J, r, 0.2, np.eye(E)), eo.gn_step_exact(J, r, 0.2),
                               rtol=1e-9, atol=1e-12)


def _voltage_oracle(C=1, rings=5, L=8, width=0.3, drive=None, scale=1.0, p2=False):
    nodes, el = ring_mesh(rings, L, width)
    if p2:
        nodes, el = promote_p2(nodes, el)
    bd = boundary_edges(el)
    ports = eo.circular_boundary_ports(nodes, el, bd, L, width).astype(np.int32)
    pat = eo.adjacent_pattern(L) if drive is None else drive
    o = eo.EITOracle(nodes, el, bd, ports, pat, pat, 1.0, scale, C, 0.4 if C > 1 else 1.0,
                     0.1 if C > 1 else 1.0, voltage=True)
    return o, nodes, el


def test_voltage_sources_pins():
    """Fixed sources (eit.cpp:89-91,159-205; eit_kernel.cu:34-54) against what the physics
    fixes: the row-replaced solve equals the symmetric elimination of the Dirichlet nodes;
    the electrode-current map is symmetric (reciprocity, drive = measurement pattern);
    equal potentials on every electrode draw no current at harmonic 0; the currents scale
    with sigma; and the Jacobian is their finite-difference derivative (C = 1)."""
    o, nodes, el = _voltage_oracle(C=2)
    g = ellipse_phantom(nodes, el, frame_seed(3, 1))
    U, phis, _ = o.solve(g)
    D = o.dirichlet
    I = np.setdiff1d(np.arange(o.N), D)
    A = o.system(g, 1)
    b = o.B @ o.Pi
    x = np.zeros_like(b)
    x[D] = b[D]
    x[I] = spla.spsolve(A[I][:, I].tocsc(), -(A[I][:, D] @ b[D]))
    np.testing.assert_allclose(phis[1], x, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(U, U.T, rtol=1e-9, atol=1e-12 * np.abs(U).max())
    U2, _, _ = o.solve(2.0 * g)
    np.testing.assert_allclose(U2, 2.0 * U, rtol=1e-10, atol=1e-14)
    # one drive column puts 1 V on every electrode: no current at harmonic 0
    dr = np.ones((8, 1))
    oc, _, _ = _voltage_oracle(C=1, drive=dr)
    Uc, _, _ = oc.solve(g)
    assert np.abs(Uc).max() < 1e-10
    # Jacobian = d U / d gamma (2D, one harmonic; linear parametrisation, ref = 1.3)
    o1, _, _ = _voltage_oracle(C=1, scale=1.3)
    U1, ph1, _ = o1.solve(g)
    J = o1.jacobian(g, ph1)
    eps = 1e-4          # central differences: truncation ~eps^2, solver round-off ~1e-14 / eps
    for e in (0, 7, len(g) // 2, len(g) - 1):
        gp, gm = g.copy(), g.copy()
        gp[e] += eps
        gm[e] -= eps
        fd = (o1.flatten_voltages(o1.solve(gp)[0]) - o1.flatten_voltages(o1.solve(gm)[0])) / (2 * eps)
        np.testing.assert_allclose(J[:, e], fd, rtol=1e-6, atol=1e-5 * np.abs(J[:, e]).max())


# ---------------------------------------------------------------------------
# complex conductivity (Equation<thrust::complex<float>>, src/models/eit.cpp:289-304)
# ---------------------------------------------------------------------------

def _complex_oracle(C=1, rings=6, L=8, w=0.3, p2=False, h=1.0, hp=1.0):
    from oracle.eit_oracle import ComplexEITOracle
    p = make_problem(rings, L, w, C=C, height=h, port_height=hp, p2=p2)
    o = ComplexEITOracle(p["nodes"], p["elements"], p["boundary"], p["ports"], p["drive"], p["meas"],
                         p["current"], 1.0, C, h, hp)
    return p, o


def _complex_gamma(E, seed=5, imag=0.3):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.5, 2.0, E) + 1j * imag * rng.uniform(-1.0, 1.0, E)


@pytest.mark.parametrize("C,h,hp", [(1, 1.0, 1.0), (3, 0.4, 0.1)])
def test_complex_real_conductivity_equals_real_model(C, h, hp):
    """A real gamma gives the real model's voltages and Jacobian (imaginary parts 0)."""
    p, o = _complex_oracle(C, h=h, hp=hp)
    r = oracle_of(p)
    g = np.random.default_rng(2).uniform(0.5, 2.0, o.E)
    U, ph, _ = o.solve(g)
    U0, ph0, _ = r.solve(g)
    np.testing.assert_allclose(U.real, U0, rtol=1e-12, atol=1e-15 * np.abs(U0).max())
    assert np.abs(U.imag).max() < 1e-14 * np.abs(U0).max()
    for q in (True, False):
        J, J0 = o.jacobian(g, ph, q), r.jacobian(g, ph0, q)
        np.testing.assert_allclose(J.real, J0, rtol=1e-11, atol=1e-14 * np.abs(J0).max())


def test_complex_holomorphic_jacobian():
    """U(gamma) is holomorphic: central differences along a real and along an imaginary
    step of gamma_e both give J[:, e] (Cauchy-Riemann).  A Jacobian with conjugated fields
    would match the real step only on real data and fail the imaginary one."""
    p, o = _complex_oracle(C=2, h=0.5, hp=0.2)
    g = _complex_gamma(o.E)
    U, ph, _ = o.solve(g)
    J = o.jacobian(g, ph, quirks=False)
    eps = 1e-6
    for e in (0, o.E // 3, o.E - 1):
        for step in (eps, 1j * eps):
            gp_, gm_ = g.copy(), g.copy()
            gp_[e] += step
            gm_[e] -= step
            d = (o.flatten_voltages(o.solve(gp_)[0]) - o.flatten_voltages(o.solve(gm_)[0])) / (2 * step)
            assert np.abs(d - J[:, e]).max() < 1e-6 * np.abs(J[:, e]).max()


def test_complex_reciprocity():
    """Adjacent drive = measurement patterns: U = U^T, not U^H (the system is complex
    symmetric, not Hermitian)."""
    p, o = _complex_oracle(C=1)
    g = _complex_gamma(o.E, seed=7, imag=0.8)
    U, ph, _ = o.solve(g)
    np.testing.assert_allclose(U, U.T, rtol=0, atol=1e-10 * np.abs(U).max())
    assert np.abs(U - U.conj().T).max() > 1e-3 * np.abs(U).max()


def test_complex_pcg_hermitian_dots():
    """The reference's complex CG conjugates the second dot operand.  On a Hermitian
    positive definite matrix that is textbook CG: the exact solution within n steps (dots
    without the conjugate do not converge there); on the EIT system with a small imaginary
    part it reaches the direct solve."""
    from oracle.eit_oracle import pcg_reference, solve_direct
    rng = np.random.default_rng(3)
    n = 12
    Z = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    H = Z @ Z.conj().T + n * np.eye(n)
    b = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    x, s = pcg_reference(sp.csr_matrix(H), b, rule="tolerance", tol=1e-12, max_iter=3 * n)
    np.testing.assert_allclose(x, np.linalg.solve(H, b), rtol=1e-9, atol=1e-11)
    assert s <= n + 2
    p, o = _complex_oracle(C=1)
    g = _complex_gamma(o.E, imag=0.05)
    A = o.system(g, 0)
    F = (o.B @ o.Pi).astype(complex)
    xc, _ = pcg_reference(A, F, 1.0 / A.diagonal(), True, rule="tolerance", tol=1e-10, max_iter=2000)
    xd = solve_direct(A, F, True)
    assert np.abs(xc - xd).max() < 1e-7 * np.abs(xd).max()


def test_complex_cocg_bilinear_dots():
    """Conjugate orthogonal CG (``bilinear``): on a complex symmetric, non-Hermitian
    matrix the exact solution within n steps, where the conjugated iteration does not
    converge; on real data both are the reference's CG; on the EIT system with a large
    imaginary part it reaches the direct solve (the dcFree harmonic included)."""
    from oracle.eit_oracle import pcg_reference, solve_direct
    rng = np.random.default_rng(8)
    n = 12
    Bm = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = Bm @ Bm.T + n * np.eye(n)
    assert np.abs(A - A.T).max() == 0 and np.abs(A - A.conj().T).max() > 1.0
    b = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    x, s = pcg_reference(sp.csr_matrix(A), b, rule="tolerance", tol=1e-12, max_iter=3 * n, bilinear=True)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-8, atol=1e-10)
    assert s <= n + 4
    xh, _ = pcg_reference(sp.csr_matrix(A), b, rule="tolerance", tol=1e-12, max_iter=3 * n)
    assert np.abs(xh - np.linalg.solve(A, b)).max() > 1e-4
    Ar = sp.csr_matrix(np.real(Bm @ Bm.T) + n * np.eye(n))
    xr0, s0 = pcg_reference(Ar, np.real(b), rule="tolerance", tol=1e-10)
    xr1, s1 = pcg_reference(Ar, np.real(b), rule="tolerance", tol=1e-10, bilinear=True)
    assert s0 == s1 and np.array_equal(xr0, xr1)
    p, o = _complex_oracle(C=1)
    g = _complex_gamma(o.E, imag=1.0)
    A = o.system(g, 0)
    F = (o.B @ o.Pi).astype(complex)
    xc, _ = pcg_reference(A, F, 1.0 / A.diagonal(), True, rule="tolerance", tol=1e-10, max_iter=2000, bilinear=True)
    xd = solve_direct(A, F, True)
    assert np.abs(xc - xd).max() < 1e-7 * np.abs(xd).max()


def test_complex_gn_step_plain_transpose():
    """Inverse::solve forms J^T J and J^T r with plain transposes (CUBLAS_OP_T): the data-
    space step equals the E x E one, equals the real step for real data, and differs
    from the Hermitian (J^H) least-squares step for complex J."""
    from oracle.eit_oracle import gn_step_exact, gn_step_normal
    rng = np.random.default_rng(4)
    m, E = 20, 60
    J = rng.standard_normal((m, E)) + 0.3j * rng.standard_normal((m, E))
    r = rng.standard_normal(m) + 0.3j * rng.standard_normal(m)
    d = gn_step_exact(J, r, 0.5)
    np.testing.assert_allclose(d, gn_step_normal(J, r, 0.5), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(J.T @ J @ d + 0.5 * d, J.T @ r, rtol=1e-10, atol=1e-12)
    dh = np.linalg.solve(J.conj().T @ J + 0.5 * np.eye(E), J.conj().T @ r)
    assert np.abs(d - dh).max() > 1e-3 * np.abs(d).max()
    Jr, rr = J.real, r.real
    np.testing.assert_allclose(gn_step_exact(Jr + 0j, rr + 0j, 0.5).real, gn_step_exact(Jr, rr, 0.5), rtol=1e-10)
