"""FP64 CPU oracle for the EIT forward/inverse path of mpFlow.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the checker.
The product path (``mpflow_amd``) never calls it.

This module restates, in float64 numpy/scipy, the algorithm of the reference files
cited at each function (paths relative to the mpFlow source tree).  The reference
cannot be compiled here (CUDA 7 + cuBLAS + Eigen + libdistmesh are absent), so this
restatement is the oracle; it is pinned by the checks listed in
``tests/test_oracle.py`` (closed forms, invariants, brute force, finite differences,
the analytic disc solution).

Two modes exist where the reference has quirks:
  * ``quirks=True`` reproduces the reference exactly (P2 boundary integral of node 2,
    Jacobian with K_e only and the single factor ``height``);
  * ``quirks=False`` is the mathematically correct variant.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

FLT_EPSILON = float(np.finfo(np.float32).eps)

# ---------------------------------------------------------------------------
# element integrals  (src/fem/basis_linear.cpp:27-91, src/fem/basis_quadratic.cpp:27-97)
# ---------------------------------------------------------------------------

# Strang-Fix 6-point rule, exact to degree 4 on the reference triangle
# (weights sum to 1, points in barycentric coordinates).
_Q4_A, _Q4_B = 0.445948490915965, 0.091576213509771
_Q4_WA, _Q4_WB = 0.223381589678011, 0.109951743655322
_Q4 = np.array([
    [_Q4_A, _Q4_A, 1 - 2 * _Q4_A, _Q4_WA],
    [_Q4_A, 1 - 2 * _Q4_A, _Q4_A, _Q4_WA],
    [1 - 2 * _Q4_A, _Q4_A, _Q4_A, _Q4_WA],
    [_Q4_B, _Q4_B, 1 - 2 * _Q4_B, _Q4_WB],
    [_Q4_B, 1 - 2 * _Q4_B, _Q4_B, _Q4_WB],
    [1 - 2 * _Q4_B, _Q4_B, _Q4_B, _Q4_WB],
])


def _monomials(x, y, k):
    if k == 3:
        return np.array([np.ones_like(x), x, y])
    return np.array([np.ones_like(x), x, y, x * x, y * y, x * y])


def _monomial_grads(x, y, k):
    one, zero = np.ones_like(x), np.zeros_like(x)
    if k == 3:
        gx = np.array([zero, one, zero])
        gy = np.array([zero, zero, one])
    else:
        gx = np.array([zero, one, zero, 2 * x, zero, y])
        gy = np.array([zero, zero, one, zero, 2 * y, x])
    return gx, gy


def basis_coefficients(xy: np.ndarray) -> np.ndarray:
    """Coefficients of the nodal basis in global monomials, column i = basis of node i.

    The reference solves the Vandermonde system per basis function by Gaussian
    elimination (``basis_linear.cpp:27-56``, ``basis_quadratic.cpp:27-59``,
    ``include/mpflow/mathematics.h:59-110``); monomials are {1,x,y} (P1) and
    {1,x,y,x^2,y^2,xy} (P2).
    """
    k = xy.shape[0]
    V = _monomials(xy[:, 0], xy[:, 1], k).T  # V[node, monomial]
    return np.linalg.solve(V, np.eye(k))


def element_integrals(xy: np.ndarray):
    """(K_e, M_e): K[i,j] = int grad(phi_i).grad(phi_j), M[i,j] = int phi_i phi_j,
    integrated exactly (degree-4 rule) over the straight triangle ``xy[:3]``."""
    k = xy.shape[0]
    C = basis_coefficients(xy)
    p0, p1, p2 = xy[0], xy[1], xy[2]
    area = 0.5 * abs((p1[0] - p0[0]) * (p2[1] - p0[1]) - (p1[1] - p0[1]) * (p2[0] - p0[0]))
    lam = _Q4[:, :3]
    w = _Q4[:, 3] * area
    X = lam @ np.array([p0[0], p1[0], p2[0]])
    Y = lam @ np.array([p0[1], p1[1], p2[1]])
    mon = _monomials(X, Y, k)            # [mono, q]
    gx, gy = _monomial_grads(X, Y, k)
    phi = C.T @ mon                      # [node, q]
    dphx, dphy = C.T @ gx, C.T @ gy
    K = (dphx * w) @ dphx.T + (dphy * w) @ dphy.T
    M = (phi * w) @ phi.T
    return K, M


def all_element_integrals(nodes, elements):
    E, k = elements.shape
    K = np.empty((E, k, k))
    M = np.empty((E, k, k))
    for e in range(E):
        K[e], M[e] = element_integrals(nodes[elements[e]])
    return K, M


def boundary_integral(s: np.ndarray, node: int, quirks: bool = True) -> float:
    """int over the edge parameter [s_0, s_last] of the 1D Lagrange basis of ``node``.

    ``s`` holds the distances of the edge nodes from the first edge node.  The
    reference's generated P2 code evaluates node 2 with node 1's coefficients
    (``src/fem/codegen/templates/basis.cpp.mako:72-84`` rendered at
    ``src/fem/basis_quadratic.cpp:135-147``); ``quirks=True`` reproduces that.
    P1 has two identical branches that both integrate to L/2 (``basis_linear.cpp:95-124``).
    """
    n = s.size
    if n == 2:
        node = 0  # both branches of the generated code use node 0's coefficients
    elif quirks and node == n - 1:
        node = n - 2
    V = np.vander(s, n, increasing=True)          # V[i, p] = s_i^p
    c = np.linalg.solve(V, np.eye(n)[:, node])     # coefficients of the basis
    a, b = s[0], s[-1]
    return float(sum(c[p] * (b ** (p + 1) - a ** (p + 1)) / (p + 1) for p in range(n)))


# ---------------------------------------------------------------------------
# ports  (src/fem/ports.cpp:32-77, src/mathematics.cpp:26-99)
# ---------------------------------------------------------------------------

def _polar_angle(x, y):
    # math::polar (src/mathematics.cpp:26-55)
    if x > 0.0:
        return math.atan(y / x)
    if x < 0.0 and y >= 0.0:
        return math.atan(y / x) + math.pi
    if x < 0.0 and y < 0.0:
        return math.atan(y / x) - math.pi
    if x == 0.0 and y > 0.0:
        return math.pi / 2
    if x == 0.0 and y < 0.0:
        return -math.pi / 2
    return 0.0


def _circle_parameter(pt, offset):
    # math::circleParameter (src/mathematics.cpp:64-77)
    r = math.hypot(pt[0], pt[1])
    ang = _polar_angle(pt[0], pt[1]) - offset / r
    ang += 2 * math.pi if ang < -math.pi else 0.0
    ang -= 2 * math.pi if ang > math.pi else 0.0
    return ang * r


def circular_boundary_ports(nodes, elements, boundary, count, width, offset=0.0,
                            clockwise=False):
    """Ports::circularBoundary: edges of ``boundary`` inside each electrode arc.

    Returns a ``count x maxEdges`` int array, -1 for unused slots (the reference's
    ``invalidIndex``).  ``boundary`` holds counter-clockwise boundary edges as rows of
    node indices; the returned entries index rows of ``boundary``.
    """
    radius = float(np.max(np.hypot(nodes[elements[:, :3], 0], nodes[elements[:, :3], 1])))
    dist = 2.0 * math.pi * radius / count
    def circ_points(off):
        phi0 = off / radius
        n = int(math.ceil((2 * math.pi) / (dist / radius)))
        phi = phi0 + np.arange(n) * (dist / radius)
        sgn = -1.0 if clockwise else 1.0
        return np.stack([radius * np.cos(phi), sgn * radius * np.sin(phi)], axis=1)
    ps = circ_points(offset + (width if clockwise else 0.0))
    pe = circ_points(offset + (0.0 if clockwise else width))
    out = []
    for port in range(count):
        sel = []
        for ei, edge in enumerate(boundary):
            pts = nodes[edge]
            par = np.sqrt(np.sum((pts - pts[0]) ** 2, axis=1))
            off0 = _circle_parameter(pts[0], 0.0)
            a = _circle_parameter(ps[port], off0)
            b = _circle_parameter(pe[port], off0)
            if a < b and par[0] - a >= -1e-9 and par[-1] - b <= 1e-9:
                sel.append(ei)
        out.append(sel)
    m = max(len(s) for s in out)
    arr = -np.ones((count, m), dtype=np.int64)
    for p, s in enumerate(out):
        arr[p, :len(s)] = s
    return arr


# ---------------------------------------------------------------------------
# assembly, excitation, pattern  (src/fem/equation.cpp:76-253, equation_kernel.cu:35-157,
#                                  src/fem/sources.cpp:27-75)
# ---------------------------------------------------------------------------

def mat_param(gamma, ref, logarithmic=False):
    """sigma(gamma) (equation_kernel.cu:35-43)."""
    if logarithmic:
        return ref * np.exp(np.log(10.0) * gamma / 10.0)
    return ref * gamma


def mat_param_deriv(gamma, ref, logarithmic=False):
    """d sigma / d gamma (equation_kernel.cu:45-53)."""
    if logarithmic:
        return (np.log(10.0) / 10.0) * ref * np.exp(np.log(10.0) * gamma / 10.0)
    return ref * np.ones_like(gamma)


def assemble(elements, Ke, N, weights):
    """sum_e w_e * Ke[e] scattered to the global N x N matrix (CSR, float64)."""
    E, k = elements.shape
    rows = np.repeat(elements, k, axis=1).ravel()
    cols = np.tile(elements, (1, k)).ravel()
    vals = (Ke * weights[:, None, None]).reshape(E, k * k).ravel()
    return sp.csr_matrix((vals, (rows, cols)), shape=(N, N))


def excitation_matrix(nodes, boundary, ports, quirks=True):
    """Gap-model excitation B (N x P): B[n,p] += int_edge phi_n / |port p|
    (Equation::initExcitationMatrix, src/fem/equation.cpp:183-215)."""
    N = nodes.shape[0]
    P = ports.shape[0]
    B = np.zeros((N, P))
    for p in range(P):
        edges = [e for e in ports[p] if e >= 0]
        length = 0.0
        for e in edges:
            pts = nodes[boundary[e]]
            length += math.sqrt(np.sum((pts[-1] - pts[0]) ** 2))
        for e in edges:
            pts = nodes[boundary[e]]
            s = np.sqrt(np.sum((pts - pts[0]) ** 2, axis=1))
            for node in range(len(boundary[e])):
                B[boundary[e][node], p] += boundary_integral(s, node, quirks) / length
    return B


def edge_integrals(s: np.ndarray):
    """Exact 1D Lagrange-basis integrals on one straight boundary edge whose nodes sit
    at distances ``s`` (s[0] = 0) from its first node: (m, v) with m[a, b] = int phi_a
    phi_b ds and v[a] = int phi_a ds over [0, s_last] (3-point Gauss-Legendre, exact for
    the degree <= 4 products of P2)."""
    n = s.size
    V = np.vander(s, n, increasing=True)
    coef = np.linalg.solve(V, np.eye(n))            # column a: coefficients of phi_a
    xg, wg = np.polynomial.legendre.leggauss(3)
    L = s[-1]
    t = 0.5 * L * (xg + 1.0)
    w = 0.5 * L * wg
    phi = np.vander(t, n, increasing=True) @ coef   # phi[q, a]
    return (phi * w[:, None]).T @ phi, (phi * w[:, None]).sum(axis=0)


def cem_blocks(nodes, boundary, ports, z):
    """Complete-electrode-model blocks for contact impedance z on every port (an
    extension: the reference has the gap model only).  With electrode potentials V_l as
    extra unknowns the system is

        [ A + Zc   -Cc ] [phi]   [ 0 ]
        [ -Cc^T     Dc ] [ V ] = [ I ]

    Zc[i, j] = sum_l (1/z) int_{e_l} phi_i phi_j, Cc[i, l] = (1/z) int_{e_l} phi_i,
    Dc[l, l] = |e_l| / z.  Returns (Zc as CSR N x N, Cc dense N x L, Dc diagonal (L,))."""
    N = nodes.shape[0]
    L = ports.shape[0]
    rows, cols, vals = [], [], []
    Cc = np.zeros((N, L))
    Dc = np.zeros(L)
    for l in range(L):
        for e in ports[l]:
            if e < 0:
                continue
            ed = boundary[e]
            pts = nodes[ed]
            sd = np.sqrt(np.sum((pts - pts[0]) ** 2, axis=1))
            m, v = edge_integrals(sd)
            for a in range(len(ed)):
                Cc[ed[a], l] += v[a] / z
                for b_ in range(len(ed)):
                    rows.append(ed[a]); cols.append(ed[b_]); vals.append(m[a, b_] / z)
            Dc[l] += sd[-1] / z
    Zc = sp.csr_matrix((vals, (rows, cols)), shape=(N, N))
    return Zc, Cc, Dc


def adjacent_pattern(L):
    """Column d = e_d - e_{d+1 mod L} (drive and measurement are both adjacent)."""
    P = np.zeros((L, L))
    for d in range(L):
        P[d, d] = 1.0
        P[(d + 1) % L, d] = -1.0
    return P


def pattern_matrix(drive, meas, current):
    """Pi = [I * Dpat | Mpat] (Sources::Sources, src/fem/sources.cpp:57-71)."""
    cur = np.broadcast_to(np.asarray(current, dtype=np.float64), (drive.shape[1],))
    return np.concatenate([drive * cur[None, :], meas], axis=1)


def ceil16(n):
    """Matrix padding rule (src/numeric/matrix.cpp:49-55, include/mpflow/mathematics.h:42-44)."""
    return n if (n % 16 == 0 or n == 1) else (n // 16 + 1) * 16


# ---------------------------------------------------------------------------
# PCG restatement (src/numeric/conjugate_gradient.cpp:61-161, conjugate_gradient_kernel.cu)
# ---------------------------------------------------------------------------

def pcg_reference(A, b, dinv=None, dc_free=False, max_iter=0, eps=FLT_EPSILON,
                  rule="reference", tol=1e-6, prec=None, x0=None, fp32=False, bilinear=False):
    """Multi-RHS Jacobi-PCG as ``ConjugateGradient::solve`` runs it, in float64.

    * x starts at ``x0`` (default 0) with r = b - A x0 (conjugate_gradient.cpp:80-82);
      dcFree adds ``1 1^T`` to the operator, and the start residual then gets
      ``+ 1 (1^T x0)`` -- addScalar adds the column sums of x0 (:84-88), the opposite
      sign of the operator's term, so a nonzero start shifts the solution by
      ``2 (1^T x0) / N`` when ``A 1 = 0``.
    * alpha = rho_old / (p.q) and beta = rho / rho_old per column; a zero denominator
      makes the updated vector 0 (``updateVectorKernel``, conjugate_gradient_kernel.cu:77-89).
    * ``rule="reference"``: stop after step s when sqrt(|rho[0]| / (b.b)[0]) < eps
      (column 0 only) and return s+1; otherwise stop after 10*N steps.
      ``rule="tolerance"``: stop when every column has ||r||/||b|| < tol.
    * ``prec`` (optional callable r -> z) replaces the diagonal preconditioner (the
      product's two-level preconditioner, ``two_level_preconditioner``).
    * complex A or b (``ConjugateGradient<thrust::complex<float>>``,
      conjugate_gradient.cpp:231-232): every dot product conjugates its second operand
      (vectorDotProductKernel, matrix_kernel.cu:405-413): rho = sum r conj(z),
      p.q = sum p conj(q), b.b = sum b conj(b); the stop rule takes |sqrt(rho / b.b)|.
      ``bilinear``: the product's alternative for complex symmetric systems, conjugate
      orthogonal CG (COCG): rho = r^T z and p.q = p^T q without conjugation (b.b and the
      tolerance rule keep the Hermitian norms).  Identical to the reference iteration
      for real data.
    * ``fp32``: the reference's own arithmetic -- matrix, right-hand side, vectors and the
      Jacobi inverse rounded to float32; A v summed row by row in float32 over the row's
      entries in stored order, each product rounded before it is added (the product's
      SpMV); the other vector operations rounded to float32 once, alpha and beta rounded
      to float32 before they scale a vector; dot products stay float64 sums of the
      float32 values (the FP64 scalars of the device solver).  Step counts of the
      FLT_EPSILON stop rule depend on this rounding.
    Returns (x, steps).
    """
    N, R = b.shape
    iters = 10 * N if max_iter == 0 else max_iter
    if dinv is None:
        dinv = np.ones(N)
    cplx = np.iscomplexobj(b) or np.iscomplexobj(dinv) or sp.csr_matrix(A).dtype.kind == "c"
    dt = np.complex128 if cplx else np.float64
    if fp32 and cplx:
        raise ValueError("fp32 emulation: real systems only")
    if fp32:
        rd = lambda v: np.asarray(v, np.float32).astype(np.float64)
        A = sp.csr_matrix(A).astype(np.float32).astype(np.float64)
        b = rd(b)
        dinv = rd(dinv)
        # the rows as (column, value) slots in stored order, padded with (row, 0)
        nnz_row = np.diff(A.indptr)
        width = int(nnz_row.max()) if N else 0
        ell_id = np.repeat(np.arange(N)[:, None], width, axis=1)
        ell_v = np.zeros((N, width))
        for i in range(N):
            s0, s1 = A.indptr[i], A.indptr[i + 1]
            ell_id[i, :s1 - s0] = A.indices[s0:s1]
            ell_v[i, :s1 - s0] = A.data[s0:s1]

        def matvec(v):
            acc = np.zeros((N, v.shape[1]), np.float32)
            for q in range(width):
                acc = acc + (ell_v[:, q:q + 1] * v[ell_id[:, q]]).astype(np.float32)
            return acc.astype(np.float64)
    else:
        rd = lambda v: v
        matvec = lambda v: A @ v
    def op(v):
        y = matvec(v)
        if dc_free:
            y = y + v.sum(axis=0, keepdims=True)
        return rd(y)
    precond = prec if prec is not None else (lambda v: rd(dinv[:, None] * v))
    if x0 is None:
        x = np.zeros((N, R), dt)
        r = np.array(b, dt)
    else:
        x = np.array(x0, dt).reshape(N, R)
        r = b - A @ x
        if dc_free:
            r = r + x.sum(axis=0, keepdims=True)
    cj = (lambda v: v) if bilinear else np.conj
    z = precond(r)
    p = z.copy()
    rho_old = np.sum(r * cj(z), axis=0)
    ref = np.real(np.sum(b * np.conj(b), axis=0))
    for step in range(iters):
        q = op(p)
        pq = np.sum(p * cj(q), axis=0)
        with np.errstate(divide="ignore", invalid="ignore"):
            a = np.where(pq != 0.0, rho_old / np.where(pq != 0, pq, 1.0), 0.0)
        a = rd(a)
        r = rd(np.where(pq != 0.0, r - a * q, 0.0))
        x = rd(np.where(pq != 0.0, x + a * p, 0.0))
        z = precond(r)
        rho = np.sum(r * cj(z), axis=0)
        if rule == "reference":
            with np.errstate(divide="ignore", invalid="ignore"):
                crit = abs(math.sqrt(abs(rho[0]) / ref[0])) if ref[0] != 0 else float("nan")
            if not (crit >= eps):
                return x, step + 1
        else:
            rr = np.real(np.sum(r * np.conj(r), axis=0))
            ok = np.where(ref > 0, rr <= (tol * tol) * ref, True)
            if np.all(ok):
                return x, step + 1
        with np.errstate(divide="ignore", invalid="ignore"):
            beta = rd(np.where(rho_old != 0.0, rho / np.where(rho_old != 0, rho_old, 1.0), 0.0))
        p = rd(np.where(rho_old != 0.0, z + beta * p, 0.0))
        rho_old = rho
    return x, iters


def _pivoted_cholesky_orth(W, drop):
    """T with W T orthonormal on the columns of W that keep more than ``drop`` of their
    squared norm after projection on the columns chosen before them (greedy pivoting on
    the largest kept fraction); the kept directions fill columns 0..k-1 of T.  Returns
    (T [Rp][Rp], k)."""
    G = W.T @ W
    Rp = G.shape[0]
    orig = np.diag(G).copy()
    S = G.copy()
    used = np.zeros(Rp, bool)
    piv, cols = [], []
    for _ in range(Rp):
        frac = np.where(~used & (orig > 0), np.diag(S) / np.where(orig > 0, orig, 1.0), -np.inf)
        p = int(np.argmax(frac))
        if not frac[p] > drop:
            break
        lpp = math.sqrt(S[p, p])
        used[p] = True
        l = np.where(used & (np.arange(Rp) != p), 0.0, S[:, p] / lpp)
        rem = ~used
        S[np.ix_(rem, rem)] -= np.outer(l[rem], l[rem])
        piv.append(p)
        cols.append(l)
    k = len(piv)
    T = np.zeros((Rp, Rp))
    if k:
        L = np.array([[cols[m][piv[a]] if m <= a else 0.0 for m in range(k)] for a in range(k)])
        Tk = sla.solve_triangular(L.T, np.eye(k), lower=False)
        T[piv, :k] = Tk
    return T, k


def block_pcg(A, B, dinv=None, dc_free=False, tol=1e-6, rule="tolerance", max_iter=0, drop=1e-10,
              fp32=False):
    """Block Jacobi-PCG over the columns of B (the product's ``mpf_block_pcg_solve``; no
    counterpart in the reference, whose ConjugateGradient::solve,
    conjugate_gradient.cpp:61-161, iterates the columns independently).  O'Leary's block
    CG with the directions orthonormalised every iteration:

        R = B, Z = M R, P = orth(Z)
        repeat:  Q = A P; alpha = (P^T Q)^-1 P^T R; X += P alpha; R -= Q alpha
                 stop (rule); Z = M R; beta = -(P^T Q)^-1 Q^T Z; P = orth(Z + P beta)

    orth: ``_pivoted_cholesky_orth`` (directions keeping less than ``drop`` of their norm
    are dropped; the kept ones occupy the first k lanes, the others are zero and get a
    unit diagonal in P^T Q).  dc_free: the operator is A + 1 1^T.  Rules as
    ``pcg_reference``: "tolerance" (every column ||r|| <= tol ||b||) or "reference"
    (sqrt(|r.z| / b.b) < FLT_EPSILON on column 0).  ``fp32``: the device's precision
    split -- A, B, the Jacobi inverse, P and Q = A P rounded to float32, T rounded to
    float32 before it forms P; X, R, Z, W, alpha and beta float64.  The lane count is that
    of the device (16, 32 or 64 >= the columns).  Returns (X, steps)."""
    A = sp.csr_matrix(A)
    N, R = B.shape
    Rp = 16 if R <= 16 else (32 if R <= 32 else 64)
    iters = 10 * N if max_iter == 0 else max_iter
    if dinv is None:
        dinv = 1.0 / A.diagonal()
    rnd = (lambda v: v.astype(np.float32).astype(np.float64)) if fp32 else (lambda v: v)
    A_op = A.astype(np.float32) if fp32 else A
    dinv = rnd(np.asarray(dinv, np.float64))
    Bp = np.zeros((N, Rp))
    Bp[:, :R] = rnd(np.asarray(B, np.float64))
    bb = (Bp * Bp).sum(0)
    X = np.zeros((N, Rp))
    Rr = Bp.copy()
    Z = dinv[:, None] * Rr

    def done(Rr, Z):
        if rule == "reference":
            return bb[0] <= 0 or math.sqrt(abs(float((Rr[:, 0] * Z[:, 0]).sum())) / bb[0]) < FLT_EPSILON
        return bool(np.all((Rr[:, :R] ** 2).sum(0) <= tol * tol * bb[:R]))

    if done(Rr, Z):
        return X[:, :R], 0
    T, k = _pivoted_cholesky_orth(Z, drop)
    P = rnd(Z @ rnd(T))
    for it in range(1, iters + 1):
        Q = rnd(np.asarray(A_op @ P.astype(np.float32 if fp32 else np.float64), np.float64))
        cP = P.sum(0)
        PQ = P.T @ Q
        PQ = 0.5 * (PQ + PQ.T)
        if dc_free:
            PQ += np.outer(cP, cP)
        PQ[k:, :] = 0.0
        PQ[:, k:] = 0.0
        PQ[np.arange(k, Rp), np.arange(k, Rp)] = 1.0
        Lc = np.linalg.cholesky(PQ)
        solve = lambda M: sla.solve_triangular(Lc.T, sla.solve_triangular(Lc, M, lower=True), lower=False)
        alpha = solve(P.T @ Rr)
        X += P @ alpha
        Rr = Rr - Q @ alpha - (np.outer(np.ones(N), cP @ alpha) if dc_free else 0.0)
        Z = dinv[:, None] * Rr
        if done(Rr, Z):
            return X[:, :R], it
        QZ = Q.T @ Z
        if dc_free:
            QZ += np.outer(cP, Z.sum(0))
        beta = -solve(QZ)
        W = Z + P @ beta
        T, k = _pivoted_cholesky_orth(W, drop)
        if k == 0:
            raise np.linalg.LinAlgError("block PCG: no independent direction left")
        P = rnd(W @ rnd(T))
    return X[:, :R], iters


def rcb_order(xy, leaf=256, agg=64):
    """Node order of recursive coordinate bisection: a set is split at the median of its
    longer coordinate extent (ties by node id), the first part holding a whole number of
    ``leaf``-node leaves, until every leaf fits; each leaf is then split the same way into
    leaves of ``agg`` nodes, each ordered along its own longer extent.  Consecutive
    ``leaf`` nodes of the order are compact tiles and their consecutive ``agg``-node runs
    compact aggregates (``tile_aggregates``; the product's device tiles and aggregates are
    built the same way)."""
    xy = np.asarray(xy, np.float64)

    def along(idx):
        pts = xy[idx]
        ax = 0 if np.ptp(pts[:, 0]) >= np.ptp(pts[:, 1]) else 1
        return idx[np.lexsort((idx, pts[:, ax]))]

    def split(idx, lf):
        if len(idx) <= lf:
            return [along(idx)] if len(idx) > 1 else [idx]
        o = along(idx)
        left = (-(-len(idx) // lf) // 2) * lf
        return split(o[:left], lf) + split(o[left:], lf)

    return np.concatenate([a for t in split(np.arange(xy.shape[0]), leaf) for a in split(t, agg)])


def tile_aggregates(N, tile=256, parts=4):
    """Coarse aggregate of each of N rows: every ``tile``-row tile split into ``parts``
    runs of tile / parts consecutive rows, run q of tile t being aggregate parts t + q --
    the product's layout (64-row quarters of 256-row tiles).  Runs past the last row are
    empty and get no number."""
    assert tile % parts == 0
    return np.arange(N) // (tile // parts)


def two_level_preconditioner(A, A1, S_diag, S1_diag, dc_free, tile=256, parts=4):
    """The product's two-level preconditioner (an extension; the reference has only the
    diagonal one, preconditioner.cpp:59-76), on aggregates of consecutive rows: the
    ``parts`` runs of tile / parts rows of every ``tile``-row tile (``tile_aggregates``;
    ``parts=1``: whole tiles):

        M^-1 r = D^-1 r + W C E1^-1 C W^T r
        D  = diag(A),  W = aggregate indicator (N x na)
        E1 = W^T A1 W  (+ n n^T with n = rows per aggregate when dc_free: the A + 1 1^T
             operator of CG's dcFree solve), A1 = the system matrix at sigma = 1
        C  = diag(c),  c_t = (sum_{i in t} S_ii / sum_{i in t} S1_ii)^(-1/2)

    (S: the stiffness part, harmonic 0).  Returns (prec, Minv) with prec(r) -> z and Minv
    the dense N x N M^-1 when N <= 2000 (else None); prec.einv holds C E1^-1 C and
    prec.agg the aggregate of each row."""
    N = A.shape[0]
    agg = tile_aggregates(N, tile, parts)
    nt = int(agg.max()) + 1
    W = sp.csr_matrix((np.ones(N), (np.arange(N), agg)), shape=(N, nt))
    E1 = (W.T @ A1 @ W).toarray()
    n = np.asarray(W.sum(axis=0)).ravel()
    if dc_free:
        E1 = E1 + np.outer(n, n)
    c = 1.0 / np.sqrt(np.asarray(W.T @ S_diag).ravel() / np.asarray(W.T @ S1_diag).ravel())
    Einv = c[:, None] * np.linalg.inv(E1) * c[None, :]
    dinv = 1.0 / A.diagonal()

    def prec(r):
        return dinv[:, None] * r + W @ (Einv @ (W.T @ r))

    prec.einv = Einv
    prec.agg = agg

    Minv = None
    if N <= 2000:
        Wd = W.toarray()
        Minv = np.diag(dinv) + Wd @ Einv @ Wd.T
    return prec, Minv


def bicgstab_reference(A, b, dinv=None, max_iter=0, eps=FLT_EPSILON, rule="reference", tol=1e-6,
                       project=False):
    """Multi-RHS right-preconditioned BiCGSTAB as ``BiCGSTAB::solve`` runs it
    (src/numeric/bicgstab.cpp:83-193), in float64, x starting at 0.  The reference ignores
    dcFree (:87).  ``project=True`` starts from r = r^ = b - mean(b) and subtracts the mean
    from nu = A y and t = A z (the device's default on the singular harmonic; a no-op in
    exact arithmetic when 1^T A = 0 and 1^T b = 0).
    Per column, with K^-1 = diag(dinv):

        rho = (r, r^);  beta = (rho / rho_old) (alpha / omega)
        p = r + beta (p - omega nu);  y = K^-1 p;  nu = A y;  alpha = rho / (nu, r^)
        s = r - alpha nu;  z = K^-1 s;  t = A z;  w = K^-1 t;  omega = (z, w) / (w, w)
        x = x + alpha y + omega z;  r = s - omega t

    with r^ = r0 = b, rho_old = alpha = omega = 1, p = nu = 0 at the start and every
    division by zero giving 0 (elementwiseDivision, matrix_kernel.cu:344-352).
    ``rule="reference"``: after step s, stop when sqrt(r.r / b.b) < eps for column 0 and
    return s+1 (:179-187).  ``rule="tolerance"``: stop when every column has
    ||r|| <= tol ||b||.  Returns (x, steps).
    """
    N, R = b.shape
    iters = 10 * N if max_iter == 0 else max_iter
    if dinv is None:
        dinv = np.ones(N)
    k = dinv[:, None]

    def op(v):
        y = A @ v
        return y - y.mean(axis=0, keepdims=True) if project else y

    def div(a, c):
        with np.errstate(divide="ignore", invalid="ignore"):
            return np.where(c != 0.0, a / np.where(c != 0.0, c, 1.0), 0.0)

    x = np.zeros((N, R))
    r = b - b.mean(axis=0, keepdims=True) if project else b.copy()
    rh = r.copy()
    rho = np.ones(R)
    alpha = np.ones(R)
    omega = np.ones(R)
    nu = np.zeros((N, R))
    p = np.zeros((N, R))
    ref = np.sum(b * b, axis=0)
    for step in range(iters):
        rho_old = rho
        rho = np.sum(r * rh, axis=0)
        beta = div(rho, rho_old) * div(alpha, omega)
        p = r + beta * (p - omega * nu)
        y = k * p
        nu = op(y)
        alpha = div(rho, np.sum(nu * rh, axis=0))
        s_ = r - alpha * nu
        z = k * s_
        t = op(z)
        w = k * t
        omega = div(np.sum(z * w, axis=0), np.sum(w * w, axis=0))
        x = x + alpha * y + omega * z
        r = s_ - omega * t
        rr = np.sum(r * r, axis=0)
        if rule == "reference":
            with np.errstate(divide="ignore", invalid="ignore"):
                crit = abs(math.sqrt(rr[0] / ref[0])) if ref[0] != 0 else float("nan")
            if not (crit >= eps):
                return x, step + 1
        else:
            ok = np.where(ref > 0, rr <= (tol * tol) * ref, True)
            if np.all(ok):
                return x, step + 1
    return x, iters


def solve_direct(A, F, dc_free):
    """Exact FP64 solve of (A [+ 1 1^T]) X = F via a sparse LU.

    With dcFree, A is the singular Neumann matrix (A 1 = 0).  Then 1^T x = 1^T F / N and
    A x = F - (1^T F / N) 1, a consistent system: it is solved with the last node
    grounded (row and column removed) and the result shifted to the required nodal sum.
    """
    N = A.shape[0]
    cplx = np.iscomplexobj(F) or sp.csr_matrix(A).dtype.kind == "c"
    F = np.asarray(F, dtype=np.complex128 if cplx else np.float64)
    if not dc_free:
        lu = spla.splu(sp.csc_matrix(A))
        return lu.solve(F)
    mean = F.sum(axis=0, keepdims=True) / N
    Fc = F - mean
    A = sp.csr_matrix(A)
    lu = spla.splu(sp.csc_matrix(A[:N - 1, :N - 1]))
    X = np.zeros_like(F)
    X[:N - 1] = lu.solve(Fc[:N - 1])
    # shift by a constant c so that 1^T x = 1^T F / N = mean:  sum + N c = mean
    X += (mean - X.sum(axis=0, keepdims=True)) / N
    return X


# ---------------------------------------------------------------------------
# 2.5D forward model and Jacobian (src/models/eit.cpp:139-245, equation_kernel.cu:197-245)
# ---------------------------------------------------------------------------

class EITOracle:
    """FP64 restatement of ``models::EIT`` with Open (current) sources, or with Fixed
    (voltage) sources when ``voltage`` (``current`` then holds the applied voltages).

    Parameters mirror ``EIT(mesh, sources, referenceValue, ..., components, height,
    portHeight)`` (src/models/eit.cpp:29-95).
    """

    def __init__(self, nodes, elements, boundary, ports, drive, meas, current,
                 reference_value=1.0, components=1, height=1.0, port_height=1.0,
                 quirks=True, logarithmic=False, contact_impedance=0.0, voltage=False):
        self.nodes = np.asarray(nodes, np.float64)
        self.elements = np.asarray(elements, np.int64)
        self.boundary = np.asarray(boundary, np.int64)
        self.ports = np.asarray(ports, np.int64)
        self.N = self.nodes.shape[0]
        self.E, self.k = self.elements.shape
        self.drive = np.asarray(drive, np.float64)
        self.meas = np.asarray(meas, np.float64)
        self.D, self.M = self.drive.shape[1], self.meas.shape[1]
        self.Pi = pattern_matrix(self.drive, self.meas, current)
        self.ref = float(reference_value)
        self.C = int(components)
        self.h = float(height)
        self.hp = float(port_height)
        self.quirks = quirks
        self.log = logarithmic
        self.Ke, self.Me = all_element_integrals(self.nodes, self.elements)
        self.Mp = ceil16(self.M)
        self.Dp = ceil16(self.D)
        self.z = float(contact_impedance)
        self.L = self.ports.shape[0]
        self.voltage = bool(voltage)
        if self.voltage and self.z > 0.0:
            raise ValueError("voltage sources use the gap model")
        if self.z > 0.0:
            # complete electrode model: unknowns [phi (N); V (L)]; current enters at the
            # electrode rows and the voltages are the electrode potentials.  Every harmonic
            # carries its own CEM system with the same electrode blocks (harmonic-wise CEM).
            self.Zc, self.Cc, self.Dc = cem_blocks(self.nodes, self.boundary, self.ports, self.z)
            self.B = np.vstack([np.zeros((self.N, self.L)), np.eye(self.L)])
            self.W = np.hstack([np.zeros((self.M, self.N)), self.meas.T])
        else:
            self.B = excitation_matrix(self.nodes, self.boundary, self.ports, quirks)
            if self.voltage:
                # applyMixedBoundaryConditionKernel (eit_kernel.cu:34-54), called in the
                # constructor before the attachment matrix (eit.cpp:89-94): the excitation
                # is clipped to 0/1 and its nonzero rows become identity rows
                self.B = (np.abs(self.B) >= 1e-9).astype(np.float64)
                self.dirichlet = np.flatnonzero(self.B.any(axis=1))
            self.W = self.meas.T @ self.B.T          # electrodesAttachmentMatrix (eit.cpp:93-94)

    def alpha_beta(self, c):
        """2.5D constants (eit.cpp:154-157)."""
        alpha = (2.0 * c * math.pi / self.h) ** 2
        beta = (1.0 / self.h) if c == 0 else \
            2.0 * math.sin(c * math.pi * self.hp / self.h) / (c * math.pi * self.hp)
        return alpha, beta

    def system(self, gamma, c):
        sig = mat_param(np.asarray(gamma, np.float64), self.ref, self.log)
        S = assemble(self.elements, self.Ke, self.N, sig)
        R = assemble(self.elements, self.Me, self.N, sig)
        alpha, _ = self.alpha_beta(c)
        A = S + alpha * R
        if self.z > 0.0:
            A = sp.bmat([[A + self.Zc, sp.csr_matrix(-self.Cc)],
                         [sp.csr_matrix(-self.Cc.T), sp.diags(self.Dc)]])
        return A.tocsr()

    def mixed_system(self, A):
        """The system matrix with the Dirichlet rows replaced by identity rows, as the
        kernel writes them (values of the row: 1 on the diagonal, 0 elsewhere)."""
        A = sp.lil_matrix(A)
        for i in self.dirichlet:
            A.rows[i] = [i]
            A.data[i] = [1.0]
        return A.tocsr()

    def fields(self, gamma, method="direct", **cg):
        """Phi_c (N x (D+M)) for every harmonic; also returns CG step counts."""
        out, steps = [], 0
        for c in range(self.C):
            A = self.system(gamma, c)
            _, beta = self.alpha_beta(c)
            if self.voltage:
                # eit.cpp:159-177: no beta, no dcFree; the row-replaced system solved directly
                if method != "direct":
                    raise ValueError("voltage sources: direct solves only")
                X = spla.spsolve(self.mixed_system(A).tocsc(), self.B @ self.Pi)
                out.append(np.asarray(X).reshape(self.N, -1))
                continue
            F = beta * (self.B @ self.Pi)
            dc = (c == 0)
            if method == "direct":
                X = solve_direct(A, F, dc)
            else:
                dinv = 1.0 / A.diagonal()
                X, s = pcg_reference(A, F, dinv, dc, **cg)
                steps += s
            out.append(X)
        return out, steps

    def voltages(self, phis, gamma=None):
        """U = sum_c W Phi_c[:, :D] (M x D) (eit.cpp:192-198, 239-244).  Voltage sources
        (eit.cpp:181-190): U = sum_c W (height A_c Phi_c)[:, :D], the electrode currents,
        with A_c the system matrix without the Dirichlet rows (``gamma`` needed)."""
        U = np.zeros((self.M, self.D))
        for c, X in enumerate(phis):
            if self.voltage:
                U += self.W @ (self.h * (self.system(gamma, c) @ X[:, :self.D]))
            else:
                U += self.W @ X[:, :self.D]
        return U

    def solve(self, gamma, method="direct", **cg):
        phis, steps = self.fields(gamma, method, **cg)
        return self.voltages(phis, gamma), phis, steps

    def jacobian(self, gamma, phis, quirks=None):
        """J[(d * ceil16(M) + m), e] (rows ceil16(D)*ceil16(M), padding rows 0).

        Reference (quirks): J = -h * sigma'_e * sum_c Phi_c,d^T K_e Phi_c,m
        (calcJacobianKernel, equation_kernel.cu:197-245; eit.cpp:180-182, 202-205).
        Exact: J = -sigma'_e * sum_c (1/beta_c) Phi_c,d^T (K_e + alpha_c M_e) Phi_c,m.
        """
        quirks = self.quirks if quirks is None else quirks
        if self.voltage and not quirks:
            raise ValueError("voltage sources: the reference Jacobian only")
        ds = mat_param_deriv(np.asarray(gamma, np.float64), self.ref, self.log)
        J = np.zeros((self.Dp, self.Mp, self.E))
        el = self.elements
        for c, X in enumerate(phis):
            alpha, beta = self.alpha_beta(c)
            Kc = self.Ke if quirks else self.Ke + alpha * self.Me
            w = self.h if quirks else 1.0 / beta
            Pd = X[el][:, :, :self.D]            # [E, k, D]
            Pm = X[el][:, :, self.D:]            # [E, k, M]
            T = np.einsum("eij,ejm->eim", Kc, Pm)
            J[:self.D, :self.M, :] += w * np.einsum("eid,eim->dme", Pd, T)
        # eit.cpp:202-205: the sign flip is for current sources only
        J *= (1.0 if self.voltage else -1.0) * ds[None, None, :]
        return J.reshape(self.Dp * self.Mp, self.E)

    def flatten_voltages(self, U):
        """U (M x D) column-major with ld ceil16(M), padded to m rows (inverse.cpp:176-193)."""
        out = np.zeros((self.Dp, self.Mp))
        out[:self.D, :self.M] = U.T
        return out.reshape(-1)


class CoupledCEMOracle(EITOracle):
    """2.5D complete electrode model in which each electrode carries ONE potential V_l for
    all harmonics (an extension; the reference has the gap model only).  The cylinder of
    height h carries u(x, y, z) = sum_c u_c(x, y) cos(k_c z), k_c = 2 pi c / h; electrode l
    is the patch e_l x [-h_p/2, h_p/2] with contact impedance z.  Each mode c has its own
    contact term against its Fourier share h beta_c V of the electrode potential, with the
    patch average of cos^2 (int_patch cos^2(k_c z) dz ~ n_c h_p / h, exact for c = 0), so
    all modes see z' = z h / h_p.  The energy

        sum_c n_c [ u_c^T A_c u_c + (1/z') int_e (u_c - h beta_c V)^2 ] - 2 I^T V

    (n_c = int_0^h cos^2(k_c z) dz: h, then h/2; A_c = S + alpha_c R, the reference's
    harmonic matrices, eit.cpp:154-157; beta_c its excitation scale) gives

        n_c (A_c + Zb / z') u_c - (n_c h beta_c / z') Bc V = 0            c = 0 .. C-1
        sum_c (n_c / z') [ (h beta_c)^2 |e| V - h beta_c Bc^T u_c ] = I

    Zb[i, j] = int_e phi_i phi_j, Bc[i, l] = int_{e_l} phi_i, |e|_l = |e_l|: symmetric and
    positive definite for C >= 2 (positive semidefinite with the constant null mode
    u_0 = 1, V = 1 for C = 1, where it is the harmonic-wise model with current I / h).
    The whole block system is solved at once here; the product solves it by a Schur
    complement on V.  Voltages: U = Mpat^T V.  Jacobian (exact): dU_md / dgamma_e =
    -sigma'_e sum_c n_c u_c^(m)T (K_e + alpha_c M_e) u_c^(d), u_c^(x) the mode-c field of
    electrode currents x (reciprocity of the symmetric block system)."""

    def __init__(self, nodes, elements, boundary, ports, drive, meas, current, reference_value=1.0,
                 components=1, height=1.0, port_height=1.0, logarithmic=False, contact_impedance=0.01):
        super().__init__(nodes, elements, boundary, ports, drive, meas, current, reference_value, components,
                         height, port_height, quirks=False, logarithmic=logarithmic, contact_impedance=0.0)
        self.z = float(contact_impedance)
        if not self.z > 0.0:
            raise ValueError("the coupled complete electrode model needs contact_impedance > 0")
        self.zp = self.z * self.h / self.hp
        # the blocks for unit contact impedance: Zb, Bc, |e|
        Zb, Bc, el = cem_blocks(self.nodes, self.boundary, self.ports, 1.0)
        self.Zb, self.Bc, self.elen = Zb.tocsr(), Bc, el
        self.n_c = np.array([self.h if c == 0 else 0.5 * self.h for c in range(self.C)])

    def system(self, gamma, c):
        """The mode-c block n_c (A_c + Zb / z') (not the coupled system)."""
        sig = mat_param(np.asarray(gamma, np.float64), self.ref, self.log)
        alpha, _ = self.alpha_beta(c)
        A = assemble(self.elements, self.Ke, self.N, sig) + alpha * assemble(self.elements, self.Me, self.N, sig)
        return (self.n_c[c] * (A + self.Zb / self.zp)).tocsr()

    def coupled_system(self, gamma):
        N, L, C = self.N, self.L, self.C
        blocks = [[None] * (C + 1) for _ in range(C + 1)]
        for c in range(C):
            _, beta = self.alpha_beta(c)
            blocks[c][c] = self.system(gamma, c)
            cpl = sp.csr_matrix(-(self.n_c[c] * self.h * beta / self.zp) * self.Bc)
            blocks[c][C] = cpl
            blocks[C][c] = cpl.T
        vv = sum(self.n_c[c] * (self.h * self.alpha_beta(c)[1]) ** 2 for c in range(C))
        blocks[C][C] = sp.diags((vv / self.zp) * self.elen)
        return sp.bmat(blocks).tocsr()

    def fields(self, gamma, method="direct", **cg):
        """Mode fields u_c for the currents Pi (N x (D+M) each) and the electrode
        potentials V (L x (D+M))."""
        if method != "direct":
            raise ValueError("coupled CEM: direct solves only")
        N, L, C = self.N, self.L, self.C
        H = self.coupled_system(gamma)
        rhs = np.zeros((C * N + L, self.Pi.shape[1]))
        rhs[C * N:] = self.Pi
        if C == 1:
            # the constant mode: currents summing to zero are consistent; the solution
            # orthogonal to the null vector is taken
            nv = np.zeros(C * N + L)
            nv[:N] = 1.0
            nv[C * N:] = 1.0
            nv /= np.linalg.norm(nv)
            rhs = rhs - np.outer(nv, nv @ rhs)
            H = H + sp.csr_matrix(np.outer(nv, nv))
        X = spla.spsolve(H.tocsc(), rhs)
        X = np.asarray(X).reshape(C * N + L, -1)
        phis = [X[c * N:(c + 1) * N] for c in range(C)]
        self.V = X[C * N:]
        return phis, 0

    def voltages(self, phis, gamma=None):
        return self.meas.T @ self.V[:, :self.D]

    def solve(self, gamma, method="direct", **cg):
        phis, steps = self.fields(gamma, method, **cg)
        return self.voltages(phis), phis, steps

    def jacobian(self, gamma, phis, quirks=None):
        ds = mat_param_deriv(np.asarray(gamma, np.float64), self.ref, self.log)
        J = np.zeros((self.Dp, self.Mp, self.E))
        el = self.elements
        for c, X in enumerate(phis):
            alpha, _ = self.alpha_beta(c)
            Kc = self.Ke + alpha * self.Me
            Pd = X[el][:, :, :self.D]
            Pm = X[el][:, :, self.D:]
            T = np.einsum("eij,ejm->eim", Kc, Pm)
            J[:self.D, :self.M, :] += self.n_c[c] * np.einsum("eid,eim->dme", Pd, T)
        J *= -ds[None, None, :]
        return J.reshape(self.Dp * self.Mp, self.E)


# ---------------------------------------------------------------------------
# Tikhonov Gauss-Newton (src/solver/inverse.cpp:60-238, src/solver/solver.cpp:142-225)
# ---------------------------------------------------------------------------

class ComplexEITOracle(EITOracle):
    """``models::EIT`` with complex conductivity (``Equation<thrust::complex<float>, ...>``,
    src/models/eit.cpp:289-304): gamma, the system matrices, fields, voltages and the
    Jacobian are complex; the excitation and the 2.5D constants stay real.  The system
    A_c = sum_e sigma_e (K_e + alpha_c M_e) is complex symmetric (no conjugation anywhere
    but in the Krylov solvers' dot products, ``pcg_reference``), and the Jacobian keeps the
    real formula with complex fields (calcJacobianKernel, equation_kernel.cu:197-245).
    Open (current) sources and the gap model only."""

    def __init__(self, *args, **kw):
        if kw.get("contact_impedance", 0.0) or kw.get("voltage", False):
            raise ValueError("complex conductivity: current sources and the gap model only")
        super().__init__(*args, **kw)

    def _sigma(self, gamma):
        g = np.asarray(gamma, np.complex128)
        return self.ref * np.exp(np.log(10.0) * g / 10.0) if self.log else self.ref * g

    def _dsigma(self, gamma):
        g = np.asarray(gamma, np.complex128)
        if self.log:
            return (np.log(10.0) / 10.0) * self.ref * np.exp(np.log(10.0) * g / 10.0)
        return self.ref * np.ones_like(g)

    def system(self, gamma, c):
        sig = self._sigma(gamma)
        alpha, _ = self.alpha_beta(c)
        return (assemble(self.elements, self.Ke, self.N, sig) + alpha * assemble(self.elements, self.Me, self.N, sig)).tocsr()

    def fields(self, gamma, method="direct", **cg):
        out, steps = [], 0
        for c in range(self.C):
            A = self.system(gamma, c)
            _, beta = self.alpha_beta(c)
            F = (beta * (self.B @ self.Pi)).astype(np.complex128)
            if method == "direct":
                out.append(solve_direct(A, F, c == 0))
            else:
                X, st = pcg_reference(A, F, 1.0 / A.diagonal(), c == 0, **cg)
                out.append(X)
                steps += st
        return out, steps

    def voltages(self, phis, gamma=None):
        U = np.zeros((self.M, self.D), np.complex128)
        for X in phis:
            U += self.W @ X[:, :self.D]
        return U

    def jacobian(self, gamma, phis, quirks=None):
        quirks = self.quirks if quirks is None else quirks
        ds = self._dsigma(gamma)
        J = np.zeros((self.Dp, self.Mp, self.E), np.complex128)
        el = self.elements
        for c, X in enumerate(phis):
            alpha, beta = self.alpha_beta(c)
            Kc = self.Ke if quirks else self.Ke + alpha * self.Me
            w = self.h if quirks else 1.0 / beta
            Pd = X[el][:, :, :self.D]
            T = np.einsum("eij,ejm->eim", Kc, X[el][:, :, self.D:])
            J[:self.D, :self.M, :] += w * np.einsum("eid,eim->dme", Pd, T)
        J *= -ds[None, None, :]
        return J.reshape(self.Dp * self.Mp, self.E)

    def flatten_voltages(self, U):
        out = np.zeros((self.Dp, self.Mp), np.complex128)
        out[:self.D, :self.M] = U.T
        return out.reshape(-1)


def gn_step_exact(J, r, lam_eff2):
    """delta = (J^T J + lam I)^-1 J^T r, computed as J^T (J J^T + lam I)^-1 r
    (Cholesky of the m x m data-space matrix, FP64).  Equal to the E x E normal
    equations of Inverse::solve by the push-through identity.  Complex J: the reference
    forms J^T J and J^T r with plain transposes (CUBLAS_OP_T, inverse.cpp:148-152,
    196-204), so the matrix is complex symmetric, not Hermitian: solved by LU."""
    G = J @ J.T + lam_eff2 * np.eye(J.shape[0])
    if np.iscomplexobj(G):
        return J.T @ np.linalg.solve(G, r)
    cf = sla.cho_factor(G)
    return J.T @ sla.cho_solve(cf, r)


def gn_step_normal(J, r, lam_eff2):
    """delta from the E x E normal equations (plain transposes) by dense Cholesky, or LU
    when J is complex (small E only)."""
    A = J.T @ J + lam_eff2 * np.eye(J.shape[1])
    if np.iscomplexobj(A):
        return np.linalg.solve(A, J.T @ r)
    return sla.cho_solve(sla.cho_factor(A), J.T @ r)


def noser_diag(J0):
    """RegularizationType::noser: R = diag(J0^T J0) (inverse.cpp:93-95)."""
    return np.einsum("ie,ie->e", J0, J0)


def gn_step_regularized(J, r, lam_eff2, R_diag):
    """delta from the E x E system (J^T J + lam R) delta = J^T r with a diagonal R by dense
    Cholesky (Inverse::updateJacobian's system matrix, inverse.cpp:60-76; small E only)."""
    A = J.T @ J + lam_eff2 * np.diag(R_diag)
    return sla.cho_solve(sla.cho_factor(A), J.T @ r)


def laplacian_matrix(nodes, elements):
    """RegularizationType::laplacian (Inverse::calcRegularizationMatrix, inverse.cpp:96-117):
    L[edge, element] = sign(edge) * |edge| over each element's three vertex edges, the
    edge's sign flipping after every use; R = L^T L (E x E, dense)."""
    el = np.asarray(elements)[:, :3]
    E = el.shape[0]
    index = {}
    elem_edges = np.empty((E, 3), np.int64)
    for e in range(E):
        for i in range(3):
            a, b = int(el[e, i]), int(el[e, (i + 1) % 3])
            key = (min(a, b), max(a, b))
            if key not in index:
                index[key] = len(index)
            elem_edges[e, i] = index[key]
    ends = np.array(sorted(index, key=index.get))
    length = np.hypot(*(nodes[ends[:, 0]] - nodes[ends[:, 1]]).T)
    L = np.zeros((len(index), E))
    signs = np.ones(len(index))
    for e in range(E):
        for i in range(3):
            k = elem_edges[e, i]
            L[k, e] = signs[k] * length[k]
            signs[k] *= -1.0
    return L.T @ L


def gn_step_general(J, r, lam_eff2, R):
    """delta from (J^T J + lam R) delta = J^T r for a dense SPD R (small E only)."""
    A = J.T @ J + lam_eff2 * R
    return sla.cho_solve(sla.cho_factor(A), J.T @ r)


def gn_step_cg_reference(J, r, lam_eff2, max_iter=0, eps=FLT_EPSILON):
    """Inverse::solve as written: A = J^T J + (sigma_ref lambda)^2 I, b = J^T r,
    unpreconditioned CG from 0 with the column-0 stop rule (inverse.cpp:212-238)."""
    A = J.T @ J + lam_eff2 * np.eye(J.shape[1])
    b = (J.T @ r)[:, None]
    x, steps = pcg_reference(A, b, None, False, max_iter, eps)
    return x[:, 0], steps


def reconstruct(oracle: EITOracle, U_meas, gamma0, lam, steps=3, quirks_jacobian=None):
    """``steps`` x Solver::solveAbsolute from gamma0; lam is regularizationFactor and the
    diagonal term is (referenceValue * lam)^2 (inverse.cpp:73-75)."""
    g = np.asarray(gamma0, np.complex128 if isinstance(oracle, ComplexEITOracle) else np.float64).copy()
    lam2 = (oracle.ref * lam) ** 2
    um = oracle.flatten_voltages(U_meas)
    for _ in range(steps):
        U, phis, _ = oracle.solve(g)
        J = oracle.jacobian(g, phis, quirks_jacobian)
        r = um - oracle.flatten_voltages(U)
        g = g + gn_step_exact(J, r, lam2)
    return g
