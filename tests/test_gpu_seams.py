"""The low-level seams on caller buffers in the reference's layouts, against FP64 dense
products and the oracle's Krylov restatement:
  * mpf_sparse_multiply  = SparseMatrix::multiply (sparse_matrix_kernel.cu:157-196)
  * mpf_ell_spmv         = the product-layout multi-harmonic S/R SpMV
  * mpf_pcg_solve        from a caller's start X0 (conjugate_gradient.cpp:80-88)
  * the reference stop rule's step count against the float32 restatement
  * the error paths: bad column ids, a zero right-hand side, the iteration cap, and
    "sparse format full" at model creation."""
import ctypes as C

import numpy as np
import pytest

import oracle.eit_oracle as eo
from mpflow_amd.phantom import frame_seed, inclusion_phantom
from tests.problems import make_problem, oracle_of

pytestmark = pytest.mark.gpu

INVALID = 0xFFFFFFFF


def _ref_layout(A, density_pad=0, shuffle_seed=None):
    """CSR -> SparseMatrix arrays (column ids uint32 [n*32], values float32 [n*32],
    density), entries of each row in CSR order or shuffled, `density_pad` extra empty
    slots in the middle of every row's used range."""
    A = A.tocsr()
    n = A.shape[0]
    rng = np.random.default_rng(shuffle_seed) if shuffle_seed is not None else None
    cid = np.full((n, 32), INVALID, np.uint32)
    val = np.zeros((n, 32), np.float32)
    density = 0
    for r in range(n):
        cols = A.indices[A.indptr[r]:A.indptr[r + 1]].astype(np.uint32)
        vals = A.data[A.indptr[r]:A.indptr[r + 1]].astype(np.float32)
        if rng is not None:
            o = rng.permutation(len(cols))
            cols, vals = cols[o], vals[o]
        k = len(cols)
        h = k // 2
        cid[r, :h] = cols[:h]
        val[r, :h] = vals[:h]
        cid[r, h + density_pad:k + density_pad] = cols[h:]
        val[r, h + density_pad:k + density_pad] = vals[h:]
        density = max(density, k + density_pad)
    return cid, val, density


def _dev(a, dtype=None):
    import torch
    t = torch.tensor(np.ascontiguousarray(a))
    return t.to("cuda") if dtype is None else t.to(device="cuda", dtype=dtype)


@pytest.mark.parametrize("rows,cols,R,pad", [(300, 300, 3, 0), (1000, 1000, 20, 2), (517, 517, 64, 0),
                                             (700, 700, 100, 1)])
def test_sparse_multiply_reference_layout(gpu_ctx, rows, cols, R, pad):
    import scipy.sparse as sp
    import torch
    from mpflow_amd import capi
    from mpflow_amd.capi import lib
    print("case: mpf_sparse_multiply rows=%d R=%d pad=%d" % (rows, R, pad), flush=True)
    rng = np.random.default_rng(rows + R)
    A = sp.random(rows, cols, density=min(1.0, 9.0 / cols), random_state=rows, format="csr")
    A = (A + sp.eye(rows, cols)).tocsr()
    A.data = rng.standard_normal(A.data.size)
    cid, val, density = _ref_layout(A, density_pad=pad, shuffle_seed=R)
    assert density <= 32
    X = rng.standard_normal((cols, R)).astype(np.float32)
    Xd = _dev(X.ravel(order="F"))
    Yd = torch.full((rows * R,), float("nan"), dtype=torch.float32, device="cuda")
    cd, vd = _dev(cid.ravel().view(np.int32)), _dev(val.ravel())
    capi.check(lib().mpf_sparse_multiply(gpu_ctx.h, C.c_void_p(vd.data_ptr()), C.c_void_p(cd.data_ptr()),
                                         C.c_void_p(Xd.data_ptr()), rows, cols, R, density,
                                         C.c_void_p(Yd.data_ptr())))
    Y = Yd.cpu().numpy().reshape(R, rows).T.astype(np.float64)
    Aref = sp.csr_matrix(A.astype(np.float32).astype(np.float64))
    Yref = Aref @ X.astype(np.float64)
    # FP32 sums of <= 12 products: a few ulps of the row's absolute sum
    bound = 1e-6 * (abs(Aref) @ np.abs(X.astype(np.float64)))
    assert np.all(np.abs(Y - Yref) <= bound + 1e-30)
    # a column id past matrix_rows is refused
    bad = cid.copy()
    bad[5, 0] = cols
    bd = _dev(bad.ravel().view(np.int32))
    rc = lib().mpf_sparse_multiply(gpu_ctx.h, C.c_void_p(vd.data_ptr()), C.c_void_p(bd.data_ptr()),
                                   C.c_void_p(Xd.data_ptr()), rows, cols, R, density, C.c_void_p(Yd.data_ptr()))
    assert rc == capi.MPF_EINVAL


def test_ell_spmv_product_layout(gpu_ctx):
    """Y[c] = (S + alpha_c R) X[c] on the slot-major layout, 3 harmonics, 32 lanes."""
    import torch
    from mpflow_amd import capi
    from mpflow_amd.capi import lib
    print("case: mpf_ell_spmv", flush=True)
    p = make_problem(16, 16, 0.1)
    o = oracle_of(p)
    g = inclusion_phantom(p["nodes"], p["elements"], frame_seed(1, 3))
    sig = eo.mat_param(g, 1.0)
    S = eo.assemble(o.elements, o.Ke, o.N, sig).tocsr()
    Rm = eo.assemble(o.elements, o.Me, o.N, sig).tocsr()
    n = o.N
    np_ = (n + 255) // 256 * 256
    width = 16
    ids = np.tile(np.arange(np_, dtype=np.int32), (width, 1))
    sv = np.zeros((width, np_), np.float32)
    rv = np.zeros((width, np_), np.float32)
    for r in range(n):
        cols = S.indices[S.indptr[r]:S.indptr[r + 1]]
        assert len(cols) <= width
        ids[:len(cols), r] = cols
        sv[:len(cols), r] = S.data[S.indptr[r]:S.indptr[r + 1]]
        rv[:len(cols), r] = [Rm[r, c] for c in cols]
    alpha = np.array([0.0, 247.0, 987.0], np.float32)
    C_, ld, R = 3, 32, 32
    rng = np.random.default_rng(5)
    X = np.zeros((C_, np_, ld), np.float32)
    X[:, :n, :R] = rng.standard_normal((C_, n, R))
    Xd, Yd = _dev(X), torch.zeros(C_ * np_ * ld, dtype=torch.float32, device="cuda")
    idd, sd, rd = _dev(ids), _dev(sv), _dev(rv)
    al = (C.c_float * 3)(*alpha)
    capi.check(lib().mpf_ell_spmv(gpu_ctx.h, np_, width, C.c_void_p(idd.data_ptr()), C.c_void_p(sd.data_ptr()),
                                  C.c_void_p(rd.data_ptr()), C_, al, C.c_void_p(Xd.data_ptr()),
                                  C.c_void_p(Yd.data_ptr()), R, ld))
    Y = Yd.cpu().numpy().reshape(C_, np_, ld)
    for c in range(C_):
        A32 = (sp_csr(sv, ids, n, np_) + float(alpha[c]) * sp_csr(rv, ids, n, np_))
        Yref = A32 @ X[c, :n, :R].astype(np.float64)
        assert np.max(np.abs(Y[c, :n, :R] - Yref)) <= 2e-6 * np.max(np.abs(Yref))


def sp_csr(vals, ids, n, np_):
    import scipy.sparse as sp
    W = vals.shape[0]
    rows = np.repeat(np.arange(np_)[None, :], W, axis=0)
    M = sp.csr_matrix((vals.ravel().astype(np.float64), (rows.ravel(), ids.ravel())), shape=(np_, np_))
    return M[:n, :n]


@pytest.fixture(scope="module")
def sysm():
    p = make_problem(16, 16, 0.1, C=2, height=0.4, port_height=0.1)
    o = oracle_of(p)
    g = inclusion_phantom(p["nodes"], p["elements"], frame_seed(1, 8))
    return o, g


def _pcg(gpu_ctx, entry, A, B, dc_free, X0=None, **opt):
    import torch
    from mpflow_amd import capi
    from mpflow_amd.capi import lib
    from mpflow_amd.eit import SolveOptions
    n, R = B.shape
    ld = (n + 15) // 16 * 16
    cid, val, density = _ref_layout(A)
    dc, dv = _dev(cid.ravel().view(np.int32)), _dev(val.ravel())
    Bd = torch.zeros(R, ld, dtype=torch.float32, device="cuda")
    Bd[:, :n] = torch.tensor(B.T, dtype=torch.float32)
    Xd = torch.zeros(R, ld, dtype=torch.float32, device="cuda")
    if X0 is not None:
        Xd[:, :n] = torch.tensor(X0.T, dtype=torch.float32)
    st = capi.SolveStats()
    o = SolveOptions(**opt)
    rc = getattr(lib(), entry)(gpu_ctx.h, n, density, C.c_void_p(dc.data_ptr()), C.c_void_p(dv.data_ptr()), R,
                               C.c_void_p(Bd.data_ptr()), ld, C.c_void_p(Xd.data_ptr()), ld, int(dc_free),
                               C.byref(o.s), C.byref(st))
    return rc, Xd[:, :n].cpu().numpy().T.astype(np.float64), st


@pytest.mark.parametrize("precision", [0, 1])
def test_pcg_solve_from_caller_start(gpu_ctx, sysm, precision):
    """X0 != 0: SPD harmonic -> the direct solution; dcFree harmonic -> the direct
    solution + 2 (1^T X0) / N (the reference's start residual b - A x0 + 1 1^T x0)."""
    from mpflow_amd import capi
    o, g = sysm
    B = o.B @ o.Pi
    rng = np.random.default_rng(9)
    X0 = (rng.standard_normal(B.shape) * 1e-3).astype(np.float32).astype(np.float64)
    for c in (1, 0):
        print("case: mpf_pcg_solve start X0 harmonic %d precision %d" % (c, precision), flush=True)
        A = o.system(g, c)
        rc, X, st = _pcg(gpu_ctx, "mpf_pcg_solve", A, B, c == 0, X0, stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-6,
                         precision=precision, refine_steps=2, refine_tol=1e-3)
        assert rc == capi.MPF_OK and st.not_converged == 0
        Xr = eo.solve_direct(A, B, c == 0)
        if c == 0:
            Xr = Xr + 2.0 * X0.sum(axis=0, keepdims=True) / o.N
        err = float(np.max(np.abs(X - Xr)) / np.max(np.abs(Xr)))
        # the restatement from the same start agrees as well
        Xo, _ = eo.pcg_reference(A, B, 1.0 / A.diagonal(), c == 0, rule="tolerance", tol=1e-10, x0=X0)
        print("harmonic %d: rel err vs closed form %.2e, vs restatement %.2e, steps %d"
              % (c, err, float(np.max(np.abs(X - Xo)) / np.max(np.abs(Xo))), st.steps))
        assert err < (2e-5 if precision == 0 else 1e-6)


@pytest.mark.parametrize("contrast", [1.0, 100.0])
def test_reference_stop_rule_step_count(gpu_ctx, sysm, contrast):
    """Reference stop rule (sqrt(|r.z| / b.b) of column 0 below FLT_EPSILON) on the dcFree
    harmonic: the device's step count equals the float32 restatement's within 4 steps.
    Variants of the rule move the restatement's count by more: every column instead of
    column 0 +14/+16, r.r instead of r.z +5/+11, no square root -89/-54 (contrast 1 / 100);
    the same rule on the last column instead of column 0 (+1/+4) is not told apart."""
    from mpflow_amd import capi
    o, g = sysm
    gg = np.where(g > 1.5, contrast, 1.0) if contrast != 1.0 else g
    A = o.system(gg, 0)
    B = o.B @ o.Pi
    _, s32 = eo.pcg_reference(A, B, 1.0 / A.diagonal(), True, fp32=True)
    _, s64 = eo.pcg_reference(A, B, 1.0 / A.diagonal(), True)
    print("case: reference stop rule contrast %g" % contrast, flush=True)
    rc, X, st = _pcg(gpu_ctx, "mpf_pcg_solve", A, B, True, stop_rule=capi.MPF_STOP_REFERENCE)
    print("device steps %d, float32 restatement %d, float64 restatement %d" % (st.steps, s32, s64))
    assert rc == capi.MPF_OK
    assert abs(st.steps - s32) <= 4
    Xd = eo.solve_direct(A, B, True)
    assert np.max(np.abs(X - Xd)) < 1e-5 * np.max(np.abs(Xd))


def test_pcg_error_paths(gpu_ctx, sysm):
    """A zero right-hand side stops after one step with X = 0 (sqrt(0/0) is not >= eps,
    conjugate_gradient.cpp:145-151); an iteration cap returns MPF_ENOCONV; a column id
    past n is refused."""
    from mpflow_amd import capi
    o, g = sysm
    A = o.system(g, 1)
    B = o.B @ o.Pi
    print("case: pcg zero rhs", flush=True)
    rc, X, st = _pcg(gpu_ctx, "mpf_pcg_solve", A, np.zeros_like(B[:, :4]), False, stop_rule=capi.MPF_STOP_REFERENCE)
    assert rc == capi.MPF_OK and st.steps == 1 and np.all(X == 0.0)
    print("case: pcg iteration cap", flush=True)
    rc, X, st = _pcg(gpu_ctx, "mpf_pcg_solve", A, B, False, stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-12,
                     max_iter=3)
    assert rc == capi.MPF_ENOCONV and st.not_converged > 0 and st.max_steps <= 3 + 16
    print("case: pcg bad column id", flush=True)
    import scipy.sparse as sp
    A2 = sp.lil_matrix(A)
    A2[3, 2] = 0.5
    A2 = A2.tocsr()
    cid, val, density = _ref_layout(A2)
    cid[3, np.argmax(cid[3] == 2)] = o.N + 7
    import torch
    from mpflow_amd.capi import lib
    from mpflow_amd.eit import SolveOptions
    dc, dv = _dev(cid.ravel().view(np.int32)), _dev(val.ravel())
    ld = (o.N + 15) // 16 * 16
    Bd = torch.zeros(4, ld, dtype=torch.float32, device="cuda")
    Xd = torch.zeros(4, ld, dtype=torch.float32, device="cuda")
    st = capi.SolveStats()
    so = SolveOptions()
    rc = lib().mpf_pcg_solve(gpu_ctx.h, o.N, density, C.c_void_p(dc.data_ptr()), C.c_void_p(dv.data_ptr()), 4,
                             C.c_void_p(Bd.data_ptr()), ld, C.c_void_p(Xd.data_ptr()), ld, 0, C.byref(so.s),
                             C.byref(st))
    assert rc == capi.MPF_EINVAL


def test_model_sparse_format_full(gpu_ctx):
    """A node shared by 40 triangles needs 41 entries in its row: model creation reports
    "sparse format full" (MPF_EFULL, sparse_matrix.cpp:417-419) with the 32-slot cap, and
    succeeds with ell_max_width = 64."""
    from mpflow_amd import capi
    from mpflow_amd.eit import EIT
    from mpflow_amd.mesh import boundary_edges
    k = 40
    t = 2 * np.pi * np.arange(k) / k
    nodes = np.vstack([[0.0, 0.0], np.stack([np.cos(t), np.sin(t)], axis=1)])
    el = np.array([[0, 1 + i, 1 + (i + 1) % k] for i in range(k)], np.int32)
    bd = boundary_edges(el)
    L = 4
    ports = np.array([[i * 10, i * 10 + 1] for i in range(L)], np.int32)
    pat = np.zeros((L, L))
    for d in range(L):
        pat[d, d], pat[(d + 1) % L, d] = 1.0, -1.0
    print("case: sparse format full", flush=True)
    with pytest.raises(capi.MpfError) as e:
        EIT(nodes, el, bd, ports, pat, pat, 1e-3, ctx=gpu_ctx)
    assert e.value.code == capi.MPF_EFULL
    m = EIT(nodes, el, bd, ports, pat, pat, 1e-3, ctx=gpu_ctx, ell_max_width=64)
    assert m.N == k + 1


def test_autotune_read_reports_choices(gpu_ctx):
    """mpf_ctx_autotune_read: after a model solve the context holds one SpMV choice per
    (lane count, width, row bucket) with a valid kernel and order; the query with no
    room reports the count."""
    import ctypes as C
    import torch
    from mpflow_amd import capi
    from mpflow_amd.eit import Context, EIT, SolveOptions
    from tests.problems import make_problem
    p = make_problem(8, 16, 0.1)
    ctx = Context()
    eit = EIT(p["nodes"], p["elements"], p["boundary"], p["ports"], p["drive"], p["meas"], p["current"], ctx=ctx)
    eit.solve(torch.ones(eit.E), SolveOptions(stop_rule=capi.MPF_STOP_TOLERANCE, tol=1e-6))
    L = capi.lib()
    n = C.c_int32()
    capi.check(L.mpf_ctx_autotune_read(ctx.h, 0, None, None, 0, C.byref(n)))
    assert n.value >= 1
    keys = (C.c_int64 * n.value)()
    vals = (C.c_int32 * n.value)()
    capi.check(L.mpf_ctx_autotune_read(ctx.h, 0, keys, vals, n.value, C.byref(n)))
    for k, v in zip(keys, vals):
        assert (k >> 40) in (4, 5, 6) and (v & 15) in (0, 1, 2) and (v >> 4) in (0, 1)
    assert L.mpf_ctx_autotune_read(ctx.h, 2, None, None, 0, C.byref(n)) == capi.MPF_EINVAL
    eit.close()
    ctx.close()
