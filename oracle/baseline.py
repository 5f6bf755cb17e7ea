"""CPU baseline leg: the same ELL PCG (Jacobi or two-level; FP32 vectors, FP64 scalars)
timed on the host.

TEST INFRASTRUCTURE ONLY (bench.py's cpu_baseline leg and tests/).  The ELL matrix is
built here from the oracle's own FP64 assembly (eit_oracle.assemble), not from the
device path.
"""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np

from . import build_baseline


def _lib():
    if not os.path.exists(build_baseline.LIB):
        build_baseline.build()
    L = C.CDLL(build_baseline.LIB)
    L.cgh_pcg.restype = C.c_int
    L.cgh_pcg.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                          C.c_int, C.c_double, C.c_int, C.c_void_p, C.c_int, C.c_void_p]
    L.cgh_threads.restype = C.c_int
    L.cgh_set_threads.argtypes = [C.c_int]
    return L


def csr_to_ell(A):
    """Row-major ELL (ids, vals) of a scipy CSR matrix; padding slots id = row, 0."""
    A = A.tocsr()
    n = A.shape[0]
    nnz_row = np.diff(A.indptr)
    width = int(nnz_row.max())
    ids = np.repeat(np.arange(n, dtype=np.int32)[:, None], width, axis=1)
    vals = np.zeros((n, width), np.float32)
    for i in range(n):
        s, e = A.indptr[i], A.indptr[i + 1]
        ids[i, :e - s] = A.indices[s:e]
        vals[i, :e - s] = A.data[s:e]
    return np.ascontiguousarray(ids), np.ascontiguousarray(vals)


def pcg(A, B, dc, tol=1e-5, max_iter=100000, threads=0, einv=None, tile=256, agg=None):
    """Solve (A + dc 1 1^T) X = B with the C++ baseline; returns (X, steps, seconds).
    ``einv`` (na x na): the two-level preconditioner's scaled coarse inverse
    (``eit_oracle.two_level_preconditioner``) over the aggregates ``agg`` of the rows
    (default: the 4 runs of tile / 4 rows of each ``tile``-row tile,
    ``eit_oracle.tile_aggregates``); None:
    Jacobi alone."""
    from .eit_oracle import tile_aggregates
    L = _lib()
    if threads:
        L.cgh_set_threads(int(threads))
    ids, vals = csr_to_ell(A)
    n, width = ids.shape
    B = np.ascontiguousarray(B, np.float32)
    R = B.shape[1]
    X = np.zeros_like(B)
    E = None if einv is None else np.ascontiguousarray(einv, np.float32)
    na, starts = 0, None
    if E is not None:
        agg = tile_aggregates(n, tile) if agg is None else np.asarray(agg)
        assert np.all(np.diff(agg) >= 0), "aggregates are ranges of consecutive rows"
        na = E.shape[0]
        starts = np.ascontiguousarray(np.r_[0, np.cumsum(np.bincount(agg, minlength=na))], np.int32)
    t0 = time.perf_counter()
    steps = L.cgh_pcg(n, width, ids.ctypes.data, vals.ctypes.data, R, B.ctypes.data, X.ctypes.data,
                      int(dc), float(tol), int(max_iter), None if E is None else E.ctypes.data, int(na),
                      None if starts is None else starts.ctypes.data)
    return X, steps, time.perf_counter() - t0


def threads():
    return _lib().cgh_threads()
