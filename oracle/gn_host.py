"""CPU baseline of whole reconstruction frames (oracle/gn_host.cpp): the device path's
algorithm -- FP32 two-level (or Jacobi) PCG with FP64 scalars, warm starts and FP64
residual refinement over all harmonics, the exact 2.5D Jacobian, and the data-space
Tikhonov step (FP32 Gram, FP64 Cholesky, FP64-residual passes) -- one frame per OpenMP
thread.

TEST INFRASTRUCTURE ONLY (bench.py's cpu_baseline leg and tests/).  Everything it needs
is built here from the oracle's own FP64 set-up (element integrals, excitation,
measurement map); nothing comes from the device path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import scipy.sparse as sp

from . import build_baseline
from .eit_oracle import CoupledCEMOracle, EITOracle, all_element_integrals, rcb_order, tile_aggregates

_I, _D, _F = C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_float)


class _Problem(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("E", C.c_int), ("k", C.c_int), ("elements", _I), ("Ke", _D), ("Me", _D),
        ("nnz", C.c_int), ("rowptr", _I), ("col", _I), ("emap", _I), ("S0", _D),
        ("C", C.c_int), ("alpha", _D), ("beta", _D), ("jw", _D),
        ("Ru", C.c_int), ("F", _D),
        ("D", C.c_int), ("M", C.c_int), ("Dp", C.c_int), ("Mp", C.c_int), ("colmap", _I), ("colscale", _D),
        ("w_off", _I), ("w_node", _I), ("w_val", _D),
        ("ref", C.c_double), ("logarithmic", C.c_int),
        ("na", C.c_int), ("astart", _I), ("cinv", _F), ("d1", _D),
        ("tol", C.c_double), ("refine_tol", C.c_double), ("refine", C.c_int), ("gn_refine", C.c_int),
        ("lam2", C.c_double),
        ("cem_coupled", C.c_int), ("L", C.c_int), ("pRu", C.c_int),
        ("bc_off", _I), ("bc_row", _I), ("bc_val", _D), ("elen", _D), ("pi", _D), ("prep", _I),
        ("coef", _D), ("tsc", _D), ("zp", C.c_double),
    ]


def _lib():
    path = build_baseline.GN_LIB
    if not os.path.exists(path):
        build_baseline.build()
    L = C.CDLL(path)
    L.gnh_reconstruct.restype = C.c_double
    L.gnh_reconstruct.argtypes = [C.POINTER(_Problem), C.c_int, _D, C.c_int, _F, C.c_int, _D, _F]
    L.gnh_threads.restype = C.c_int
    return L


def _unique_columns(F):
    """Columns of F that are scalar multiples of an earlier column share one solve:
    returns (kept column indices, colmap, colscale) with F[:, j] = colscale[j] *
    F[:, keep[colmap[j]]]."""
    keep, colmap, colscale = [], [], []
    for j in range(F.shape[1]):
        f = F[:, j]
        hit = False
        for u, i in enumerate(keep):
            g = F[:, i]
            piv = int(np.argmax(np.abs(g)))
            if g[piv] == 0.0:
                continue
            s = f[piv] / g[piv]
            if np.max(np.abs(f - s * g)) <= 1e-12 * max(np.max(np.abs(f)), 1e-300):
                colmap.append(u)
                colscale.append(s)
                hit = True
                break
        if not hit:
            colmap.append(len(keep))
            colscale.append(1.0)
            keep.append(j)
    return keep, np.asarray(colmap, np.int32), np.asarray(colscale, np.float64)


class GnHost:
    """Host reconstruction of a gap-model or complete-electrode-model EIT problem (the
    mesh renumbered by recursive coordinate bisection so that the two-level
    preconditioner's aggregates are 64-row ranges, as on the device)."""

    def __init__(self, nodes, elements, boundary, ports, drive, meas, current, reference_value=1.0,
                 components=1, height=1.0, port_height=1.0, contact_impedance=0.0, two_level=True,
                 tol=1e-4, refine=1, refine_tol=1e-2, gn_refine=4, lam=1e-3, cem="harmonic"):
        nodes = np.asarray(nodes, np.float64)
        order = rcb_order(nodes, 256)
        perm = np.empty_like(order)
        perm[order] = np.arange(order.size)
        nodes = nodes[order]
        elements = perm[np.asarray(elements)]
        boundary = perm[np.asarray(boundary)]
        coupled = contact_impedance > 0.0 and cem == "coupled"
        if coupled:
            o = CoupledCEMOracle(nodes, elements, boundary, ports, drive, meas, current, reference_value, components,
                                 height, port_height, contact_impedance=contact_impedance)
        else:
            o = EITOracle(nodes, elements, boundary, ports, drive, meas, current, reference_value, components,
                          height, port_height, quirks=False, contact_impedance=contact_impedance)
        self.o = o
        self.coupled = coupled
        # harmonic-wise CEM: the electrode potentials follow the nodes
        N = o.N + (o.L if o.z > 0.0 and not coupled else 0)
        k = o.k
        # coupled: the mode blocks without their factor n_c, A_c + Zb / z'
        sys1 = (lambda c: o.system(np.ones(o.E), c) / o.n_c[c]) if coupled else (lambda c: o.system(np.ones(o.E), c))
        A1 = sys1(0).tocsr()
        A1.sort_indices()
        self.N, self.E, self.k = N, o.E, k
        Ke, Me = o.Ke, o.Me
        # CSR slot of every element entry
        el = o.elements
        rows = np.repeat(el, k, axis=1).reshape(-1)
        cols = np.tile(el, (1, k)).reshape(-1)
        indptr, indices = A1.indptr, A1.indices
        emap = np.empty(rows.size, np.int64)
        for t in range(rows.size):
            r = rows[t]
            lo, hi = indptr[r], indptr[r + 1]
            emap[t] = lo + np.searchsorted(indices[lo:hi], cols[t])
        # sigma-independent entries (complete electrode model blocks) added to S
        self.S0 = None
        if o.z > 0.0:
            S1 = (A1 - self._mesh_part(o, np.ones(o.E), N)).tocsr()
            S1.eliminate_zeros()
            S1.sort_indices()
            S0 = np.zeros(indices.size)
            for r in range(N):
                lo, hi = indptr[r], indptr[r + 1]
                for q in range(S1.indptr[r], S1.indptr[r + 1]):
                    S0[lo + np.searchsorted(indices[lo:hi], S1.indices[q])] += S1.data[q]
            self.S0 = np.ascontiguousarray(S0)
        self.rowptr = np.ascontiguousarray(indptr, np.int32)
        self.col = np.ascontiguousarray(indices, np.int32)
        self.emap = np.ascontiguousarray(emap, np.int32)
        self.elements = np.ascontiguousarray(el, np.int32)
        self.Ke = np.ascontiguousarray(Ke.reshape(o.E, k * k))
        self.Me = np.ascontiguousarray(Me.reshape(o.E, k * k))
        C_ = o.C
        ab = [o.alpha_beta(c) for c in range(C_)]
        self.alpha = np.ascontiguousarray([a for a, _ in ab], np.float64)
        self.beta = np.ascontiguousarray([b for _, b in ab], np.float64)
        self.jw = np.ascontiguousarray(1.0 / self.beta)
        if coupled:
            # unit electrode sources, unscaled; Jacobian weights n_c; the pattern's unique
            # columns index the combined fields
            F = o.Bc
            self.F = np.ascontiguousarray(F)
            self.jw = np.ascontiguousarray(o.n_c, np.float64)
            keep, colmap, colscale = _unique_columns(o.Pi)
            self.beta_rhs = np.ones(C_)
        else:
            F = o.B @ o.Pi
            keep, colmap, colscale = _unique_columns(F)
            self.F = np.ascontiguousarray(F[:, keep])
            self.beta_rhs = self.beta
        self.colmap, self.colscale = colmap, colscale
        self.keep = keep
        Wd = np.asarray(o.W)
        Wc = sp.csr_matrix(Wd)
        self.w_off = np.ascontiguousarray(Wc.indptr, np.int32)
        self.w_node = np.ascontiguousarray(Wc.indices, np.int32)
        self.w_val = np.ascontiguousarray(Wc.data, np.float64)
        # two-level: aggregates = 64-row quarters of 256-row tiles, coarse inverses at sigma = 1
        self.na = 0
        self.astart = np.zeros(1, np.int32)
        self.cinv = np.zeros(1, np.float32)
        self.d1 = np.zeros(1, np.float64)
        if two_level:
            agg = tile_aggregates(N, 256)
            na = int(agg.max()) + 1
            W = sp.csr_matrix((np.ones(N), (np.arange(N), agg)), shape=(N, na))
            n_a = np.asarray(W.sum(axis=0)).ravel()
            cinv = []
            for c in range(C_):
                E1 = (W.T @ sys1(c) @ W).toarray()
                if c == 0 and not coupled:
                    E1 = E1 + np.outer(n_a, n_a)
                cinv.append(np.linalg.inv(E1))
            S1 = sys1(0).diagonal()
            self.na = na
            self.astart = np.ascontiguousarray(np.r_[0, np.cumsum(np.bincount(agg, minlength=na))], np.int32)
            self.cinv = np.ascontiguousarray(np.stack(cinv), np.float32)
            self.d1 = np.ascontiguousarray(np.asarray(W.T @ S1).ravel(), np.float64)
        self.p = _Problem()
        p = self.p
        p.N, p.E, p.k = N, o.E, k
        p.elements = self.elements.ctypes.data_as(_I)
        p.Ke, p.Me = self.Ke.ctypes.data_as(_D), self.Me.ctypes.data_as(_D)
        p.nnz = int(self.col.size)
        p.rowptr, p.col, p.emap = (a.ctypes.data_as(_I) for a in (self.rowptr, self.col, self.emap))
        p.S0 = self.S0.ctypes.data_as(_D) if self.S0 is not None else None
        p.C = C_
        self.beta_rhs = np.ascontiguousarray(self.beta_rhs, np.float64)
        p.alpha, p.beta, p.jw = (a.ctypes.data_as(_D) for a in (self.alpha, self.beta_rhs, self.jw))
        p.Ru = self.F.shape[1]
        p.F = self.F.ctypes.data_as(_D)
        p.D, p.M, p.Dp, p.Mp = o.D, o.M, o.Dp, o.Mp
        p.colmap, p.colscale = self.colmap.ctypes.data_as(_I), self.colscale.ctypes.data_as(_D)
        p.w_off, p.w_node, p.w_val = (self.w_off.ctypes.data_as(_I), self.w_node.ctypes.data_as(_I),
                                      self.w_val.ctypes.data_as(_D))
        p.ref, p.logarithmic = o.ref, int(o.log)
        p.na = self.na
        p.astart, p.cinv, p.d1 = self.astart.ctypes.data_as(_I), self.cinv.ctypes.data_as(_F), self.d1.ctypes.data_as(_D)
        p.tol, p.refine_tol, p.refine, p.gn_refine = tol, refine_tol, refine, gn_refine
        p.lam2 = (o.ref * lam) ** 2
        self.m = o.Dp * o.Mp
        p.cem_coupled = int(coupled)
        if coupled:
            L = o.L
            Bc = sp.csc_matrix(o.Bc)
            self.bc_off = np.ascontiguousarray(Bc.indptr, np.int32)
            self.bc_row = np.ascontiguousarray(Bc.indices, np.int32)
            self.bc_val = np.ascontiguousarray(Bc.data, np.float64)
            self.elen = np.ascontiguousarray(o.elen, np.float64)
            self.pi = np.ascontiguousarray(o.Pi, np.float64)
            self.prep = np.ascontiguousarray(keep, np.int32)
            hb = np.array([o.h * b for b in self.beta])
            self.coef = np.ascontiguousarray(o.n_c * hb * hb / o.zp)
            self.tsc = np.ascontiguousarray(hb / o.zp)
            p.L, p.pRu = L, len(keep)
            p.bc_off, p.bc_row = self.bc_off.ctypes.data_as(_I), self.bc_row.ctypes.data_as(_I)
            p.bc_val, p.elen = self.bc_val.ctypes.data_as(_D), self.elen.ctypes.data_as(_D)
            p.pi, p.prep = self.pi.ctypes.data_as(_D), self.prep.ctypes.data_as(_I)
            p.coef, p.tsc = self.coef.ctypes.data_as(_D), self.tsc.ctypes.data_as(_D)
            p.zp = o.zp

    @staticmethod
    def _mesh_part(o, gamma, N):
        """S(gamma) of the mesh alone, padded to the N rows of the system."""
        from .eit_oracle import assemble, mat_param
        S = assemble(o.elements, o.Ke, o.N, mat_param(gamma, o.ref, o.log)).tocsr()
        S.resize((N, N))
        return S

    def reconstruct(self, Umeas, steps=3, threads=0, outputs=False):
        """Umeas: frames x m flattened measured voltages (EIT::result layout).  Returns
        (gamma frames x E, wall seconds), and with ``outputs`` also the voltages
        (frames x m) and Jacobian (frames x m x E) of each frame's last step."""
        Um = np.ascontiguousarray(np.atleast_2d(Umeas), np.float64)
        frames = Um.shape[0]
        g = np.ones((frames, self.E), np.float32)
        Uo = np.zeros((frames, self.m)) if outputs else None
        Jo = np.zeros((frames, self.E, self.m), np.float32) if outputs else None
        t = _lib().gnh_reconstruct(C.byref(self.p), frames, Um.ctypes.data_as(_D), int(steps),
                                   g.ctypes.data_as(_F), int(threads),
                                   Uo.ctypes.data_as(_D) if outputs else None,
                                   Jo.ctypes.data_as(_F) if outputs else None)
        if outputs:
            return g.astype(np.float64), t, Uo, np.transpose(Jo, (0, 2, 1)).astype(np.float64)
        return g.astype(np.float64), t


def threads():
    return _lib().gnh_threads()
