"""Python host mirror of mpFlow's EIT API over the C ABI.

Names and argument meaning follow the reference classes:
  ``EIT``    -> models::EIT            (src/models/eit.cpp:29-213)
  ``Solver`` -> solver::Solver         (src/solver/solver.cpp:28-225)
Device buffers are torch tensors on ``cuda``; torch only provides memory and the
stream, every computation runs in ``libmpflow_amd.so``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import check, lib

__all__ = ["Context", "EIT", "Solver", "Batch", "BatchSolver", "circular_ports", "SolveOptions"]


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class Context:
    """mpf_ctx: device + HIP stream.  With ``use_torch_stream`` the context runs on torch's
    current stream so that torch-allocated buffers are ordered with the library's work."""

    def __init__(self, device: int = 0, use_torch_stream: bool = True):
        import torch
        self.device = device
        stream = None
        if use_torch_stream:
            stream = C.c_void_p(torch.cuda.current_stream(device).cuda_stream)
        h = C.c_void_p()
        check(lib().mpf_ctx_create(device, stream, C.byref(h)))
        self.h = h

    def synchronize(self):
        check(lib().mpf_ctx_synchronize(self.h))

    def close(self):
        if self.h:
            lib().mpf_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def circular_ports(nodes, elements, boundary, count, width, offset=0.0, clockwise=False):
    """Ports::circularBoundary (src/fem/ports.cpp:32-77) via the C ABI; returns a
    count x maxEdges int32 array of boundary-edge rows, -1 = unused."""
    nodes = np.ascontiguousarray(nodes, np.float64)
    elements = np.ascontiguousarray(elements, np.int32)
    boundary = np.ascontiguousarray(boundary, np.int32)
    w = C.c_int32()
    L = lib()
    args = (nodes.shape[0], nodes.ctypes.data, elements.shape[0], elements.shape[1], elements.ctypes.data,
            boundary.shape[0], boundary.shape[1], boundary.ctypes.data, int(count), float(width), float(offset),
            int(bool(clockwise)))
    check(L.mpf_ports_circular(*args, None, 0, C.byref(w)))
    out = np.empty((count, max(w.value, 1)), np.int32)
    check(L.mpf_ports_circular(*args, out.ctypes.data, out.size, C.byref(w)))
    return out


class SolveOptions:
    def __init__(self, stop_rule=capi.MPF_STOP_REFERENCE, tol=1e-6, max_iter=0, precision=capi.MPF_PREC_FP32,
                 refine_steps=2, refine_tol=1e-3, krylov=capi.MPF_KRYLOV_CG, check_every=16, warm_start=0,
                 bicgstab_ignore_dc=0, graphs=0, preconditioner=capi.MPF_PRECOND_JACOBI, complex_hermitian=0):
        self.s = capi.SolveOpts()
        lib().mpf_solve_opts_default(C.byref(self.s))
        self.s.stop_rule, self.s.tol, self.s.max_iter = stop_rule, tol, max_iter
        self.s.precision, self.s.refine_steps, self.s.refine_tol = precision, refine_steps, refine_tol
        self.s.krylov, self.s.check_every = krylov, check_every
        self.s.warm_start = int(warm_start)
        self.s.bicgstab_ignore_dc = int(bicgstab_ignore_dc)
        self.s.graphs = int(graphs)
        self.s.preconditioner = int(preconditioner)
        self.s.complex_hermitian = int(complex_hermitian)

    def copy(self, **kw):
        o = SolveOptions()
        C.memmove(C.byref(o.s), C.byref(self.s), C.sizeof(self.s))
        for k, v in kw.items():
            setattr(o.s, k, v)
        return o


class EIT:
    """models::EIT on one mesh, for ``frames`` frames at once.  ``source`` is "current"
    (Open sources, the default: ``current`` is the injected current per drive column) or
    "voltage" (Fixed sources, eit.cpp:89-91,159-194: ``current`` is then the applied
    voltage per drive column and the results are electrode currents)."""

    def __init__(self, nodes, elements, boundary, ports, drive_pattern, meas_pattern, current,
                 reference_value=1.0, components=1, height=1.0, port_height=1.0, logarithmic=False,
                 quirks=True, jacobian_exact=False, frames=1, ell_max_width=0, ctx: Context | None = None,
                 keep_node_order=False, contact_impedance=0.0, source="current", cem="harmonic",
                 complex_gamma=False):
        import torch
        self.ctx = ctx or Context()
        self.torch = torch
        self._keep = dict(
            nodes=np.ascontiguousarray(nodes, np.float64),
            elements=np.ascontiguousarray(elements, np.int32),
            boundary=np.ascontiguousarray(boundary, np.int32),
            ports=np.ascontiguousarray(ports, np.int32),
            drive=np.ascontiguousarray(drive_pattern, np.float64),
            meas=np.ascontiguousarray(meas_pattern, np.float64),
        )
        k = self._keep
        D = k["drive"].shape[1]
        cur = np.ascontiguousarray(np.broadcast_to(np.asarray(current, np.float64), (D,)))
        k["current"] = cur
        d = capi.ModelDesc()
        d.n_nodes, d.n_elements = k["nodes"].shape[0], k["elements"].shape[0]
        d.nodes_per_element = k["elements"].shape[1]
        d.nodes, d.elements = _ptr(k["nodes"], C.c_double), _ptr(k["elements"], C.c_int32)
        d.n_boundary, d.nodes_per_edge = k["boundary"].shape
        d.boundary = _ptr(k["boundary"], C.c_int32)
        d.n_ports, d.port_max_edges = k["ports"].shape
        d.port_edges = _ptr(k["ports"], C.c_int32)
        d.n_drive, d.n_meas = D, k["meas"].shape[1]
        d.drive_pattern, d.meas_pattern = _ptr(k["drive"], C.c_double), _ptr(k["meas"], C.c_double)
        d.drive_current = _ptr(cur, C.c_double)
        d.reference_value, d.components = float(reference_value), int(components)
        d.height, d.port_height = float(height), float(port_height)
        d.logarithmic, d.quirks, d.jacobian_exact = int(logarithmic), int(quirks), int(jacobian_exact)
        d.ell_max_width, d.frames = int(ell_max_width), int(frames)
        d.keep_node_order = int(bool(keep_node_order))
        d.contact_impedance = float(contact_impedance)
        if cem not in ("harmonic", "coupled"):
            raise ValueError("cem must be 'harmonic' or 'coupled'")
        d.cem_coupled = int(cem == "coupled" and contact_impedance > 0.0)
        self.cem = cem if contact_impedance > 0.0 else None
        if source not in ("current", "voltage"):
            raise ValueError("source must be 'current' or 'voltage'")
        d.source_type = capi.MPF_SOURCE_VOLTAGE if source == "voltage" else capi.MPF_SOURCE_CURRENT
        self.source = source
        # complex conductivity (Equation<thrust::complex<float>>): complex gamma, voltages
        # and Jacobian (torch complex64 / complex128 tensors)
        d.complex_gamma = int(bool(complex_gamma))
        self.complex = bool(complex_gamma)
        h = C.c_void_p()
        check(lib().mpf_model_create(self.ctx.h, C.byref(d), C.byref(h)))
        self.h = h
        dims = (C.c_int64 * 16)()
        check(lib().mpf_model_dims(h, dims))
        (self.N, self.E, self.np, self.width, self.m, self.Mp, self.Dp, self.R, self.Rp, self.frames,
         self.C, self.nnz, self.R_unique, self.wpad) = [int(v) for v in dims[:14]]
        self.D, self.M = D, k["meas"].shape[1]
        self.reference_value = float(reference_value)
        self.logarithmic = bool(logarithmic)
        dev = torch.device("cuda", self.ctx.device)
        self.dtype = torch.complex64 if self.complex else torch.float32
        self.dtype64 = torch.complex128 if self.complex else torch.float64
        self.result = torch.zeros(self.frames, self.m, dtype=self.dtype, device=dev)
        self.result64 = torch.zeros(self.frames, self.m, dtype=self.dtype64, device=dev)
        self.jacobian = None
        self.last_stats = None

    def _dev(self):
        return self.torch.device("cuda", self.ctx.device)

    def solve(self, gamma, options: SolveOptions | None = None, jacobian=False):
        """EIT::solve: voltages (frames x m, column-major M x D with ld ceil16(M)).
        ``gamma``: (frames, E) or (E,) float32 cuda tensor (complex64 for a complex model)."""
        torch = self.torch
        g = gamma.reshape(self.frames, self.E).to(self._dev(), self.dtype).contiguous()
        Jp = None
        if jacobian:
            if self.jacobian is None:
                self.jacobian = torch.empty(self.frames, self.E, self.m, dtype=self.dtype, device=self._dev())
            Jp = C.c_void_p(self.jacobian.data_ptr())
        st = capi.SolveStats()
        opts = (options or SolveOptions()).s
        rc = lib().mpf_eit_solve(self.h, C.c_void_p(g.data_ptr()), C.c_void_p(self.result.data_ptr()),
                                 C.c_void_p(self.result64.data_ptr()), Jp, C.byref(opts), C.byref(st))
        check(rc, allow=(capi.MPF_OK, capi.MPF_ENOCONV))
        self.last_stats = st
        return self.result

    def update(self, gamma, alpha=0.0, column_ids=None, values=None):
        """Equation::update of one frame (mpf_fem_assemble): the system matrix
        S(gamma) + alpha R(gamma) as the SparseMatrix arrays (rows x 32 column ids, values),
        caller's node numbering.  Returns (column_ids uint32-as-int32, values) cuda tensors."""
        torch = self.torch
        g = gamma.reshape(self.E).to(self._dev(), torch.float32).contiguous()
        if values is None:
            values = torch.empty(self.N, 32, dtype=torch.float32, device=self._dev())
        if column_ids is None:
            column_ids = torch.empty(self.N, 32, dtype=torch.int32, device=self._dev())
        check(lib().mpf_fem_assemble(self.h, C.c_void_p(g.data_ptr()), float(alpha),
                                     C.c_void_p(column_ids.data_ptr()), C.c_void_p(values.data_ptr())))
        return column_ids, values

    def apply_measurement_pattern(self, phi, result, additive=False):
        """EIT::applyMeasurementPattern (mpf_eit_measure): result (ceil16(M) x D,
        column-major) (+)= W phi[:, :D]; phi is a (D + M, ld) tensor (column-major rows x
        (D + M) fields, ld >= rows)."""
        check(lib().mpf_eit_measure(self.h, C.c_void_p(phi.data_ptr()), int(phi.shape[-1]), int(bool(additive)),
                                    C.c_void_p(result.data_ptr())))
        return result

    def calc_jacobian(self, phi, gamma, jacobian, additive=False):
        """Equation::calcJacobian (mpf_jacobian) for one harmonic's fields ``phi`` (as in
        ``apply_measurement_pattern``); ``jacobian`` is E x m (column-major m x E)."""
        torch = self.torch
        g = gamma.reshape(self.E).to(self._dev(), torch.float32).contiguous()
        check(lib().mpf_jacobian(self.h, C.c_void_p(phi.data_ptr()), int(phi.shape[-1]), C.c_void_p(g.data_ptr()),
                                 int(bool(additive)), C.c_void_p(jacobian.data_ptr())))
        return jacobian

    def voltages(self, U=None, dtype=None):
        """(frames, M, D) view of the padded result."""
        U = self.result64 if U is None else U
        u = U.reshape(self.frames, self.Dp, self.Mp)[:, :self.D, :self.M].transpose(1, 2)
        return u if dtype is None else u.to(dtype)

    def jacobian_matrix(self):
        """(frames, m, E) Jacobian (stored transposed-contiguous as E x m per frame)."""
        return self.jacobian.transpose(1, 2)

    def set_sensitivity_mask(self, mask=None):
        """The configuration's ``noSensitivity`` mask (Solver::fromConfig,
        src/solver/solver.cpp:107-113): an (m, E) array multiplied elementwise into the
        Jacobians of the noser regularisation and of the differential preparation; None
        clears it."""
        torch = self.torch
        if mask is None:
            check(lib().mpf_set_sensitivity_mask(self.h, None))
            self._mask = None
            return
        t = torch.as_tensor(mask, dtype=torch.float32)
        if tuple(t.shape) != (self.m, self.E):
            raise ValueError("sensitivity mask: shape %s, expected (%d, %d)" % (tuple(t.shape), self.m, self.E))
        self._mask = t.t().contiguous().to(self._dev())          # E x m: column-major m x E
        check(lib().mpf_set_sensitivity_mask(self.h, C.c_void_p(self._mask.data_ptr())))

    def close(self):
        if getattr(self, "h", None):
            lib().mpf_model_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Solver:
    """solver::Solver with identity Tikhonov regularisation; ``solve_absolute`` is one
    Gauss-Newton step per frame (Solver::solveAbsolute, src/solver/solver.cpp:195-225)."""

    def __init__(self, eit: EIT, regularization_factor: float, gn_refine: int = 4,
                 options: SolveOptions | None = None):
        torch = eit.torch
        self.eit = eit
        self.regularization_factor = float(regularization_factor)
        self.gn_refine = int(gn_refine)
        self.options = options or SolveOptions()
        # the reference distribution: gamma = 1 (linear) or 0 (logarithmic), solver.cpp:55-62
        init = 0.0 if eit.logarithmic else 1.0
        self.material_distribution = torch.full((eit.frames, eit.E), init, dtype=eit.dtype, device=eit._dev())
        self.measurement = torch.zeros(eit.frames, eit.m, dtype=eit.dtype64, device=eit._dev())
        self.calculation = torch.zeros(eit.frames, eit.m, dtype=eit.dtype64, device=eit._dev())
        self.last_stats = None

    def set_regularization(self, regularization_type: str = "identity", gamma_ref=None):
        """Inverse::setRegularizationParameter's type (src/solver/inverse.cpp:83-135):
        "identity", "noser" (R = diag(J0^T J0) at ``gamma_ref``, default the reference
        distribution; computed now with one forward solve and kept) or "laplacian"
        (R = L^T L over the elements' shared edges, inverse.cpp:96-117)."""
        types = {"identity": capi.MPF_REG_IDENTITY, "noser": capi.MPF_REG_NOSER,
                 "laplacian": capi.MPF_REG_LAPLACIAN}
        if regularization_type not in types:
            raise ValueError("regularization type must be one of %s" % sorted(types))
        g = None
        if gamma_ref is not None:
            g = gamma_ref.reshape(self.eit.E).to(self.eit._dev(), self.eit.torch.float32).contiguous()
        st = capi.SolveStats()
        check(lib().mpf_gn_set_regularization(self.eit.h, types[regularization_type],
                                              None if g is None else C.c_void_p(g.data_ptr()),
                                              C.byref(self.options.s), C.byref(st)))

    def pre_solve(self, gamma_ref=None):
        """Solver::preSolve (solver.cpp:136-160): forward solve, Jacobian and factored
        normal matrix at the reference distribution (default 1), kept for
        ``solve_differential``."""
        torch = self.eit.torch
        if gamma_ref is None:
            gamma_ref = torch.full((self.eit.E,), 0.0 if self.eit.logarithmic else 1.0, dtype=torch.float32,
                                   device=self.eit._dev())
        self.reference_distribution = gamma_ref.reshape(self.eit.E).to(self.eit._dev(), torch.float32).contiguous()
        st = capi.SolveStats()
        check(lib().mpf_diff_prepare(self.eit.h, C.c_void_p(self.reference_distribution.data_ptr()),
                                     self.regularization_factor, C.byref(self.options.s), C.byref(st)))

    def solve_differential(self):
        """Solver::solveDifferential (solver.cpp:163-188) for every frame's measurement:
        material_distribution = reference + one linearised step, reusing pre_solve's
        Jacobian and factorisation."""
        check(lib().mpf_diff_solve(self.eit.h, C.c_void_p(self.measurement.data_ptr()), self.gn_refine,
                                   C.c_void_p(self.material_distribution.data_ptr())))
        return self.material_distribution

    def solve_absolute(self, warm_start: bool = False):
        """One Gauss-Newton step.  ``warm_start`` starts the forward solves from the fields of
        the previous step (mixed precision only); the first step of a frame should not."""
        st = capi.SolveStats()
        opts = self.options.copy(warm_start=int(bool(warm_start)))
        rc = lib().mpf_gn_step(self.eit.h, C.c_void_p(self.material_distribution.data_ptr()),
                               C.c_void_p(self.measurement.data_ptr()), self.regularization_factor,
                               self.gn_refine, C.byref(opts.s),
                               C.c_void_p(self.calculation.data_ptr()), C.byref(st))
        check(rc, allow=(capi.MPF_OK, capi.MPF_ENOCONV))
        self.last_stats = st
        return self.material_distribution


def _ptrs(tensors):
    arr = (C.c_void_p * len(tensors))()
    for i, t in enumerate(tensors):
        arr[i] = None if t is None else t.data_ptr()
    return arr


class Batch:
    """Several ``EIT`` models (any mesh, electrode and harmonic counts) solved together:
    one set of Krylov launches per iteration for all of them (mpf_batch_*).  The models
    must share one ``Context``; close the batch before its models."""

    def __init__(self, models):
        self.models = list(models)
        ctx = self.models[0].ctx
        arr = (C.c_void_p * len(self.models))(*[m.h.value for m in self.models])
        h = C.c_void_p()
        check(lib().mpf_batch_create(ctx.h, arr, len(self.models), C.byref(h)))
        self.h = h
        self.last_stats = None

    def solve(self, gammas, options: SolveOptions | None = None, jacobian=False):
        """EIT::solve of every model; ``gammas[i]`` as for ``EIT.solve``.  Results land in
        each model's ``result`` / ``result64`` / ``jacobian``."""
        torch = self.models[0].torch
        gs, Js = [], []
        for m, g in zip(self.models, gammas):
            gs.append(g.reshape(m.frames, m.E).to(m._dev(), torch.float32).contiguous())
            if jacobian and m.jacobian is None:
                m.jacobian = torch.empty(m.frames, m.E, m.m, dtype=torch.float32, device=m._dev())
            Js.append(m.jacobian if jacobian else None)
        st = capi.SolveStats()
        opts = (options or SolveOptions()).s
        rc = lib().mpf_batch_solve(self.h, _ptrs(gs), _ptrs([m.result for m in self.models]),
                                   _ptrs([m.result64 for m in self.models]), _ptrs(Js), C.byref(opts),
                                   C.byref(st))
        check(rc, allow=(capi.MPF_OK, capi.MPF_ENOCONV))
        self.last_stats = st
        return [m.result for m in self.models]

    def close(self):
        if getattr(self, "h", None):
            lib().mpf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchSolver:
    """``Solver.solve_absolute`` of several solvers (one per model of a ``Batch``) with the
    forward solves batched."""

    def __init__(self, batch: Batch, solvers):
        self.batch = batch
        self.solvers = list(solvers)
        assert [s.eit for s in self.solvers] == batch.models
        self.last_stats = None

    def solve_absolute(self, warm_start: bool = False):
        s0 = self.solvers[0]
        st = capi.SolveStats()
        opts = s0.options.copy(warm_start=int(bool(warm_start)))
        rc = lib().mpf_batch_gn_step(self.batch.h, _ptrs([s.material_distribution for s in self.solvers]),
                                     _ptrs([s.measurement for s in self.solvers]), s0.regularization_factor,
                                     s0.gn_refine, C.byref(opts.s),
                                     _ptrs([s.calculation for s in self.solvers]), C.byref(st))
        check(rc, allow=(capi.MPF_OK, capi.MPF_ENOCONV))
        self.last_stats = st
        return [s.material_distribution for s in self.solvers]
