"""ctypes binding of ``include/mpflow_amd.h`` (the C ABI of ``libmpflow_amd.so``)."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpflow_amd.so")

MPF_OK, MPF_EINVAL, MPF_ENOMEM, MPF_EHIP, MPF_ENOCONV, MPF_EFULL = range(6)
MPF_STOP_REFERENCE, MPF_STOP_TOLERANCE = 0, 1
MPF_PREC_FP32, MPF_PREC_MIXED = 0, 1
MPF_KRYLOV_CG, MPF_KRYLOV_BICGSTAB, MPF_KRYLOV_BLOCK_CG = 0, 1, 2
MPF_PRECOND_JACOBI, MPF_PRECOND_TWO_LEVEL = 0, 1
MPF_REG_IDENTITY, MPF_REG_NOSER, MPF_REG_LAPLACIAN = 0, 1, 2
MPF_SOURCE_CURRENT, MPF_SOURCE_VOLTAGE = 0, 1

EXPORTED = [
    "mpf_version", "mpf_last_error", "mpf_ctx_create", "mpf_ctx_destroy", "mpf_ctx_synchronize",
    "mpf_ctx_stream", "mpf_solve_opts_default", "mpf_ports_circular", "mpf_model_create",
    "mpf_model_destroy", "mpf_model_dims", "mpf_eit_solve", "mpf_gn_step", "mpf_ell_spmv",
    "mpf_gram_bf16", "mpf_gemv_t", "mpf_spd_solve", "mpf_ctx_profile", "mpf_ctx_profile_read", "mpf_ctx_autotune_read", "mpf_ctx_stream_probe",
    "mpf_element_integrals", "mpf_batch_create", "mpf_batch_destroy", "mpf_batch_solve",
    "mpf_batch_gn_step", "mpf_mlp_create", "mpf_mlp_destroy", "mpf_mlp_fwd", "mpf_mlp_bwd",
    "mpf_mlp_adam", "mpf_mlp_param_count", "mpf_mlp_params", "mpf_pcg_solve", "mpf_bicgstab_solve", "mpf_block_pcg_solve",
    "mpf_fem_assemble", "mpf_eit_measure", "mpf_jacobian", "mpf_gn_set_regularization", "mpf_diff_prepare",
    "mpf_diff_solve", "mpf_normal_solve", "mpf_set_sensitivity_mask", "mpf_sparse_multiply",
    "mpf_normal_solve_noalloc",
]


class ModelDesc(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int32), ("n_elements", C.c_int32), ("nodes_per_element", C.c_int32),
        ("nodes", C.POINTER(C.c_double)), ("elements", C.POINTER(C.c_int32)),
        ("n_boundary", C.c_int32), ("nodes_per_edge", C.c_int32), ("boundary", C.POINTER(C.c_int32)),
        ("n_ports", C.c_int32), ("port_max_edges", C.c_int32), ("port_edges", C.POINTER(C.c_int32)),
        ("n_drive", C.c_int32), ("n_meas", C.c_int32),
        ("drive_pattern", C.POINTER(C.c_double)), ("meas_pattern", C.POINTER(C.c_double)),
        ("drive_current", C.POINTER(C.c_double)),
        ("reference_value", C.c_double), ("components", C.c_int32), ("height", C.c_double),
        ("port_height", C.c_double), ("logarithmic", C.c_int32), ("quirks", C.c_int32),
        ("jacobian_exact", C.c_int32), ("ell_max_width", C.c_int32), ("frames", C.c_int32),
        ("keep_node_order", C.c_int32), ("contact_impedance", C.c_double), ("source_type", C.c_int32),
        ("cem_coupled", C.c_int32),
        ("complex_gamma", C.c_int32),
    ]


class SolveOpts(C.Structure):
    _fields_ = [
        ("stop_rule", C.c_int32), ("tol", C.c_double), ("max_iter", C.c_int32),
        ("precision", C.c_int32), ("refine_steps", C.c_int32), ("refine_tol", C.c_double),
        ("krylov", C.c_int32), ("check_every", C.c_int32), ("warm_start", C.c_int32),
        ("bicgstab_ignore_dc", C.c_int32), ("graphs", C.c_int32), ("preconditioner", C.c_int32),
        ("complex_hermitian", C.c_int32),
    ]


class SolveStats(C.Structure):
    _fields_ = [
        ("steps", C.c_int64), ("max_steps", C.c_int64), ("not_converged", C.c_int32),
        ("refine_residual", C.c_double), ("seconds", C.c_double),
    ]


_lib = None


def lib() -> C.CDLL:
    """Load ``libmpflow_amd.so``.  Raises if it has not been built: the product path has
    no CPU fallback."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libmpflow_amd.so is not built; run `python -m mpflow_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    L.mpf_version.restype = C.c_char_p
    L.mpf_last_error.restype = C.c_char_p
    L.mpf_ctx_create.argtypes = [C.c_int, vp, C.POINTER(vp)]
    L.mpf_ctx_destroy.argtypes = [vp]
    L.mpf_ctx_synchronize.argtypes = [vp]
    L.mpf_ctx_stream.argtypes = [vp]
    L.mpf_ctx_stream.restype = vp
    L.mpf_solve_opts_default.argtypes = [C.POINTER(SolveOpts)]
    L.mpf_solve_opts_default.restype = None
    L.mpf_ports_circular.argtypes = [i32, vp, i32, i32, vp, i32, i32, vp, i32, dbl, dbl, i32, vp, i32,
                                     C.POINTER(i32)]
    L.mpf_model_create.argtypes = [vp, C.POINTER(ModelDesc), C.POINTER(vp)]
    L.mpf_model_destroy.argtypes = [vp]
    L.mpf_model_dims.argtypes = [vp, C.POINTER(i64)]
    L.mpf_eit_solve.argtypes = [vp, vp, vp, vp, vp, C.POINTER(SolveOpts), C.POINTER(SolveStats)]
    L.mpf_gn_step.argtypes = [vp, vp, vp, dbl, i32, C.POINTER(SolveOpts), vp, C.POINTER(SolveStats)]
    L.mpf_ell_spmv.argtypes = [vp, i32, i32, vp, vp, vp, i32, vp, vp, vp, i32, i32]
    u32 = C.c_uint32
    L.mpf_sparse_multiply.argtypes = [vp, vp, vp, vp, u32, u32, u32, u32, vp]
    L.mpf_gram_bf16.argtypes = [vp, i32, i32, i32, vp, i32, vp]
    L.mpf_gemv_t.argtypes = [vp, i32, i32, vp, i32, vp, vp]
    L.mpf_spd_solve.argtypes = [vp, i32, vp, dbl, vp, vp]
    L.mpf_ctx_profile.argtypes = [vp, i32]
    L.mpf_ctx_profile_read.argtypes = [vp, C.POINTER(dbl)]
    L.mpf_ctx_stream_probe.argtypes = [vp, i32, i32, vp, vp, C.c_int64, i32, i32, C.POINTER(dbl)]
    L.mpf_ctx_autotune_read.argtypes = [vp, i32, C.POINTER(C.c_int64), C.POINTER(i32), i32, C.POINTER(i32)]
    L.mpf_element_integrals.argtypes = [i32, vp, vp, vp]
    L.mpf_batch_create.argtypes = [vp, vp, i32, C.POINTER(vp)]
    L.mpf_batch_destroy.argtypes = [vp]
    L.mpf_batch_solve.argtypes = [vp, vp, vp, vp, vp, C.POINTER(SolveOpts), C.POINTER(SolveStats)]
    L.mpf_batch_gn_step.argtypes = [vp, vp, vp, dbl, i32, C.POINTER(SolveOpts), vp, C.POINTER(SolveStats)]
    L.mpf_mlp_create.argtypes = [vp, i32, vp, C.c_uint64, C.POINTER(vp)]
    L.mpf_mlp_destroy.argtypes = [vp]
    L.mpf_mlp_fwd.argtypes = [vp, i32, vp, vp]
    L.mpf_mlp_bwd.argtypes = [vp, vp, C.POINTER(dbl)]
    L.mpf_mlp_adam.argtypes = [vp, C.c_float, C.c_float, C.c_float, C.c_float]
    L.mpf_mlp_param_count.argtypes = [vp]
    L.mpf_mlp_param_count.restype = i64
    L.mpf_mlp_params.argtypes = [vp, i32, i32, vp]
    for f in (L.mpf_pcg_solve, L.mpf_bicgstab_solve, L.mpf_block_pcg_solve):
        f.argtypes = [vp, i32, i32, vp, vp, i32, vp, i32, vp, i32, i32, C.POINTER(SolveOpts), C.POINTER(SolveStats)]
    L.mpf_fem_assemble.argtypes = [vp, vp, dbl, vp, vp]
    L.mpf_eit_measure.argtypes = [vp, vp, i32, i32, vp]
    L.mpf_jacobian.argtypes = [vp, vp, i32, vp, i32, vp]
    L.mpf_gn_set_regularization.argtypes = [vp, i32, vp, C.POINTER(SolveOpts), C.POINTER(SolveStats)]
    L.mpf_diff_prepare.argtypes = [vp, vp, dbl, C.POINTER(SolveOpts), C.POINTER(SolveStats)]
    L.mpf_diff_solve.argtypes = [vp, vp, i32, vp]
    L.mpf_set_sensitivity_mask.argtypes = [vp, vp]
    L.mpf_normal_solve.argtypes = [vp, i32, i32, vp, i32, dbl, vp, i32, vp]
    L.mpf_normal_solve_noalloc.argtypes = [vp, i32, i32, vp, i32, dbl, vp, i32, vp]
    _lib = L
    return L


class MpfError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("mpflow_amd error %d: %s" % (code, msg))
        self.code = code


def check(rc: int, allow=(MPF_OK,)):
    if rc not in allow:
        raise MpfError(rc, lib().mpf_last_error().decode())
    return rc
