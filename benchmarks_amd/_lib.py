"""ctypes binding of the C-ABI in include/mfx.h (benchmarks_amd/libmfx.so).

The library is the product: there is no fallback.  If libmfx.so is missing
every entry point raises, naming the build step.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmfx.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "mfx.h")

MFX_OK = 0
STATUS_NAMES = {0: "MFX_OK", 1: "MFX_ERR_INVALID_ARG", 2: "MFX_ERR_OOM", 3: "MFX_ERR_HIP",
                4: "MFX_ERR_NOT_CONVERGED", 5: "MFX_ERR_UNSUPPORTED", 6: "MFX_ERR_BREAKDOWN"}
INVALID_DOF = 0xFFFFFFFF
QUAD_GAUSS, QUAD_GLL = 0, 1
GEOM_STORED, GEOM_ON_THE_FLY = 0, 1


class MfxError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class BoxDesc(C.Structure):
    _fields_ = [("degree", C.c_int), ("n_q_points_1d", C.c_int), ("quadrature", C.c_int),
                ("geometry", C.c_int), ("n_cells", C.c_uint32 * 3), ("lo", C.c_double * 3),
                ("hi", C.c_double * 3), ("deform", C.c_double), ("with_fp32", C.c_int)]


class OpDesc(C.Structure):
    _fields_ = [("degree", C.c_int), ("n_q_points_1d", C.c_int), ("quadrature", C.c_int),
                ("geometry", C.c_int), ("n_cells", C.c_uint32), ("n_dofs", C.c_uint64),
                ("map", C.c_void_p), ("cell_vertices", C.c_void_p),
                ("constrained", C.c_void_p), ("n_constrained", C.c_uint64),
                ("with_fp32", C.c_int)]


class OpInfo(C.Structure):
    _fields_ = [("degree", C.c_int), ("n_q_points_1d", C.c_int), ("quadrature", C.c_int),
                ("geometry", C.c_int), ("n_cells", C.c_uint32), ("n_dofs", C.c_uint64),
                ("n_constrained", C.c_uint64), ("map", C.c_void_p), ("G", C.c_void_p),
                ("B", C.c_void_p), ("D", C.c_void_p), ("cell_vertices", C.c_void_p),
                ("constrained", C.c_void_p), ("xq", C.c_void_p), ("wq", C.c_void_p),
                ("n_patches", C.c_uint32), ("patch_cells", C.c_int * 3), ("patch_map", C.c_void_p)]


class CgResult(C.Structure):
    _fields_ = [("steps", C.c_int), ("converged", C.c_int), ("initial_value", C.c_double),
                ("last_value", C.c_double), ("seconds", C.c_double)]


class MgDesc(C.Structure):
    _fields_ = [("fine", BoxDesc), ("smoother_degree", C.c_int), ("smoothing_range", C.c_double),
                ("eig_cg_n_iterations", C.c_int), ("eig_safety", C.c_double),
                ("coarse_max_dofs", C.c_uint32), ("fine_smoother_n_q_points_1d", C.c_int)]


class MgLevelInfo(C.Structure):
    _fields_ = [("degree", C.c_int), ("n_q_points_1d", C.c_int), ("n_cells", C.c_uint32 * 3),
                ("n_dofs", C.c_uint64), ("lambda_max", C.c_double), ("lambda_min", C.c_double),
                ("op", C.c_void_p)]


class BatchInstance(C.Structure):
    _fields_ = [("degree", C.c_int), ("n_cells", C.c_uint32 * 3), ("lo", C.c_double * 3),
                ("hi", C.c_double * 3), ("deform", C.c_double)]


class BatchTuning(C.Structure):
    _fields_ = [("degree", C.c_int), ("n_cells", C.c_uint32), ("n_variants", C.c_int),
                ("variant", C.c_int), ("default_variant", C.c_int),
                ("cells_per_workgroup", C.c_int * 8), ("prefetch", C.c_int * 8),
                ("geometry", C.c_int * 8), ("ms", C.c_double * 8)]


_P = C.c_void_p
_SIGS = {
    "mfx_last_error": (C.c_char_p, []),
    "mfx_version": (C.c_int, []),
    "mfx_basis_tables": (C.c_int, [C.c_int, C.c_int, C.c_int, _P, _P, _P, _P]),
    "mfx_bk3_apply_f64": (C.c_int, [C.c_int, C.c_int, _P, _P, _P, _P, _P, _P, C.c_uint32, _P]),
    "mfx_bk3_apply_f32": (C.c_int, [C.c_int, C.c_int, _P, _P, _P, _P, _P, _P, C.c_uint32, _P]),
    "mfx_bk5_apply_f64": (C.c_int, [C.c_int, _P, _P, _P, _P, _P, C.c_uint32, _P]),
    "mfx_bk5_apply_f32": (C.c_int, [C.c_int, _P, _P, _P, _P, _P, C.c_uint32, _P]),
    "mfx_bk3_evec_f64": (C.c_int, [C.c_int, C.c_int, _P, _P, _P, _P, _P, C.c_uint32, _P]),
    "mfx_bk5_evec_f64": (C.c_int, [C.c_int, _P, _P, _P, _P, C.c_uint32, _P]),
    "mfx_gather_f64": (C.c_int, [_P, C.c_uint32, C.c_int, _P, _P, _P]),
    "mfx_scatter_add_f64": (C.c_int, [_P, C.c_uint32, C.c_int, _P, _P, _P]),
    "mfx_op_create_box": (C.c_int, [C.POINTER(BoxDesc), C.POINTER(C.c_void_p)]),
    "mfx_op_create": (C.c_int, [C.POINTER(OpDesc), C.POINTER(C.c_void_p)]),
    "mfx_op_destroy": (C.c_int, [_P]),
    "mfx_op_n_dofs": (C.c_uint64, [_P]),
    "mfx_op_n_cells": (C.c_uint32, [_P]),
    "mfx_op_get_info": (C.c_int, [_P, C.POINTER(OpInfo)]),
    "mfx_op_set_deterministic": (C.c_int, [_P, C.c_int]),
    "mfx_op_vmult": (C.c_int, [_P, _P, _P, _P]),
    "mfx_op_vmult_add": (C.c_int, [_P, _P, _P, _P]),
    "mfx_op_vmult_f32": (C.c_int, [_P, _P, _P, _P]),
    "mfx_op_diagonal": (C.c_int, [_P, _P, _P]),
    "mfx_op_inverse_diagonal": (C.c_int, [_P, _P, _P]),
    "mfx_op_rhs": (C.c_int, [_P, C.c_int, _P, _P]),
    "mfx_op_dof_coordinates": (C.c_int, [_P, _P, _P]),
    "mfx_cg_solve": (C.c_int, [_P, _P, _P, _P, C.c_int, C.c_double, C.c_double,
                               C.POINTER(CgResult), _P]),
    "mfx_cg_solve_noalloc": (C.c_int, [_P, _P, _P, _P, C.c_int, C.c_double, C.c_double,
                                       C.POINTER(CgResult), _P]),
    "mfx_cg_noalloc_max_dofs": (C.c_uint64, []),
    "mfx_vec_random": (C.c_int, [_P, C.c_uint64, C.c_uint64, _P]),
    "mfx_vec_sine_modes": (C.c_int, [_P, C.c_uint64, C.c_int, _P, _P, _P, _P]),
    "mfx_vec_zero_constrained": (C.c_int, [_P, _P, _P]),
    "mfx_vec_dot": (C.c_int, [_P, _P, C.c_uint64, C.POINTER(C.c_double), _P]),
    "mfx_mg_create": (C.c_int, [C.POINTER(MgDesc), C.POINTER(C.c_void_p)]),
    "mfx_mg_destroy": (C.c_int, [_P]),
    "mfx_mg_n_levels": (C.c_int, [_P]),
    "mfx_mg_get_level": (C.c_int, [_P, C.c_int, C.POINTER(MgLevelInfo)]),
    "mfx_mg_vmult": (C.c_int, [_P, _P, _P, _P]),
    "mfx_mg_solve": (C.c_int, [_P, _P, _P, C.c_int, C.c_double, C.c_double,
                               C.POINTER(CgResult), _P]),
    "mfx_transfer_prolongate_f64": (C.c_int, [_P, _P, _P, _P, _P]),
    "mfx_transfer_restrict_f64": (C.c_int, [_P, _P, _P, _P, _P]),
    "mfx_chebyshev_f32": (C.c_int, [_P, _P, _P, _P, C.c_int, C.c_double, C.c_double, C.c_int,
                                    _P]),
    "mfx_batch_create": (C.c_int, [C.POINTER(BatchInstance), C.c_int, C.c_int,
                                   C.POINTER(C.c_void_p)]),
    "mfx_batch_destroy": (C.c_int, [_P]),
    "mfx_batch_n_instances": (C.c_int, [_P]),
    "mfx_batch_n_dofs": (C.c_uint64, [_P]),
    "mfx_batch_offsets": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "mfx_batch_instance_operator": (_P, [_P, C.c_int]),
    "mfx_batch_get_tuning": (C.c_int, [_P, C.c_int, C.POINTER(BatchTuning)]),
    "mfx_batch_set_variant": (C.c_int, [_P, C.c_int, C.c_int]),
    "mfx_batch_vmult": (C.c_int, [_P, _P, _P, _P]),
    "mfx_batch_inverse_diagonal": (C.c_int, [_P, _P, _P]),
    "mfx_batch_cg_solve": (C.c_int, [_P, _P, _P, _P, C.c_int, C.c_double, C.c_double, C.c_int,
                                     C.POINTER(CgResult), _P]),
    "mfx_nn_n_params": (C.c_uint64, []),
    "mfx_nn_create": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "mfx_nn_destroy": (C.c_int, [_P]),
    "mfx_nn_get_params": (C.c_int, [_P, _P]),
    "mfx_nn_set_params": (C.c_int, [_P, _P]),
    "mfx_nn_set_output_scale": (C.c_int, [_P, C.c_double]),
    "mfx_nn_output_scale": (C.c_double, [_P]),
    "mfx_nn_set_mode": (C.c_int, [_P, C.c_int, C.c_int]),
    "mfx_nn_set_coarse": (C.c_int, [_P, _P]),
    "mfx_nn_smoother_info": (C.c_int, [_P, _P, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double), _P]),
    "mfx_nn_forward": (C.c_int, [_P, C.POINTER(C.c_uint32), _P, _P, _P]),
    "mfx_nn_gradient": (C.c_int, [_P, C.POINTER(C.c_uint32), _P, _P, _P, _P]),
    "mfx_nn_train_step": (C.c_int, [_P, C.POINTER(C.c_uint32), C.c_int, _P, _P, C.c_double,
                                    C.POINTER(C.c_double), _P]),
    "mfx_nn_train_energy": (C.c_int, [_P, _P, C.c_int, _P, _P, C.c_double,
                                     C.POINTER(C.c_double), _P]),
    "mfx_nn_vmult": (C.c_int, [_P, _P, _P, _P, _P]),
    "mfx_nn_fcg_solve": (C.c_int, [_P, _P, _P, _P, C.c_int, C.c_double, C.c_double,
                                   C.POINTER(CgResult), _P]),
    "mfx_nn_restricted_residual": (C.c_int, [_P, _P, _P, _P, _P]),
    "mfx_nn_set_fallback": (C.c_int, [_P, _P, C.c_int, C.c_double]),
    "mfx_nn_fallback_step": (C.c_int, [_P]),
}

_lib = None


def lib() -> C.CDLL:
    """Load libmfx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        # torch ships its own HIP runtime; loading it first makes libmfx bind to
        # that one, so the process holds a single HIP runtime.
        import torch  # noqa: F401
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                "g.build()'` (or `make -C benchmarks_amd/csrc`). There is no CPU fallback.")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != MFX_OK:
        raise MfxError(status, lib().mfx_last_error().decode())


def exported_names() -> list:
    return list(_SIGS)
