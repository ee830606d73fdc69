"""ctypes binding of the libpfgpu C-ABI (include/pfgpu.h).

The library must have been built (``python -m polyfem_amd.build_lib``); there
is no fallback: loading fails loudly when it is missing.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PFGPU_LIB", os.path.join(HERE, "libpfgpu.so"))

PFGPU_OK = 0
STATUS_NAMES = {0: "OK", 1: "INVALID_ARGUMENT", 2: "HIP", 3: "INVALID_STATE", 4: "UNSUPPORTED", 5: "OUT_OF_MEMORY"}
OK, ERR_INVALID_ARGUMENT, ERR_HIP, ERR_INVALID_STATE, ERR_UNSUPPORTED, ERR_OUT_OF_MEMORY = range(6)
NEOHOOKEAN, LINEAR_ELASTICITY, SAINT_VENANT, FIXED_COROTATIONAL, MOONEY_RIVLIN = 0, 1, 2, 3, 4
FP64, MIXED = 0, 1
PATH_FUSED, PATH_PLAIN, PATH_FACTORED = 0, 1, 2


class MeshDesc(C.Structure):
    _fields_ = [("order", C.c_int32), ("quadrature_order", C.c_int32), ("n_elements", C.c_int32),
                ("n_nodes", C.c_int32), ("elem_nodes", C.c_void_p), ("n_vertices", C.c_int32),
                ("vertices", C.c_void_p), ("elem_vertices", C.c_void_p)]


class MaterialDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("lam", C.c_void_p), ("mu", C.c_void_p),
                ("young", C.c_double), ("poisson", C.c_double),
                ("c1", C.c_void_p), ("c2", C.c_void_p), ("k", C.c_void_p),
                ("c1_uniform", C.c_double), ("c2_uniform", C.c_double), ("k_uniform", C.c_double)]


class Info(C.Structure):
    _fields_ = [("order", C.c_int32), ("n_local", C.c_int32), ("n_quad", C.c_int32), ("material", C.c_int32),
                ("precision", C.c_int32), ("path", C.c_int32), ("n_elements", C.c_int32), ("n_nodes", C.c_int32),
                ("n_vertices", C.c_int32), ("n_dofs", C.c_int32), ("nnz", C.c_int64), ("device_bytes", C.c_int64)]


class NewtonParams(C.Structure):
    _fields_ = [("max_iterations", C.c_int32), ("grad_norm", C.c_double), ("cg_max_iterations", C.c_int32),
                ("cg_tolerance", C.c_double), ("line_search", C.c_int32), ("armijo_c", C.c_double),
                ("max_line_search", C.c_int32), ("use_graph", C.c_int32), ("relative_grad_norm", C.c_double),
                ("regularization_min", C.c_double), ("regularization_max", C.c_double)]


class NewtonReport(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("cg_iterations", C.c_int32), ("line_search_halvings", C.c_int32),
                ("converged", C.c_int32), ("energy", C.c_double), ("grad_norm", C.c_double),
                ("last_alpha", C.c_double), ("gradient_steps", C.c_int32), ("cg_stops_tolerance", C.c_int32),
                ("cg_stops_curvature", C.c_int32), ("cg_stops_cap", C.c_int32), ("initial_grad_norm", C.c_double),
                ("regularized_solves", C.c_int32), ("last_regularization", C.c_double)]


class TransientParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("time_steps", C.c_int32), ("checkpoint_every", C.c_int32)]


class TransientStats(C.Structure):
    _fields_ = [(f, C.c_int64) for f in ("steps", "newton_iterations", "cg_iterations", "line_search_halvings",
                                         "recompute_newton_iterations", "recompute_cg_iterations",
                                         "adjoint_cg_iterations")]


class TuneReport(C.Structure):
    _fields_ = [("path", C.c_int32), ("gather_threads", C.c_int32), ("hessian_threads", C.c_int32),
                ("candidates", C.c_int32), ("ms", C.c_double), ("ms_plain_default", C.c_double),
                ("ms_fused", C.c_double), ("ms_factored", C.c_double)]


class PfgpuError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"pfgpu {STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


SOLVER_HESSIAN_AUTO, SOLVER_HESSIAN_ASSEMBLED, SOLVER_HESSIAN_MATRIX_FREE, SOLVER_HESSIAN_ASSEMBLED_PSD = 0, 1, 2, 3
PRECOND_NONE, PRECOND_BLOCK_JACOBI = 0, 1

vp = C.c_void_p
_SIGNATURES = {
    "pfgpu_create": [C.POINTER(MeshDesc), C.POINTER(MaterialDesc), C.c_int32, C.POINTER(vp)],
    "pfgpu_destroy": [vp],
    "pfgpu_info": [vp, C.POINTER(Info)],
    "pfgpu_set_stream": [vp, vp],
    "pfgpu_set_path": [vp, C.c_int32],
    "pfgpu_set_lame": [vp, vp, vp],
    "pfgpu_set_mooney_rivlin": [vp, vp, vp, vp],
    "pfgpu_value": [vp, vp, vp],
    "pfgpu_value_per_element": [vp, vp, vp],
    "pfgpu_gradient": [vp, vp, vp],
    "pfgpu_hessian_nnz": [vp, vp],
    "pfgpu_hessian_pattern": [vp, vp, vp],
    "pfgpu_hessian_values": [vp, vp, C.c_int32, vp],
    "pfgpu_energy_gradient_hessian": [vp, vp, vp, vp, vp],
    "pfgpu_hvp": [vp, vp, vp, vp],
    "pfgpu_material_vjp": [vp, vp, vp, vp],
    "pfgpu_shape_vjp": [vp, vp, vp, vp],
    "pfgpu_material_jvp": [vp, vp, vp, vp, vp],
    "pfgpu_value_dev": [vp, vp, vp],
    "pfgpu_value_noalloc": [vp, vp, vp],
    "pfgpu_gradient_dev": [vp, vp, vp],
    "pfgpu_hessian_values_dev": [vp, vp, vp],
    "pfgpu_energy_gradient_hessian_dev": [vp, vp, vp, vp, vp],
    "pfgpu_hvp_dev": [vp, vp, vp, vp],
    "pfgpu_material_vjp_dev": [vp, vp, vp, vp],
    "pfgpu_shape_vjp_dev": [vp, vp, vp, vp],
    "pfgpu_hessian_pattern_dev": [vp, vp, vp],
    "pfgpu_synchronize": [vp],
    "pfgpu_newton_solve": [vp, vp, vp, C.POINTER(NewtonParams), C.POINTER(NewtonReport)],
    "pfgpu_newton_solve_dev": [vp, vp, vp, C.POINTER(NewtonParams), C.POINTER(NewtonReport)],
    "pfgpu_cg_solve": [vp, vp, vp, vp, vp, C.c_int32, C.c_double, C.POINTER(C.c_int32), C.POINTER(C.c_double)],
    "pfgpu_cg_solve_dev": [vp, vp, vp, vp, vp, C.c_int32, C.c_double, C.c_int32, C.POINTER(C.c_int32),
                           C.POINTER(C.c_double)],
    "pfgpu_set_load": [vp, vp],
    "pfgpu_set_vertices": [vp, vp],
    "pfgpu_get_transient_stats": [vp, vp],
    "pfgpu_set_solver_options": [vp, C.c_int32, C.c_int32],
    "pfgpu_autotune": [vp, vp, C.POINTER(TuneReport)],
    "pfgpu_set_tuning": [vp, C.c_int32, C.c_int32, C.c_int32],
    "pfgpu_newton_solve_batch": [vp, vp, vp, C.c_int32, vp, vp, C.POINTER(NewtonParams), vp],
    "pfgpu_batch_create": [C.c_int32, vp, vp, C.c_int32, C.POINTER(vp)],
    "pfgpu_batch_newton": [vp, vp, vp, C.POINTER(NewtonParams), vp],
    "pfgpu_energy_lame_gradient": [vp, vp, vp],
    "pfgpu_set_density": [vp, vp, C.c_double],
    "pfgpu_set_mass_lumping": [vp, C.c_int32],
    "pfgpu_mass_matvec": [vp, vp, vp],
    "pfgpu_mass_matvec_dev": [vp, vp, vp],
    "pfgpu_transient_forward": [vp, vp, vp, vp, vp, C.POINTER(TransientParams), C.POINTER(NewtonParams), vp, vp],
    "pfgpu_transient_adjoint": [vp, vp, vp, C.c_int32, C.c_double, vp, vp, vp],
}

_lib = None


def exported_symbols():
    return sorted(list(_SIGNATURES) + ["pfgpu_last_error"])


def lib():
    """Load libpfgpu.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is missing: build it with `python -m polyfem_amd.build_lib`")
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.pfgpu_last_error.restype = C.c_char_p
        L.pfgpu_last_error.argtypes = []
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != PFGPU_OK:
        raise PfgpuError(status, lib().pfgpu_last_error().decode())
