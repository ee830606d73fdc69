"""Host-side mirror of PolyFEM's assembler and form API, backed by libpfgpu.

Names, argument meaning and error behaviour follow the reference:

* ``make_assembler`` -- ``AssemblerUtils::make_assembler`` (assembler/AssemblerUtils.cpp:55-122)
* ``NLAssembler`` -- ``ElasticityNLAssembler`` (assembler/Assembler.hpp:53-376):
  ``assemble_energy``, ``assemble_energy_per_element``, ``assemble_gradient``,
  ``assemble_hessian``, ``update_lame_params``
* ``ElasticForm`` -- ``solver::ElasticForm`` (solver/forms/ElasticForm.cpp:287-328,
  Form.hpp:26-56): ``value``, ``first_derivative``, ``second_derivative``
* ``force_material_derivative`` / ``force_shape_derivative`` --
  ``ElasticForceDerivative`` (optimization/force_derivatives/ElasticForceDerivative.cpp:35-264)

Every computation runs in the HIP library; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from ._ffi import check, lib

_NAMES = {"NeoHookean": _ffi.NEOHOOKEAN, "LinearElasticity": _ffi.LINEAR_ELASTICITY, "SaintVenant": _ffi.SAINT_VENANT,
          "FixedCorotational": _ffi.FIXED_COROTATIONAL, "MooneyRivlin": _ffi.MOONEY_RIVLIN}
_PRECISIONS = {"fp64": _ffi.FP64, "mixed": _ffi.MIXED}
_PATHS = {"fused": _ffi.PATH_FUSED, "plain": _ffi.PATH_PLAIN, "factored": _ffi.PATH_FACTORED}


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} entries, got {a.size}")
    return a


class NLAssembler:
    """Elasticity assembler on one tet mesh (P1 or P2, P1 geometry)."""

    def __init__(self, name: str, mesh, lame=None, young: float = None, poisson: float = None,
                 precision: str = "fp64", path: str = None, quadrature_order: int = 0, params=None):
        """``lame=(lambda, mu)`` (scalars or per element) or ``young`` and
        ``poisson``; MooneyRivlin takes ``params=(c1, c2, k)`` instead."""
        if name not in _NAMES:
            raise ValueError(f"unknown material {name!r}; expected one of {sorted(_NAMES)}")
        self._name = name
        self.mesh = mesh
        E = mesh.n_elements
        self._en = np.ascontiguousarray(mesh.elem_nodes, dtype=np.int32)
        self._vx = np.ascontiguousarray(mesh.vertices, dtype=np.float64)
        self._ev = np.ascontiguousarray(mesh.elem_vertices, dtype=np.int32)
        md = _ffi.MeshDesc(mesh.order, quadrature_order, E, mesh.n_nodes, _ptr(self._en), self._vx.shape[0],
                           _ptr(self._vx), _ptr(self._ev))
        if name == "MooneyRivlin":
            if params is None:
                raise ValueError("MooneyRivlin takes params=(c1, c2, k)")
            self._mr = [_f64(np.broadcast_to(np.asarray(v, dtype=np.float64), (E,))) for v in params]
            mat = _ffi.MaterialDesc(_NAMES[name], None, None, 0.0, 0.0, _ptr(self._mr[0]), _ptr(self._mr[1]),
                                    _ptr(self._mr[2]), 0.0, 0.0, 0.0)
        elif lame is not None:
            lam = _f64(np.broadcast_to(np.asarray(lame[0], dtype=np.float64), (E,)))
            mu = _f64(np.broadcast_to(np.asarray(lame[1], dtype=np.float64), (E,)))
            mat = _ffi.MaterialDesc(_NAMES[name], _ptr(lam), _ptr(mu), 0.0, 0.0)
        else:
            if young is None or poisson is None:
                raise ValueError("give either lame=(lambda, mu) or young and poisson")
            mat = _ffi.MaterialDesc(_NAMES[name], None, None, float(young), float(poisson))
        h = C.c_void_p()
        check(lib().pfgpu_create(C.byref(md), C.byref(mat), _PRECISIONS[precision], C.byref(h)))
        self._h = h
        self.precision = precision
        info = _ffi.Info()
        check(lib().pfgpu_info(self._h, C.byref(info)))
        # None keeps the library's default path for the material
        self.path = {v: k for k, v in _PATHS.items()}[info.path]
        if path is not None:
            self.set_path(path)
        self.n_dofs = info.n_dofs
        self.n_elements = info.n_elements
        self.n_vertices = info.n_vertices
        self.n_quad = info.n_quad
        self.nnz = info.nnz
        self._value_dtype = np.float32 if precision == "mixed" else np.float64

    # -- lifetime ------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            check(lib().pfgpu_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def name(self) -> str:
        return self._name

    def is_linear(self) -> bool:
        return self._name == "LinearElasticity"

    def size(self) -> int:
        return 3

    def info(self) -> _ffi.Info:
        info = _ffi.Info()
        check(lib().pfgpu_info(self._h, C.byref(info)))
        return info

    def set_path(self, path: str):
        check(lib().pfgpu_set_path(self._h, _PATHS[path]))
        self.path = path

    def autotune_dev(self, x_ptr: int) -> dict:
        """Times the launch-shape candidates of the E+G+H assembly at the
        device state x and keeps the fastest (pfgpu_autotune)."""
        r = _ffi.TuneReport()
        check(lib().pfgpu_autotune(self._h, C.c_void_p(x_ptr), C.byref(r)))
        self.path = {v: k for k, v in _PATHS.items()}[r.path]
        return {"path": self.path, "gather_threads": r.gather_threads, "hessian_threads": r.hessian_threads,
                "candidates": r.candidates, "ms": r.ms, "ms_plain_default": r.ms_plain_default,
                "ms_fused": r.ms_fused if r.ms_fused >= 0 else None,
                "ms_factored": r.ms_factored if r.ms_factored >= 0 else None}

    def set_tuning(self, path: str, gather_threads: int = 256, hessian_threads: int = 256):
        check(lib().pfgpu_set_tuning(self._h, _PATHS[path], gather_threads, hessian_threads))
        self.path = path

    def set_solver_options(self, hessian: str = "auto", preconditioner: str = "block_jacobi"):
        """Linear algebra of the solves' CG: hessian "auto" (assembled for P1,
        matrix-free for P2), "assembled", "matrix_free" or "assembled_psd"
        (element Hessians projected onto the PSD cone); preconditioner
        "block_jacobi" (inverse 3x3 diagonal blocks) or "none"."""
        hm = {"auto": _ffi.SOLVER_HESSIAN_AUTO, "assembled": _ffi.SOLVER_HESSIAN_ASSEMBLED,
              "matrix_free": _ffi.SOLVER_HESSIAN_MATRIX_FREE, "assembled_psd": _ffi.SOLVER_HESSIAN_ASSEMBLED_PSD}
        pm = {"none": _ffi.PRECOND_NONE, "block_jacobi": _ffi.PRECOND_BLOCK_JACOBI}
        if hessian not in hm or preconditioner not in pm:
            raise ValueError("unknown solver option")
        check(lib().pfgpu_set_solver_options(self._h, hm[hessian], pm[preconditioner]))

    def set_stream(self, stream_handle: int):
        check(lib().pfgpu_set_stream(self._h, C.c_void_p(stream_handle)))

    def update_mooney_rivlin_params(self, c1, c2, k):
        """Per-element MooneyRivlin c1, c2, k (scalars broadcast)."""
        v = [_f64(np.broadcast_to(np.asarray(a, dtype=np.float64), (self.n_elements,))) for a in (c1, c2, k)]
        check(lib().pfgpu_set_mooney_rivlin(self._h, _ptr(v[0]), _ptr(v[1]), _ptr(v[2])))

    def update_lame_params(self, lambdas, mus):
        lam = _f64(lambdas, self.n_elements)
        mu = _f64(mus, self.n_elements)
        check(lib().pfgpu_set_lame(self._h, _ptr(lam), _ptr(mu)))

    # -- assembly (host arrays) ---------------------------------------------
    def assemble_energy(self, x) -> float:
        x = _f64(x, self.n_dofs)
        out = C.c_double()
        check(lib().pfgpu_value(self._h, _ptr(x), C.byref(out)))
        return out.value

    def assemble_energy_per_element(self, x) -> np.ndarray:
        x = _f64(x, self.n_dofs)
        out = np.empty(self.n_elements)
        check(lib().pfgpu_value_per_element(self._h, _ptr(x), _ptr(out)))
        return out

    def assemble_gradient(self, x) -> np.ndarray:
        x = _f64(x, self.n_dofs)
        g = np.empty(self.n_dofs)
        check(lib().pfgpu_gradient(self._h, _ptr(x), _ptr(g)))
        return g

    def hessian_pattern(self):
        outer = np.empty(self.n_dofs + 1, dtype=np.int32)
        inner = np.empty(self.nnz, dtype=np.int32)
        check(lib().pfgpu_hessian_pattern(self._h, _ptr(outer), _ptr(inner)))
        return outer, inner

    def hessian_values(self, x, project_to_psd: bool = False) -> np.ndarray:
        x = _f64(x, self.n_dofs)
        vals = np.empty(self.nnz, dtype=self._value_dtype)
        check(lib().pfgpu_hessian_values(self._h, _ptr(x), int(project_to_psd), _ptr(vals)))
        return vals

    def assemble_hessian(self, x, project_to_psd: bool = False):
        """CSC matrix, like NLAssembler::assemble_hessian's StiffnessMatrix."""
        import scipy.sparse as sp
        outer, inner = self.hessian_pattern()
        vals = self.hessian_values(x, project_to_psd)
        return sp.csc_matrix((vals, inner, outer), shape=(self.n_dofs, self.n_dofs))

    def assemble_energy_gradient_hessian(self, x):
        x = _f64(x, self.n_dofs)
        e = C.c_double()
        g = np.empty(self.n_dofs)
        vals = np.empty(self.nnz, dtype=self._value_dtype)
        check(lib().pfgpu_energy_gradient_hessian(self._h, _ptr(x), C.byref(e), _ptr(g), _ptr(vals)))
        return e.value, g, vals

    def hessian_vector_product(self, x, v) -> np.ndarray:
        x = _f64(x, self.n_dofs)
        v = _f64(v, self.n_dofs)
        y = np.empty(self.n_dofs)
        check(lib().pfgpu_hvp(self._h, _ptr(x), _ptr(v), _ptr(y)))
        return y

    def force_material_derivative(self, x, adjoint) -> np.ndarray:
        """term[e] (lambda) and term[E+e] (mu), ElasticForceDerivative.cpp:97-129."""
        x = _f64(x, self.n_dofs)
        p = _f64(adjoint, self.n_dofs)
        out = np.empty(2 * self.n_elements)
        check(lib().pfgpu_material_vjp(self._h, _ptr(x), _ptr(p), _ptr(out)))
        return out

    def energy_lame_gradient(self, x) -> np.ndarray:
        """dE/dlambda_e (first E entries) and dE/dmu_e (last E), the explicit
        dependence of the elastic energy on the per-element Lame parameters."""
        x = _f64(x, self.n_dofs)
        out = np.empty(2 * self.n_elements)
        check(lib().pfgpu_energy_lame_gradient(self._h, _ptr(x), _ptr(out)))
        return out

    def set_vertices(self, vertices):
        """New rest positions of the mesh vertices (n_vertices x 3), same
        topology: the element geometry (and the mass matrix, if set) follow."""
        v = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
        if v.shape[0] != self.n_vertices:
            raise ValueError("one position per mesh vertex expected")
        check(lib().pfgpu_set_vertices(self._h, _ptr(v)))
        self._vx = v

    def set_load(self, f):
        """External load f (n_dofs) for the Newton solve, which then minimises
        E(x) - f^T x; None removes it."""
        if f is None:
            check(lib().pfgpu_set_load(self._h, None))
            return
        f = _f64(f, self.n_dofs)
        check(lib().pfgpu_set_load(self._h, _ptr(f)))

    def force_shape_derivative(self, x, adjoint) -> np.ndarray:
        """term[vertex*3 + d], ElasticForceDerivative.cpp:135-264."""
        x = _f64(x, self.n_dofs)
        p = _f64(adjoint, self.n_dofs)
        out = np.empty(3 * self.n_vertices)
        check(lib().pfgpu_shape_vjp(self._h, _ptr(x), _ptr(p), _ptr(out)))
        return out

    def material_jvp(self, x, dlambda, dmu) -> np.ndarray:
        x = _f64(x, self.n_dofs)
        dl = _f64(dlambda, self.n_elements)
        dm = _f64(dmu, self.n_elements)
        y = np.empty(self.n_dofs)
        check(lib().pfgpu_material_jvp(self._h, _ptr(x), _ptr(dl), _ptr(dm), _ptr(y)))
        return y

    # -- assembly (device pointers, asynchronous on the handle's stream) ------
    def energy_dev(self, x_ptr: int, e_ptr: int):
        check(lib().pfgpu_value_dev(self._h, C.c_void_p(x_ptr), C.c_void_p(e_ptr)))

    def gradient_dev(self, x_ptr: int, g_ptr: int):
        check(lib().pfgpu_gradient_dev(self._h, C.c_void_p(x_ptr), C.c_void_p(g_ptr)))

    def hessian_values_dev(self, x_ptr: int, v_ptr: int):
        check(lib().pfgpu_hessian_values_dev(self._h, C.c_void_p(x_ptr), C.c_void_p(v_ptr)))

    def energy_gradient_hessian_dev(self, x_ptr: int, e_ptr: int, g_ptr: int, v_ptr: int):
        check(lib().pfgpu_energy_gradient_hessian_dev(self._h, C.c_void_p(x_ptr), C.c_void_p(e_ptr),
                                                      C.c_void_p(g_ptr), C.c_void_p(v_ptr)))

    def hvp_dev(self, x_ptr: int, v_ptr: int, y_ptr: int):
        check(lib().pfgpu_hvp_dev(self._h, C.c_void_p(x_ptr), C.c_void_p(v_ptr), C.c_void_p(y_ptr)))

    def material_vjp_dev(self, x_ptr: int, p_ptr: int, out_ptr: int):
        check(lib().pfgpu_material_vjp_dev(self._h, C.c_void_p(x_ptr), C.c_void_p(p_ptr), C.c_void_p(out_ptr)))

    def shape_vjp_dev(self, x_ptr: int, p_ptr: int, out_ptr: int):
        check(lib().pfgpu_shape_vjp_dev(self._h, C.c_void_p(x_ptr), C.c_void_p(p_ptr), C.c_void_p(out_ptr)))

    def synchronize(self):
        check(lib().pfgpu_synchronize(self._h))


def make_assembler(name: str, mesh, **kwargs) -> NLAssembler:
    """AssemblerUtils::make_assembler for the two materials of this library."""
    return NLAssembler(name, mesh, **kwargs)


class ElasticForm:
    """solver::ElasticForm: value / first_derivative / second_derivative scaled
    by weight()/scale (Form.hpp:26-56).  The linear-elastic Hessian is assembled
    once and cached, as compute_cached_stiffness does (ElasticForm.cpp:411-418)."""

    def __init__(self, assembler: NLAssembler, weight: float = 1.0, scale: float = 1.0):
        self.assembler = assembler
        self._weight = weight
        self._scale = scale
        self._cached_stiffness = None

    def weight(self) -> float:
        return self._weight

    def set_weight(self, w: float):
        self._weight = w

    def _w(self):
        return self._weight / self._scale

    def value_unweighted(self, x):
        return self.assembler.assemble_energy(x)

    def value(self, x):
        return self._w() * self.value_unweighted(x)

    def value_per_element_unweighted(self, x):
        return self.assembler.assemble_energy_per_element(x)

    def first_derivative_unweighted(self, x):
        return self.assembler.assemble_gradient(x)

    def first_derivative(self, x):
        return self._w() * self.first_derivative_unweighted(x)

    def second_derivative_unweighted(self, x):
        if self.assembler.is_linear():
            if self._cached_stiffness is None:
                self._cached_stiffness = self.assembler.assemble_hessian(np.zeros(self.assembler.n_dofs))
            return self._cached_stiffness
        return self.assembler.assemble_hessian(x)

    def second_derivative(self, x):
        return self._w() * self.second_derivative_unweighted(x)

    def is_step_valid(self, x0, x1) -> bool:
        """ElasticForm::is_step_valid (ElasticForm.cpp:388-405): reject NaN gradients."""
        try:
            g = self.assembler.assemble_gradient(x1)
        except _ffi.PfgpuError as e:
            if e.status == 3:
                return False
            raise
        return bool(np.all(np.isfinite(g)))
