"""ctypes wrapper of the FP64 oracle (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` use this module; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")

NEOHOOKEAN = 0
LINEAR_ELASTICITY = 1
SAINT_VENANT = 2

_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _ip, C.c_int, C.c_int, _dp, _ip, _dp, _dp]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_n_quad.argtypes = [C.c_void_p]
        L.orc_n_local.argtypes = [C.c_void_p]
        L.orc_quadrature.argtypes = [C.c_int, _dp, _dp]
        L.orc_basis_grad.argtypes = [C.c_int, _dp, _dp]
        L.orc_basis_val.argtypes = [C.c_int, _dp, _dp]
        L.orc_geometry.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.orc_element_energy.restype = C.c_double
        L.orc_element_energy.argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_element_gradient.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.orc_element_hessian.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.orc_energy.restype = C.c_double
        L.orc_energy.argtypes = [C.c_void_p, _dp]
        L.orc_energy_per_element.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_gradient.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_hessian_nnz.restype = C.c_int64
        L.orc_hessian_nnz.argtypes = [C.c_void_p]
        L.orc_hessian.argtypes = [C.c_void_p, _dp, _ip, _ip, _dp]
        L.orc_material_vjp.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_shape_vjp.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_triplets_to_csc.restype = C.c_int64
        L.orc_triplets_to_csc.argtypes = [C.c_int, C.c_int64, _ip, _ip, _dp, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_baseline_create.restype = C.c_void_p
        L.orc_baseline_create.argtypes = [C.c_void_p, C.c_int]
        L.orc_baseline_destroy.argtypes = [C.c_void_p]
        L.orc_baseline_threads.argtypes = [C.c_void_p]
        L.orc_baseline_energy.restype = C.c_double
        L.orc_baseline_energy.argtypes = [C.c_void_p, _dp]
        L.orc_baseline_gradient.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_baseline_hessian.restype = C.c_int64
        L.orc_baseline_hessian.argtypes = [C.c_void_p, _dp, C.c_void_p]
        _lib = L
    return _lib


def quadrature(order: int):
    pts = np.zeros((64, 3))
    w = np.zeros(64)
    n = lib().orc_quadrature(order, pts, w)
    return pts[:n].copy(), w[:n].copy()


def basis_grad(order: int, xi):
    n = 4 if order == 1 else 10
    g = np.zeros((n, 3))
    lib().orc_basis_grad(order, np.ascontiguousarray(xi, dtype=np.float64), g)
    return g


def basis_val(order: int, xi):
    n = 4 if order == 1 else 10
    v = np.zeros(n)
    lib().orc_basis_val(order, np.ascontiguousarray(xi, dtype=np.float64), v)
    return v


def triplets_to_csc(n, rows, cols, vals):
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    L = lib()
    nnz = L.orc_triplets_to_csc(n, rows.size, rows, cols, vals, None, None, None)
    outer = np.zeros(n + 1, dtype=np.int32)
    inner = np.zeros(nnz, dtype=np.int32)
    values = np.zeros(nnz)
    L.orc_triplets_to_csc(n, rows.size, rows, cols, vals, outer.ctypes.data, inner.ctypes.data, values.ctypes.data)
    return outer, inner, values


class Oracle:
    """FP64 restatement of the reference's assembler on one mesh."""

    def __init__(self, mesh, lam, mu, material=NEOHOOKEAN, quad_order=0):
        self.mesh = mesh
        E = mesh.n_elements
        self._keep = [np.ascontiguousarray(mesh.elem_nodes, dtype=np.int32),
                      np.ascontiguousarray(mesh.vertices, dtype=np.float64),
                      np.ascontiguousarray(mesh.elem_vertices, dtype=np.int32),
                      np.ascontiguousarray(np.broadcast_to(np.asarray(lam, dtype=np.float64), (E,))),
                      np.ascontiguousarray(np.broadcast_to(np.asarray(mu, dtype=np.float64), (E,)))]
        en, vx, ev, la, m = self._keep
        self.material = material
        self.p = lib().orc_create(mesh.order, quad_order, material, E, en, mesh.n_nodes,
                                  vx.shape[0], vx, ev, la, m)
        if not self.p:
            raise ValueError("oracle rejected the problem")
        self.n_elements = E
        self.n_dofs = mesh.n_dofs
        self.n_local = lib().orc_n_local(self.p)
        self.n_quad = lib().orc_n_quad(self.p)

    def __del__(self):
        # at interpreter teardown the module globals may already be gone
        p = getattr(self, "p", None)
        self.p = None
        if p:
            try:
                lib().orc_destroy(p)
            except Exception:
                pass

    @staticmethod
    def _x(x):
        return np.ascontiguousarray(x, dtype=np.float64)

    def geometry(self, e):
        j = np.zeros(9 * self.n_quad)
        d = np.zeros(self.n_quad)
        lib().orc_geometry(self.p, e, j, d)
        return j.reshape(self.n_quad, 3, 3), d

    def element_energy(self, e, x):
        return lib().orc_element_energy(self.p, e, self._x(x))

    def element_gradient(self, e, x):
        g = np.zeros(3 * self.n_local)
        lib().orc_element_gradient(self.p, e, self._x(x), g)
        return g

    def element_hessian(self, e, x):
        H = np.zeros((3 * self.n_local, 3 * self.n_local))
        lib().orc_element_hessian(self.p, e, self._x(x), H)
        return H

    def energy(self, x):
        return lib().orc_energy(self.p, self._x(x))

    def energy_per_element(self, x):
        out = np.zeros(self.n_elements)
        lib().orc_energy_per_element(self.p, self._x(x), out)
        return out

    def gradient(self, x):
        g = np.zeros(self.n_dofs)
        lib().orc_gradient(self.p, self._x(x), g)
        return g

    def hessian(self, x):
        """(outer, inner, values) of the CSC Hessian; the pattern is symmetric."""
        nnz = lib().orc_hessian_nnz(self.p)
        outer = np.zeros(self.n_dofs + 1, dtype=np.int32)
        inner = np.zeros(nnz, dtype=np.int32)
        values = np.zeros(nnz)
        lib().orc_hessian(self.p, self._x(x), outer, inner, values)
        return outer, inner, values

    def hessian_psd(self, x):
        """(outer, inner, values) of the Hessian summed from element Hessians
        projected onto the PSD cone, as NLAssembler::assemble_hessian with
        project_to_psd (Assembler.cpp:645-771) calling ipc::project_to_psd of
        ipc-toolkit@b40e9c0 (CLAMP): A = V diag(d) V^T, eigenvalues <= 0 set to
        0, and A unchanged when its smallest eigenvalue is >= 0.  The
        eigen-decomposition is numpy.linalg.eigh (LAPACK)."""
        n = self.n_local
        rows, cols, vals = [], [], []
        for e in range(self.n_elements):
            H = self.element_hessian(e, x)
            d, V = np.linalg.eigh(H)
            if d[0] < 0.0:
                H = (V * np.where(d > 0.0, d, 0.0)) @ V.T
            g = (3 * self.mesh.elem_nodes[e][:, None] + np.arange(3)).reshape(-1)
            rows.append(np.repeat(g, 3 * n))
            cols.append(np.tile(g, 3 * n))
            vals.append(H.reshape(-1))
        return triplets_to_csc(self.n_dofs, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))

    def hessian_scipy(self, x):
        import scipy.sparse as sp
        outer, inner, values = self.hessian(x)
        return sp.csc_matrix((values, inner, outer), shape=(self.n_dofs, self.n_dofs))

    def material_vjp(self, x, adjoint):
        out = np.zeros(2 * self.n_elements)
        lib().orc_material_vjp(self.p, self._x(x), self._x(adjoint), out)
        return out

    def shape_vjp(self, x, adjoint):
        out = np.zeros(3 * self.mesh.vertices.shape[0])
        lib().orc_shape_vjp(self.p, self._x(x), self._x(adjoint), out)
        return out


class Baseline:
    """Reference-algorithm CPU baseline: threaded assembly with per-thread storage."""

    def __init__(self, oracle: Oracle, n_threads: int):
        self.o = oracle
        self.b = lib().orc_baseline_create(oracle.p, n_threads)
        self.threads = lib().orc_baseline_threads(self.b)

    def __del__(self):
        b = getattr(self, "b", None)
        self.b = None
        if b:
            try:
                lib().orc_baseline_destroy(b)
            except Exception:
                pass

    def energy(self, x):
        return lib().orc_baseline_energy(self.b, Oracle._x(x))

    def gradient(self, x):
        g = np.zeros(self.o.n_dofs)
        lib().orc_baseline_gradient(self.b, Oracle._x(x), g)
        return g

    def hessian(self, x, values=None):
        ptr = values.ctypes.data if values is not None else None
        return lib().orc_baseline_hessian(self.b, Oracle._x(x), ptr)
