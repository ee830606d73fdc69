"""Host-side mirror of the callers of the assembly: the nonlinear solve of the
elastic form and the static adjoint solve, backed by libpfgpu.

* ``NLProblem`` -- ``solver::NLProblem`` (solver/NLProblem.cpp:583-629):
  ``full_to_reduced`` / ``reduced_to_full`` over the Dirichlet dofs.
* ``NewtonSolver`` -- polysolve's Newton solver with a CG linear solver and a
  backtracking (Armijo) line search, as configured by the reference's
  ``"nonlinear": {"solver": "Newton", "line_search": {"method": "Backtracking"}}``;
  runs entirely on the GPU (``pfgpu_newton_solve``).
* ``solve_static_adjoint`` -- ``solve_static_adjoint`` (solver/StateDiff.cpp:230-318):
  H(u)^T p = dJ/du with zero adjoint on Dirichlet dofs (``pfgpu_cg_solve``).

Every computation runs in the HIP library; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _ffi
from ._ffi import check, lib


def dirichlet_mask(n_dofs: int, fixed_dofs) -> np.ndarray:
    """uint8 mask over the dofs, 1 on Dirichlet dofs."""
    m = np.zeros(n_dofs, dtype=np.uint8)
    idx = np.asarray(fixed_dofs, dtype=np.int64).ravel()
    if idx.size and (idx.min() < 0 or idx.max() >= n_dofs):
        raise ValueError("Dirichlet dof out of range")
    m[idx] = 1
    return m


class NLProblem:
    """Reduced / full dof maps of a problem with Dirichlet dofs (NLProblem.cpp:583-629)."""

    def __init__(self, n_dofs: int, fixed_dofs, fixed_values=None):
        self.n_dofs = n_dofs
        self.mask = dirichlet_mask(n_dofs, fixed_dofs)
        self.fixed = np.flatnonzero(self.mask)
        self.free = np.flatnonzero(self.mask == 0)
        vals = np.zeros(self.fixed.size) if fixed_values is None else np.asarray(fixed_values, dtype=np.float64)
        if vals.size != self.fixed.size:
            raise ValueError("one value per Dirichlet dof expected")
        self.fixed_values = vals

    def full_to_reduced(self, full):
        return np.asarray(full, dtype=np.float64)[self.free]

    def reduced_to_full(self, reduced):
        full = np.zeros(self.n_dofs)
        full[self.free] = reduced
        full[self.fixed] = self.fixed_values
        return full


@dataclass
class NewtonReport:
    iterations: int
    cg_iterations: int
    line_search_halvings: int
    converged: bool
    energy: float
    grad_norm: float
    last_alpha: float
    gradient_steps: int
    cg_stops_tolerance: int = 0
    cg_stops_curvature: int = 0
    cg_stops_cap: int = 0
    initial_grad_norm: float = 0.0
    regularized_solves: int = 0
    last_regularization: float = 0.0


def _report(r) -> "NewtonReport":
    return NewtonReport(r.iterations, r.cg_iterations, r.line_search_halvings, bool(r.converged), r.energy,
                        r.grad_norm, r.last_alpha, r.gradient_steps, r.cg_stops_tolerance, r.cg_stops_curvature,
                        r.cg_stops_cap, r.initial_grad_norm, r.regularized_solves, r.last_regularization)


class NewtonSolver:
    """Newton-CG on the free dofs with a backtracking line search (GPU-resident)."""

    def __init__(self, max_iterations: int = 10, grad_norm: float = 1e-8, cg_max_iterations: int = 1000,
                 cg_tolerance: float = 1e-6, line_search: str = "Backtracking", armijo_c: float = 1e-4,
                 max_line_search: int = 30, use_graph: bool = True, relative_grad_norm: float = 0.0,
                 regularization=None):
        """``regularization=(delta_min, delta_max)`` turns on the regularized
        Newton fallback (H + delta D on non-positive curvature, D the
        block-diagonal of H); None keeps plain truncated Newton-CG."""
        if line_search not in ("Backtracking", "None"):
            raise ValueError("line_search must be 'Backtracking' or 'None'")
        rmin, rmax = regularization if regularization is not None else (0.0, 0.0)
        self.params = _ffi.NewtonParams(max_iterations, grad_norm, cg_max_iterations, cg_tolerance,
                                        1 if line_search == "Backtracking" else 0, armijo_c, max_line_search,
                                        1 if use_graph else 0, relative_grad_norm, rmin, rmax)

    def minimize(self, assembler, x, problem: NLProblem):
        """Returns (x*, report); x holds the initial guess on the free dofs and
        is overwritten with the problem's Dirichlet values on the fixed dofs."""
        xx = np.array(x, dtype=np.float64, copy=True).ravel()
        if xx.size != assembler.mesh.n_dofs:
            raise ValueError("x has the wrong size")
        xx[problem.fixed] = problem.fixed_values
        rep = _ffi.NewtonReport()
        check(lib().pfgpu_newton_solve(assembler._h, xx.ctypes.data_as(C.c_void_p),
                                       problem.mask.ctypes.data_as(C.c_void_p), C.byref(self.params), C.byref(rep)))
        return xx, _report(rep)


def cg_solve(assembler, x, rhs, fixed_mask=None, max_iterations: int = 1000, tolerance: float = 1e-8):
    """y with H(x)_ff y_f = rhs_f and y = 0 on fixed dofs (matrix-free CG).
    Returns (y, iterations, relative residual)."""
    n = assembler.mesh.n_dofs
    xx = np.ascontiguousarray(x, dtype=np.float64)
    b = np.ascontiguousarray(rhs, dtype=np.float64)
    if xx.size != n or b.size != n:
        raise ValueError("wrong vector size")
    y = np.zeros(n)
    it = C.c_int32(0)
    rel = C.c_double(0)
    m = None if fixed_mask is None else np.ascontiguousarray(fixed_mask, dtype=np.uint8)
    check(lib().pfgpu_cg_solve(assembler._h, xx.ctypes.data_as(C.c_void_p),
                               None if m is None else m.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p),
                               y.ctypes.data_as(C.c_void_p), max_iterations, tolerance, C.byref(it), C.byref(rel)))
    return y, it.value, rel.value


def solve_static_adjoint(assembler, u, dJdu, problem: NLProblem, max_iterations: int = 2000, tolerance: float = 1e-10):
    """Adjoint p of a static solve: H(u) p = dJ/du on the free dofs, p = 0 on
    Dirichlet dofs (StateDiff.cpp:230-318; the Hessian is symmetric)."""
    p, it, rel = cg_solve(assembler, u, dJdu, problem.mask, max_iterations, tolerance)
    return p


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class TransientSimulation:
    """Implicit-Euler elastodynamics with a checkpointed discrete adjoint.

    Mirrors the reference's transient solve (time_integrator/ImplicitEuler.cpp:13-31,
    solver/forms/InertiaForm.cpp:17-35 with the elastic form weighted by dt^2)
    and its backward recurrence (optimization/StateDiff.cpp:320-413,
    AdjointTools.cpp:852-887 and 992-1010).  The mass matrix is the
    consistent one (assembler/Mass.cpp:4-20) with density ``rho``, its
    row-sum lumping with ``lump_mass=True`` or ``"row_sum"`` (the reference's
    solver.advanced.lump_mass_matrix, State.cpp:1542-1545), or the HRZ
    diagonal with ``lump_mass="hrz"`` (HRZMass, Mass.cpp:48-86)."""

    def __init__(self, assembler, rho=1000.0, dt=1e-3, time_steps=10, checkpoint_every=1,
                 newton: NewtonSolver = None, lump_mass: bool = False):
        self.a = assembler
        E = assembler.mesh.n_elements
        kinds = {False: 0, True: 1, "consistent": 0, "row_sum": 1, "hrz": 2}
        if lump_mass not in kinds:
            raise ValueError("lump_mass must be False, True, 'consistent', 'row_sum' or 'hrz'")
        check(lib().pfgpu_set_mass_lumping(assembler._h, kinds[lump_mass]))
        if np.ndim(rho) == 0:
            check(lib().pfgpu_set_density(assembler._h, None, float(rho)))
        else:
            r = np.ascontiguousarray(rho, dtype=np.float64)
            if r.size != E:
                raise ValueError("one density per element expected")
            check(lib().pfgpu_set_density(assembler._h, _p(r), 0.0))
        self.params = _ffi.TransientParams(dt, time_steps, checkpoint_every)
        self.newton = newton or NewtonSolver(max_iterations=50, grad_norm=1e-10, cg_tolerance=1e-12)

    def mass_matvec(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        y = np.zeros_like(v)
        check(lib().pfgpu_mass_matvec(self.a._h, _p(v), _p(y)))
        return y

    def forward(self, x0, v0, problem: NLProblem = None, load=None):
        n = self.a.mesh.n_dofs
        x0 = np.array(x0, dtype=np.float64).ravel()
        v0 = np.ascontiguousarray(v0, dtype=np.float64).ravel()
        mask = None
        if problem is not None:
            x0[problem.fixed] = problem.fixed_values
            mask = problem.mask
        f = None if load is None else np.ascontiguousarray(load, dtype=np.float64)
        xT, vT = np.zeros(n), np.zeros(n)
        check(lib().pfgpu_transient_forward(self.a._h, _p(x0), _p(v0), _p(mask), _p(f), C.byref(self.params),
                                            C.byref(self.newton.params), _p(xT), _p(vT)))
        return xT, vT

    def stats(self) -> dict:
        """Newton and CG work of the last forward run and adjoint."""
        st = _ffi.TransientStats()
        check(lib().pfgpu_get_transient_stats(self.a._h, C.byref(st)))
        return {f: int(getattr(st, f)) for f, _ in st._fields_}

    def adjoint(self, dJdx=None, dJdv=None, cg_max_iterations=5000, cg_tolerance=1e-12):
        """(dJ/d(lambda, mu), dJ/dx0, dJ/dv0) of J(x_T, v_T) through the last forward run."""
        n, E = self.a.mesh.n_dofs, self.a.mesh.n_elements
        gx = None if dJdx is None else np.ascontiguousarray(dJdx, dtype=np.float64)
        gv = None if dJdv is None else np.ascontiguousarray(dJdv, dtype=np.float64)
        dth, dx0, dv0 = np.zeros(2 * E), np.zeros(n), np.zeros(n)
        check(lib().pfgpu_transient_adjoint(self.a._h, _p(gx), _p(gv), cg_max_iterations, cg_tolerance,
                                            _p(dth), _p(dx0), _p(dv0)))
        return dth, dx0, dv0


def concatenate_meshes(meshes):
    """One mesh holding several meshes with disjoint node and vertex ranges;
    returns (mesh, element offsets, node offsets)."""
    from .mesh import TetMesh
    order = meshes[0].order
    if any(m.order != order for m in meshes):
        raise ValueError("a batch holds meshes of one order")
    eo, no, vo = [0], [0], [0]
    for m in meshes:
        eo.append(eo[-1] + m.n_elements)
        no.append(no[-1] + m.n_nodes)
        vo.append(vo[-1] + m.vertices.shape[0])
    big = TetMesh(order=order,
                  vertices=np.concatenate([m.vertices for m in meshes]),
                  elem_vertices=np.concatenate([m.elem_vertices + vo[i] for i, m in enumerate(meshes)]).astype(np.int32),
                  nodes=np.concatenate([m.nodes for m in meshes]),
                  elem_nodes=np.concatenate([m.elem_nodes + no[i] for i, m in enumerate(meshes)]).astype(np.int32))
    return big, np.array(eo, dtype=np.int32), np.array(no, dtype=np.int32)


class BatchSolver:
    """Independent static problems solved together (many meshes of uneven
    size, P1 and P2, NeoHookean and linear elasticity, in one batch handle:
    every element kernel and node gather is one launch over all problems,
    pfgpu_batch_create / pfgpu_batch_newton).  ``problems`` is a list of
    dicts with keys ``mesh``, ``material`` ("NeoHookean" |
    "LinearElasticity"), ``lame`` (lambda, mu; scalars or per element),
    ``fixed_dofs``, ``fixed_values`` and optionally ``x0``."""

    def __init__(self, problems, precision="mixed", newton: NewtonSolver = None):
        from .assembler import _NAMES, _PRECISIONS
        self.problems = problems
        self.newton = newton or NewtonSolver(max_iterations=20, grad_norm=1e-8)
        n = len(problems)
        self._keep = []
        mds = (_ffi.MeshDesc * n)()
        mats = (_ffi.MaterialDesc * n)()
        self.nl = []
        for i, p in enumerate(problems):
            m = p["mesh"]
            E = m.n_elements
            en = np.ascontiguousarray(m.elem_nodes, dtype=np.int32)
            vx = np.ascontiguousarray(m.vertices, dtype=np.float64)
            ev = np.ascontiguousarray(m.elem_vertices, dtype=np.int32)
            lam = np.ascontiguousarray(np.broadcast_to(np.asarray(p["lame"][0], dtype=np.float64), (E,)))
            mu = np.ascontiguousarray(np.broadcast_to(np.asarray(p["lame"][1], dtype=np.float64), (E,)))
            self._keep += [en, vx, ev, lam, mu]
            mds[i] = _ffi.MeshDesc(m.order, 0, E, m.n_nodes, en.ctypes.data, vx.shape[0], vx.ctypes.data,
                                   ev.ctypes.data)
            mats[i] = _ffi.MaterialDesc(_NAMES[p["material"]], lam.ctypes.data, mu.ctypes.data, 0.0, 0.0,
                                        None, None, None, 0.0, 0.0, 0.0)
            self.nl.append(NLProblem(m.n_dofs, p["fixed_dofs"], p["fixed_values"]))
        h = C.c_void_p()
        check(lib().pfgpu_batch_create(n, mds, mats, _PRECISIONS[precision], C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            check(lib().pfgpu_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self):
        """Returns (list of x per problem, list of NewtonReport per problem)."""
        n = len(self.problems)
        xs = []
        for p, pr in zip(self.problems, self.nl):
            x = np.array(p["x0"], dtype=np.float64) if "x0" in p else np.zeros(p["mesh"].n_dofs)
            x[pr.fixed] = pr.fixed_values
            xs.append(x)
        xp = (C.c_void_p * n)(*[x.ctypes.data for x in xs])
        mp = (C.c_void_p * n)(*[pr.mask.ctypes.data for pr in self.nl])
        r = (_ffi.NewtonReport * n)()
        check(lib().pfgpu_batch_newton(self._h, xp, mp, C.byref(self.newton.params), r))
        return xs, [_report(rr) for rr in r]
