"""GPU tests of vertex updates and of the shape gradient through static solves.

pfgpu_set_vertices must leave a handle that computes what a handle created on
the moved mesh computes (bit for bit: the geometry is recomputed by the same
code), including the mass matrix and a CG solve whose graph was captured
before the move.  The shape gradient of a static objective,
dJ/dX = dJ/dX|explicit - force_shape_derivative(u, p) with H p = dJ/du,
must match central differences of J through full Newton solves.
"""
import dataclasses

import numpy as np
import pytest

from polyfem_amd import NewtonSolver, NLAssembler, NLProblem, TransientSimulation, cg_solve, solve_static_adjoint
from polyfem_amd.mesh import BASE_SEED, face_nodes, kuhn_box, random_lame, smooth_displacement

pytestmark = pytest.mark.gpu


def node_of_vertex(m):
    """P1/P2 node id of every mesh vertex (the first four local nodes are the vertices)."""
    v2n = np.empty(m.vertices.shape[0], dtype=np.int64)
    v2n[m.elem_vertices.ravel()] = m.elem_nodes[:, :4].ravel()
    return v2n


@pytest.mark.parametrize("order,precision", [(1, "fp64"), (2, "fp64"), (2, "mixed")])
def test_set_vertices_matches_fresh_handle(order, precision):
    m = kuhn_box(3, 2, 3, order=order)
    rng = np.random.default_rng(BASE_SEED + 60)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 61)
    x = smooth_displacement(m.nodes, 0.02, 1e-3, BASE_SEED + 62)
    V = m.vertices + rng.normal(size=m.vertices.shape) * 0.01
    fixed = np.zeros(m.n_dofs, dtype=np.uint8)
    fixed[:6] = 1
    b = rng.normal(size=m.n_dofs)
    print(f"set_vertices order={order} {precision}", flush=True)
    a = NLAssembler("NeoHookean", m, lame=(lam, mu), precision=precision)
    sim = TransientSimulation(a, rho=900.0)
    cg_solve(a, x, b, fixed, max_iterations=40, tolerance=0.0)  # captures the CG graph on the old geometry
    a.set_vertices(V)
    fresh = NLAssembler("NeoHookean", dataclasses.replace(m, vertices=V), lame=(lam, mu), precision=precision)
    sim_f = TransientSimulation(fresh, rho=900.0)
    assert a.assemble_energy(x) == fresh.assemble_energy(x)
    assert np.array_equal(a.assemble_gradient(x), fresh.assemble_gradient(x))
    assert np.array_equal(a.hessian_values(x), fresh.hessian_values(x))
    assert np.array_equal(sim.mass_matvec(b), sim_f.mass_matvec(b))
    y, it, _ = cg_solve(a, x, b, fixed, max_iterations=40, tolerance=0.0)
    yf, itf, _ = cg_solve(fresh, x, b, fixed, max_iterations=40, tolerance=0.0)
    assert it == itf and 0 < it <= 40
    assert np.array_equal(y, yf)


def test_set_lame_after_captured_solve():
    """A CG graph captured before pfgpu_set_lame reads the new parameters."""
    m = kuhn_box(2, 2, 3, order=2)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 63)
    lam2, mu2 = random_lame(m.n_elements, seed=BASE_SEED + 64)
    x = smooth_displacement(m.nodes, 0.02, 1e-3, BASE_SEED + 65)
    b = np.random.default_rng(BASE_SEED + 66).normal(size=m.n_dofs)
    print("set_lame after a captured CG", flush=True)
    a = NLAssembler("NeoHookean", m, lame=(lam, mu), precision="fp64")
    cg_solve(a, x, b, None, max_iterations=30, tolerance=0.0)
    a.update_lame_params(lam2, mu2)
    fresh = NLAssembler("NeoHookean", m, lame=(lam2, mu2), precision="fp64")
    y, _, _ = cg_solve(a, x, b, None, max_iterations=30, tolerance=0.0)
    yf, _, _ = cg_solve(fresh, x, b, None, max_iterations=30, tolerance=0.0)
    assert np.array_equal(y, yf)


@pytest.mark.parametrize("order", [1, 2])
def test_shape_gradient_matches_finite_differences(order):
    """J = 1/2 sum_S |X_v + u_v - X*_v|^2 over the vertices v of the plane
    z = 1/3 (bottom clamped, top pulled): the explicit term plus the adjoint
    shape term against central differences of J through Newton solves on the
    moved mesh (pfgpu_set_vertices)."""
    m = kuhn_box(2, 2, 3, order=order)
    bot, top = face_nodes(m, 2, 0.0), face_nodes(m, 2, 1.0)
    fixed = np.concatenate([3 * bot[:, None] + np.arange(3), 3 * top[:, None] + np.arange(3)]).ravel()
    vals = np.concatenate([np.zeros(3 * bot.size), np.tile([0.0, 0.0, 0.05], top.size)])
    prob = NLProblem(m.n_dofs, fixed, vals)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 67)
    v2n = node_of_vertex(m)
    S = np.flatnonzero(np.abs(m.vertices[:, 2] - 1.0 / 3.0) < 1e-9)
    Sd = (3 * v2n[S][:, None] + np.arange(3)).ravel()
    rng = np.random.default_rng(BASE_SEED + 68)
    Xstar = m.vertices[S] + rng.normal(size=(S.size, 3)) * 1e-3
    solver = NewtonSolver(max_iterations=30, grad_norm=1e-9, cg_max_iterations=5000, cg_tolerance=1e-12)
    print(f"shape gradient FD order={order}", flush=True)
    a = NLAssembler("NeoHookean", m, lame=(lam, mu), precision="fp64")

    def solve(V, u0):
        a.set_vertices(V)
        u, rep = solver.minimize(a, u0, prob)
        assert rep.converged
        return u

    def J(V, u):
        return 0.5 * np.sum((V[S] + u[Sd].reshape(-1, 3) - Xstar) ** 2)

    V0 = m.vertices.copy()
    u = solve(V0, np.zeros(m.n_dofs))
    r = V0[S] + u[Sd].reshape(-1, 3) - Xstar
    dJdu = np.zeros(m.n_dofs)
    dJdu[Sd] = r.ravel()
    p = solve_static_adjoint(a, u, dJdu, prob, tolerance=1e-13)
    explicit = np.zeros_like(V0)
    explicit[S] = r
    grad = explicit.ravel() - a.force_shape_derivative(u, p)
    h = 1e-5
    for _ in range(3):
        dV = rng.normal(size=V0.shape) * 0.02
        jp = J(V0 + h * dV, solve(V0 + h * dV, u))
        jm = J(V0 - h * dV, solve(V0 - h * dV, u))
        fd = (jp - jm) / (2 * h)
        lin = grad @ dV.ravel()
        assert abs(fd - lin) <= 1e-5 * abs(lin) + 1e-14, (fd, lin)


def test_degenerate_vertices_are_rejected_and_leave_the_handle_unchanged():
    from polyfem_amd import PfgpuError
    m = kuhn_box(2, 2, 2, order=1)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 69)
    x = smooth_displacement(m.nodes, 0.02, 1e-3, BASE_SEED + 70)
    print("degenerate vertex update", flush=True)
    a = NLAssembler("NeoHookean", m, lame=(lam, mu), precision="fp64")
    e0, g0 = a.assemble_energy(x), a.assemble_gradient(x)
    V = m.vertices.copy()
    ev = m.elem_vertices[0]
    V[ev[3]] = V[ev[0]]  # collapses element 0 (and its neighbours through that vertex)
    with pytest.raises(PfgpuError):
        a.set_vertices(V)
    assert a.assemble_energy(x) == e0 and np.array_equal(a.assemble_gradient(x), g0)
    with pytest.raises(PfgpuError):
        NLAssembler("NeoHookean", dataclasses.replace(m, vertices=V), lame=(lam, mu), precision="fp64")
