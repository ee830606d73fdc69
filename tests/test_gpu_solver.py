"""GPU tests of the solves that call the assembly: Newton-CG on the free dofs
and the static adjoint (runs on an MI355X).

The exact Newton step of the oracle (a sparse direct solve of the FP64
oracle Hessian restricted to the free dofs) pins the Newton-CG step; central
finite differences of the objective through full nonlinear solves pin the
adjoint gradient.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import pyoracle as po
from polyfem_amd import NewtonSolver, NLAssembler, NLProblem, cg_solve, solve_static_adjoint
from polyfem_amd.mesh import BASE_SEED, face_nodes, kuhn_box, random_lame, smooth_displacement

pytestmark = pytest.mark.gpu
MAT = {"NeoHookean": po.NEOHOOKEAN, "LinearElasticity": po.LINEAR_ELASTICITY}


def clamped_box(order, grid=(3, 3, 3), pull=0.1):
    """Bottom face fixed, top face moved by (0, 0, pull) (the small-box Dirichlet setup)."""
    m = kuhn_box(*grid, order=order)
    bot, top = face_nodes(m, 2, 0.0), face_nodes(m, 2, 1.0)
    fixed = np.concatenate([3 * bot[:, None] + np.arange(3), 3 * top[:, None] + np.arange(3)]).ravel()
    vals = np.concatenate([np.zeros(3 * bot.size), np.tile([0.0, 0.0, pull], top.size)])
    return m, NLProblem(m.n_dofs, fixed, vals)


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(np.asarray(b))), 1e-300))


@pytest.mark.parametrize("order,name", [(1, "NeoHookean"), (2, "NeoHookean"), (1, "LinearElasticity")])
def test_newton_step_is_exact_newton_step(order, name):
    m, prob = clamped_box(order)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 3)
    x0 = smooth_displacement(m.nodes, 0.02, 1e-3, BASE_SEED + 4)
    x0[prob.fixed] = prob.fixed_values
    print(f"newton step order={order} {name} dofs={m.n_dofs}", flush=True)
    a = NLAssembler(name, m, lame=(lam, mu), precision="fp64")
    x1, rep = NewtonSolver(max_iterations=1, grad_norm=0.0, cg_max_iterations=20000, cg_tolerance=1e-13,
                           line_search="None").minimize(a, x0, prob)
    assert rep.iterations == 1 and rep.last_alpha == 1.0
    o = po.Oracle(m, lam, mu, material=MAT[name])
    H = o.hessian_scipy(x0).tocsr()[prob.free][:, prob.free]
    g = o.gradient(x0)[prob.free]
    ref = x0.copy()
    ref[prob.free] += spla.spsolve(H.tocsc(), -g)
    assert np.array_equal(x1[prob.fixed], prob.fixed_values)
    assert rel(x1 - x0, ref - x0) < 1e-8


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("path", ["fused", "plain", "factored"])
def test_newton_converges(precision, path):
    m, prob = clamped_box(2, (2, 2, 3))
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 5)
    x0 = prob.reduced_to_full(np.zeros(prob.free.size))
    o = po.Oracle(m, lam, mu)
    g0 = np.linalg.norm(o.gradient(x0)[prob.free])
    # FP32 element kernels leave ~1e-6 relative noise in the gradient
    rtol = 1e-10 if precision == "fp64" else 1e-5
    print(f"newton converge {precision} {path} |g0|={g0:.3e}", flush=True)
    a = NLAssembler("NeoHookean", m, lame=(lam, mu), precision=precision, path=path)
    x, rep = NewtonSolver(max_iterations=25, grad_norm=rtol * g0, cg_max_iterations=4000,
                          cg_tolerance=1e-10 if precision == "fp64" else 1e-6).minimize(a, x0, prob)
    assert rep.converged, rep
    assert np.linalg.norm(o.gradient(x)[prob.free]) <= (rtol if precision == "fp64" else 10 * rtol) * g0
    assert o.energy(x) < o.energy(x0)


def test_newton_large_pull_and_roundoff_regime():
    # a large pull, then Newton iterations past the point where FP64 energy
    # differences resolve the decrease: the solve must keep going without a
    # line-search failure and end at a finite, lower energy
    m, prob = clamped_box(1, (2, 2, 2), pull=0.6)
    a = NLAssembler("NeoHookean", m, young=1e6, poisson=0.4, precision="fp64")
    x0 = prob.reduced_to_full(np.zeros(prob.free.size))
    g0 = np.linalg.norm(a.assemble_gradient(x0)[prob.free])
    x, rep = NewtonSolver(max_iterations=12, grad_norm=0.0, cg_tolerance=1e-12).minimize(a, x0, prob)
    print("large-pull case:", rep, flush=True)
    assert np.isfinite(rep.energy) and rep.iterations == 12 and rep.grad_norm <= 1e-9 * g0
    assert np.isfinite(a.assemble_energy(x)) and a.assemble_energy(x) < a.assemble_energy(x0)


def test_cg_graph_replays_match_eager():
    torch = pytest.importorskip("torch")
    from polyfem_amd import _ffi
    m, prob = clamped_box(2, (2, 2, 2))
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 6)
    x = smooth_displacement(m.nodes, 0.02, 1e-3, BASE_SEED + 7)
    a = NLAssembler("NeoHookean", m, lame=(lam, mu), precision="mixed")
    rng = np.random.default_rng(BASE_SEED + 8)
    dev = torch.device("cuda:0")
    dx = torch.tensor(x, device=dev)
    dm = torch.tensor(prob.mask, device=dev)
    out = {}
    for trial in range(3):
        b = rng.normal(size=m.n_dofs)
        db = torch.tensor(b, device=dev)
        for graph in (0, 1, 1, 0, 1):
            y = torch.zeros(m.n_dofs, dtype=torch.float64, device=dev)
            it, rr = C.c_int32(0), C.c_double(0)
            _ffi.check(_ffi.lib().pfgpu_cg_solve_dev(a._h, C.c_void_p(dx.data_ptr()), C.c_void_p(dm.data_ptr()),
                                                     C.c_void_p(db.data_ptr()), C.c_void_p(y.data_ptr()), 300, 0.0,
                                                     graph, C.byref(it), C.byref(rr)))
            _ffi.check(_ffi.lib().pfgpu_synchronize(a._h))
            yy = y.cpu().numpy()
            dx.add_(0.0)  # eager work of the program between replays
            torch.cuda.synchronize()
            if graph == 0 and (trial, 0) not in out:
                out[(trial, 0)] = yy
                out[(trial, 1)] = it.value
            # the iterations taken (the device stops at a non-positive
            # curvature once the residual has vanished to rounding)
            assert 0 < it.value <= 300 and it.value == out[(trial, 1)]
            assert np.array_equal(yy, out[(trial, 0)]), f"trial {trial} graph={graph} differs from eager"


def test_static_adjoint_residual():
    m, prob = clamped_box(2, (2, 2, 3))
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 9)
    u = smooth_displacement(m.nodes, 0.02, 1e-3, BASE_SEED + 10)
    u[prob.fixed] = prob.fixed_values
    a = NLAssembler("NeoHookean", m, lame=(lam, mu), precision="fp64")
    rhs = np.random.default_rng(BASE_SEED + 11).normal(size=m.n_dofs)
    p = solve_static_adjoint(a, u, rhs, prob, tolerance=1e-12)
    assert np.all(p[prob.fixed] == 0.0)
    H = po.Oracle(m, lam, mu).hessian_scipy(u).tocsr()[prob.free][:, prob.free]
    r = H @ p[prob.free] - rhs[prob.free]
    assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(rhs[prob.free])


def test_adjoint_material_gradient_matches_finite_differences():
    """dJ/d(lambda, mu) of J(u(theta)) = 1/2 |u_S - u*|^2 with u(theta) the
    static equilibrium: adjoint H p = dJ/du, dJ/dtheta = -p^T dg/dtheta
    (the material VJP), checked by central differences of J through full
    Newton solves."""
    m, prob = clamped_box(1, (2, 2, 3), pull=0.05)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 12)
    S = np.flatnonzero(np.abs(m.nodes[:, 2] - 1.0 / 3.0) < 1e-9)
    Sd = (3 * S[:, None] + np.arange(3)).ravel()
    target = np.random.default_rng(BASE_SEED + 13).normal(size=Sd.size) * 1e-3
    solver = NewtonSolver(max_iterations=30, grad_norm=1e-9, cg_max_iterations=5000, cg_tolerance=1e-12)

    def solve(lm, mm):
        a = NLAssembler("NeoHookean", m, lame=(lm, mm), precision="fp64")
        u, rep = solver.minimize(a, np.zeros(m.n_dofs), prob)
        assert rep.converged
        return a, u

    def J(u):
        return 0.5 * np.sum((u[Sd] - target) ** 2)

    a, u = solve(lam, mu)
    dJdu = np.zeros(m.n_dofs)
    dJdu[Sd] = u[Sd] - target
    p = solve_static_adjoint(a, u, dJdu, prob, tolerance=1e-13)
    grad = a.force_material_derivative(u, p)
    rng = np.random.default_rng(BASE_SEED + 14)
    for _ in range(3):
        dl = rng.normal(size=m.n_elements) * 0.05 * lam
        dm = rng.normal(size=m.n_elements) * 0.05 * mu
        h = 1e-4
        jp = J(solve(lam + h * dl, mu + h * dm)[1])
        jm = J(solve(lam - h * dl, mu - h * dm)[1])
        fd = (jp - jm) / (2 * h)
        lin = grad[:m.n_elements] @ dl + grad[m.n_elements:] @ dm
        assert abs(fd - lin) <= 1e-5 * max(abs(lin), 1e-12) + 1e-14, (fd, lin)


@pytest.mark.parametrize("order,precision", [(1, "fp64"), (2, "fp64"), (1, "mixed")])
def test_assembled_and_matrix_free_operators_agree(order, precision):
    """One unpreconditioned CG step from 0 is y = (b.b / b.Ab) b: the
    block-sparse product of the assembled values and the matrix-free element
    product give the same step; full solves agree with and without the
    block-Jacobi preconditioner."""
    m, prob = clamped_box(order, (2, 3, 2))
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 70)
    x = smooth_displacement(m.nodes, 0.02, 1e-3, BASE_SEED + 71)
    x[prob.fixed] = prob.fixed_values
    b = np.random.default_rng(BASE_SEED + 72).normal(size=m.n_dofs)
    print(f"operators order={order} {precision}", flush=True)
    a = NLAssembler("NeoHookean", m, lame=(lam, mu), precision=precision)
    one, full = {}, {}
    for hess in ("assembled", "matrix_free"):
        for pre in ("none", "block_jacobi"):
            a.set_solver_options(hess, pre)
            if pre == "none":
                one[hess] = cg_solve(a, x, b, prob.mask, max_iterations=1, tolerance=0.0)[0]
            full[(hess, pre)], it, r = cg_solve(a, x, b, prob.mask, max_iterations=2000, tolerance=1e-11)
            assert r <= 1e-11 or precision == "mixed", (hess, pre, r)
    tol1 = 1e-12 if precision == "fp64" else 1e-5
    assert rel(one["assembled"], one["matrix_free"]) < tol1
    ref = full[("matrix_free", "none")]
    tol = 1e-9 if precision == "fp64" else 1e-3
    for k, y in full.items():
        assert rel(y, ref) < tol, k
    # the solve: H_ff y_f = b_f (FP64 oracle Hessian)
    if precision == "fp64":
        H = po.Oracle(m, lam, mu).hessian_scipy(x).tocsr()[prob.free][:, prob.free]
        res = H @ full[("assembled", "block_jacobi")][prob.free] - b[prob.free]
        assert np.linalg.norm(res) <= 1e-9 * np.linalg.norm(b[prob.free])


@pytest.mark.parametrize("order", [1, 2])
def test_saint_venant_newton_converges(order):
    """Newton on SaintVenant (plain path): quadratic convergence to the
    minimiser; the final gradient agrees with the oracle's."""
    m, prob = clamped_box(order, (2, 2, 3), pull=0.1)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 90)
    a = NLAssembler("SaintVenant", m, lame=(lam, mu), precision="fp64")
    print(f"SaintVenant Newton order={order}", flush=True)
    x0 = np.zeros(m.n_dofs)
    x0[prob.fixed] = prob.fixed_values
    g0 = np.linalg.norm(a.assemble_gradient(x0)[prob.free])
    x, rep = NewtonSolver(max_iterations=30, grad_norm=1e-9 * g0, cg_max_iterations=5000,
                          cg_tolerance=1e-12).minimize(a, np.zeros(m.n_dofs), prob)
    assert rep.converged and rep.iterations <= 10, rep
    go = po.Oracle(m, lam, mu, material=po.SAINT_VENANT).gradient(x)
    assert np.linalg.norm(go[prob.free]) <= 1e-8 * g0


def test_newton_with_psd_projected_operator():
    """The assembled operator from PSD-projected element Hessians (the
    reference's project_to_psd option): Newton still converges to the same
    minimiser from a compressed start."""
    m, prob = clamped_box(1, (2, 2, 3), pull=-0.2)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 95)
    print("Newton with the PSD-projected operator", flush=True)
    sols = []
    for hess in ("assembled", "assembled_psd"):
        a = NLAssembler("NeoHookean", m, lame=(lam, mu), precision="fp64")
        a.set_solver_options(hess, "block_jacobi")
        x0 = np.zeros(m.n_dofs)
        x0[2::3] = -0.2 * m.nodes[:, 2]
        x0[prob.fixed] = prob.fixed_values
        g0 = np.linalg.norm(a.assemble_gradient(x0)[prob.free])
        x, rep = NewtonSolver(max_iterations=60, grad_norm=1e-9 * g0, cg_max_iterations=5000,
                              cg_tolerance=1e-12).minimize(a, x0, prob)
        assert rep.converged, (hess, rep)
        sols.append(x)
    assert rel(sols[1], sols[0]) < 1e-7


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("name", ["NeoHookean", "LinearElasticity"])
def test_preconditioner_from_node_diagonals(order, precision, name):
    """On the factored path a matrix-free operator takes its block-Jacobi
    blocks from the factor records (no assembly); the plain path reads them
    from the assembled values.  Fixed-count PCG iterates agree."""
    m, prob = clamped_box(order, (3, 2, 3))
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 73)
    x = smooth_displacement(m.nodes, 0.02, 1e-3, BASE_SEED + 74)
    x[prob.fixed] = prob.fixed_values
    b = np.random.default_rng(BASE_SEED + 75).normal(size=m.n_dofs)
    print(f"node-diagonal preconditioner order={order} {precision} {name}", flush=True)
    ys = {}
    for path in ("plain", "factored"):
        a = NLAssembler(name, m, lame=(lam, mu), precision=precision, path=path)
        a.set_solver_options("matrix_free", "block_jacobi")
        ys[path], it, _ = cg_solve(a, x, b, prob.mask, max_iterations=15, tolerance=0.0)
        assert it == 15
    assert rel(ys["factored"], ys["plain"]) < (1e-10 if precision == "fp64" else 1e-4)


def test_captured_value_survives_other_entry_points():
    """A graph captured over pfgpu_value_dev stays valid after other entry
    points ran (the shared scratch is sized at creation): replays between
    E+G+H, gradient and Hv calls match the eager energy bit for bit."""
    torch = pytest.importorskip("torch")
    m = kuhn_box(3, 3, 2, order=2)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 60)
    x = smooth_displacement(m.nodes, 0.03, 1e-3, BASE_SEED + 61)
    a = NLAssembler("NeoHookean", m, lame=(lam, mu), precision="mixed")
    dev = torch.device("cuda:0")
    s = torch.cuda.Stream(device=dev)
    a.set_stream(s.cuda_stream)
    dx = torch.tensor(x, device=dev)
    e = torch.zeros(1, dtype=torch.float64, device=dev)
    g = torch.zeros(m.n_dofs, dtype=torch.float64, device=dev)
    vals = torch.zeros(a.nnz, dtype=torch.float32, device=dev)
    y = torch.zeros(m.n_dofs, dtype=torch.float64, device=dev)
    eg = torch.zeros(1, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    a.energy_dev(dx.data_ptr(), e.data_ptr())  # warm-up outside the capture
    a.synchronize()
    ref = float(e.cpu()[0])
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        a.energy_dev(dx.data_ptr(), e.data_ptr())
    print("captured pfgpu_value_dev", flush=True)
    for k in range(4):
        e.zero_()
        # other entry points between the replays (their first calls build tables)
        a.energy_gradient_hessian_dev(dx.data_ptr(), eg.data_ptr(), g.data_ptr(), vals.data_ptr())
        a.gradient_dev(dx.data_ptr(), g.data_ptr())
        a.hvp_dev(dx.data_ptr(), g.data_ptr(), y.data_ptr())
        a.synchronize()
        with torch.cuda.stream(s):
            graph.replay()
        s.synchronize()
        assert float(e.cpu()[0]) == ref, k
    a.set_stream(0)


def indefinite_case(order=2, amp=0.15):
    """A clamped box whose start has an indefinite Hessian (P2, amplitude
    0.15: 67 negative eigenvalues of H_ff out of 741)."""
    m, prob = clamped_box(order)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 3)
    x0 = smooth_displacement(m.nodes, amp, 1e-3, BASE_SEED + 4)
    x0[prob.fixed] = prob.fixed_values
    return m, prob, lam, mu, x0


@pytest.mark.parametrize("path", ["factored", "plain"])
def test_regularized_step_solves_the_shifted_system(path):
    """On non-positive curvature the direction is solved again on
    H + delta D (D the node 3x3 diagonal blocks of H, fixed rows and columns
    zeroed); the step taken equals the sparse-direct solution of that system
    for the delta reported."""
    m, prob, lam, mu, x0 = indefinite_case()
    a = NLAssembler("NeoHookean", m, lame=(lam, mu), precision="fp64", path=path)
    x1, rep = NewtonSolver(max_iterations=1, grad_norm=0.0, cg_max_iterations=20000, cg_tolerance=1e-12,
                           line_search="None", regularization=(1e-2, 1e6)).minimize(a, x0, prob)
    print("regularized step:", rep, flush=True)
    delta = rep.last_regularization
    assert rep.regularized_solves >= 1 and delta > 0 and rep.cg_stops_tolerance == 1
    o = po.Oracle(m, lam, mu)
    H = o.hessian_scipy(x0).tocsr()
    n = m.n_nodes
    blocks = np.zeros((n, 3, 3))
    Hd = H.toarray() if m.n_dofs <= 4000 else None
    for i in range(n):
        blocks[i] = Hd[3 * i:3 * i + 3, 3 * i:3 * i + 3]
    import scipy.sparse as sp
    D = sp.block_diag(list(blocks)).tocsr()
    f = prob.free
    A = (H + delta * D).tocsr()[f][:, f]
    step = spla.spsolve(A.tocsc(), -o.gradient(x0)[f])
    # D is held in FP32: the system differs from the FP64 one by ~1e-7 delta D
    assert rel((x1 - x0)[f], step) < 1e-5
    assert np.array_equal(x1[prob.fixed], prob.fixed_values)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_regularized_newton_converges_from_an_indefinite_start(precision):
    m, prob, lam, mu, x0 = indefinite_case()
    o = po.Oracle(m, lam, mu)
    g0 = np.linalg.norm(o.gradient(x0)[prob.free])
    rtol = 1e-9 if precision == "fp64" else 1e-5
    a = NLAssembler("NeoHookean", m, lame=(lam, mu), precision=precision)
    x, rep = NewtonSolver(max_iterations=80, grad_norm=0.0, relative_grad_norm=rtol, cg_max_iterations=4000,
                          cg_tolerance=1e-8, regularization=(1e-2, 1e6)).minimize(a, x0, prob)
    print(f"regularized Newton {precision}:", rep, flush=True)
    assert rep.converged and rep.regularized_solves >= 1, rep
    assert np.linalg.norm(o.gradient(x)[prob.free]) <= (1 if precision == "fp64" else 10) * rtol * g0
    assert o.energy(x) < o.energy(x0)
