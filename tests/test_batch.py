"""Batched solves of independent problems (many meshes of uneven size, mixed
order and material) in one batch handle (pfgpu_batch_create /
pfgpu_batch_newton).

CPU: the concatenation of meshes keeps each problem's nodes and elements
together.  GPU: against the FP64 oracle, every problem's reduced gradient at
the returned x is below the tolerance; one batched Newton step equals each
problem's sparse-direct Newton step (Dirichlet dofs removed,
NLProblem.cpp:583-629); the batch equals the problems solved one by one.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from polyfem_amd.mesh import BASE_SEED, face_nodes, kuhn_box, random_lame
from polyfem_amd.solver import concatenate_meshes

MAT = {"NeoHookean": po.NEOHOOKEAN, "LinearElasticity": po.LINEAR_ELASTICITY}


def make_problems(seed=BASE_SEED + 40, grids=None):
    rng = np.random.default_rng(seed)
    grids = grids or [(2, 2, 2), (3, 2, 2), (2, 3, 3), (3, 3, 2), (2, 2, 4), (4, 2, 2)]
    probs = []
    for i, g in enumerate(grids):
        order = 1 if i % 2 == 0 else 2
        material = "LinearElasticity" if (i // 2) % 2 == 0 else "NeoHookean"
        m = kuhn_box(*g, order=order)
        lam, mu = random_lame(m.n_elements, seed=seed + i)
        bot, top = face_nodes(m, 2, 0.0), face_nodes(m, 2, 1.0)
        pull = rng.uniform(0.02, 0.08)
        fixed = np.concatenate([3 * bot[:, None] + np.arange(3), 3 * top[:, None] + np.arange(3)]).ravel()
        vals = np.concatenate([np.zeros(3 * bot.size), np.tile([0.0, 0.0, pull], top.size)])
        x0 = np.zeros(m.n_dofs)
        x0[2::3] = pull * m.nodes[:, 2]
        probs.append(dict(mesh=m, material=material, lame=(lam, mu), fixed_dofs=fixed, fixed_values=vals, x0=x0))
    return probs


def test_concatenate_meshes():
    ms = [kuhn_box(2, 2, 2, order=2), kuhn_box(3, 2, 1, order=2)]
    big, eo, no = concatenate_meshes(ms)
    assert list(eo) == [0, ms[0].n_elements, ms[0].n_elements + ms[1].n_elements]
    assert list(no) == [0, ms[0].n_nodes, ms[0].n_nodes + ms[1].n_nodes]
    for s, m in enumerate(ms):
        en = big.elem_nodes[eo[s]:eo[s + 1]]
        assert en.min() >= no[s] and en.max() < no[s + 1]
        assert np.array_equal(en - no[s], m.elem_nodes)
        assert np.array_equal(big.nodes[no[s]:no[s + 1]], m.nodes)


def reduced_grad0(p):
    from polyfem_amd import NLProblem
    o = po.Oracle(p["mesh"], p["lame"][0], p["lame"][1], material=MAT[p["material"]])
    pr = NLProblem(p["mesh"].n_dofs, p["fixed_dofs"], p["fixed_values"])
    x0 = p["x0"].copy()
    x0[pr.fixed] = pr.fixed_values
    return o, pr, x0, np.linalg.norm(o.gradient(x0)[pr.free])


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_mixed_batch_against_the_oracle(precision):
    """P1/P2 x NH/LE problems of uneven size, one of more than 10k tets; at
    the batch's solution every problem's oracle gradient on the free dofs is
    below the tolerance relative to its initial gradient."""
    from polyfem_amd import BatchSolver, NewtonSolver
    probs = make_problems(seed=BASE_SEED + 45, grids=[(2, 2, 2), (3, 2, 2), (12, 12, 12), (3, 3, 2), (2, 2, 4),
                                                      (4, 2, 2), (5, 3, 2)])
    assert max(p["mesh"].n_elements for p in probs) >= 10000
    refs = [reduced_grad0(p) for p in probs]
    rtol = 1e-9 if precision == "fp64" else 1e-5
    newton = NewtonSolver(max_iterations=40, grad_norm=0.0, relative_grad_norm=rtol, cg_max_iterations=5000,
                          cg_tolerance=1e-12 if precision == "fp64" else 1e-7)
    print(f"mixed batch of {len(probs)} problems, {precision}", flush=True)
    xs, reps = BatchSolver(probs, precision=precision, newton=newton).solve()
    for p, x, rep, (o, pr, x0, g0) in zip(probs, xs, reps, refs):
        assert rep.converged, (p["material"], p["mesh"].order, rep)
        assert np.array_equal(x[pr.fixed], pr.fixed_values)
        # the oracle's gradient at the batch's solution (FP64); in mixed
        # precision the solution is as accurate as the FP32 gradient allows
        gtol = 1e-7 if precision == "fp64" else 1e-4
        assert np.linalg.norm(o.gradient(x)[pr.free]) <= gtol * g0, (p["material"], p["mesh"].order)


@pytest.mark.gpu
def test_one_batched_newton_step_is_the_sparse_direct_step():
    """One Newton iteration without line search on every problem at once:
    each problem's step is -H_ff^{-1} g_f of its oracle Hessian."""
    import scipy.sparse.linalg as spla
    from polyfem_amd import BatchSolver, NewtonSolver
    probs = make_problems(seed=BASE_SEED + 46)
    refs = [reduced_grad0(p) for p in probs]
    newton = NewtonSolver(max_iterations=1, grad_norm=0.0, cg_max_iterations=20000, cg_tolerance=1e-14,
                          line_search="None")
    xs, reps = BatchSolver(probs, precision="fp64", newton=newton).solve()
    for p, x, (o, pr, x0, g0) in zip(probs, xs, refs):
        H = o.hessian_scipy(x0).tocsr()
        g = o.gradient(x0)
        f = pr.free
        step = spla.spsolve(H[f][:, f].tocsc(), -g[f])
        assert np.abs((x - x0)[f] - step).max() <= 1e-8 * np.abs(step).max(), (p["material"], p["mesh"].order)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_matches_individual_solves(precision):
    from polyfem_amd import BatchSolver, NewtonSolver, NLAssembler, NLProblem
    probs = make_problems()
    g0 = [reduced_grad0(p)[3] for p in probs]
    rtol = 1e-10 if precision == "fp64" else 1e-5
    newton = NewtonSolver(max_iterations=30, grad_norm=rtol * min(g0), cg_max_iterations=5000,
                          cg_tolerance=1e-12 if precision == "fp64" else 1e-7)
    print(f"batch of {len(probs)} problems, {precision}", flush=True)
    xs, reps = BatchSolver(probs, precision=precision, newton=newton).solve()
    for p, x, rep in zip(probs, xs, reps):
        a = NLAssembler(p["material"], p["mesh"], lame=p["lame"], precision=precision)
        pr = NLProblem(p["mesh"].n_dofs, p["fixed_dofs"], p["fixed_values"])
        xi, ri = newton.minimize(a, p["x0"], pr)
        assert rep.converged and ri.converged, (rep, ri)
        tol = 1e-7 if precision == "fp64" else 1e-3
        assert np.max(np.abs(x - xi)) <= tol * np.max(np.abs(xi)), (p["material"], p["mesh"].order)
        assert np.array_equal(x[pr.fixed], pr.fixed_values)


@pytest.mark.gpu
def test_batch_problems_of_different_stiffness():
    """Stiffness spread of 100x across the batch: each problem's CG runs to its
    own relative tolerance, so every problem takes the Newton iterations it
    takes alone and reaches the same minimiser."""
    from polyfem_amd import BatchSolver, NewtonSolver, NLAssembler, NLProblem
    probs = make_problems(seed=BASE_SEED + 50)
    for i, p in enumerate(probs):
        s = 10.0 ** (2.0 * i / (len(probs) - 1))
        p["lame"] = (p["lame"][0] * s, p["lame"][1] * s)
    g0 = [reduced_grad0(p)[3] for p in probs]
    newton = NewtonSolver(max_iterations=30, grad_norm=1e-8 * min(g0), cg_max_iterations=5000, cg_tolerance=1e-10)
    print(f"batch of {len(probs)} problems, stiffness spread 100x", flush=True)
    xs, reps = BatchSolver(probs, precision="fp64", newton=newton).solve()
    for p, x, rep in zip(probs, xs, reps):
        a = NLAssembler(p["material"], p["mesh"], lame=p["lame"], precision="fp64")
        pr = NLProblem(p["mesh"].n_dofs, p["fixed_dofs"], p["fixed_values"])
        xi, ri = newton.minimize(a, p["x0"], pr)
        assert rep.converged and ri.converged, (rep, ri)
        assert abs(rep.iterations - ri.iterations) <= 1, (rep, ri)
        assert np.max(np.abs(x - xi)) <= 1e-7 * np.max(np.abs(xi))
