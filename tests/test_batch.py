"""Batched solves of independent problems (many meshes of uneven size).

CPU: the concatenation of meshes keeps each problem's nodes and elements
together.  GPU: a batch solved in one handle per (order, material) group gives
each problem the same minimiser as solving it alone.
"""
import numpy as np
import pytest

from polyfem_amd.mesh import BASE_SEED, face_nodes, kuhn_box, random_lame
from polyfem_amd.solver import concatenate_meshes


def make_problems(seed=BASE_SEED + 40):
    rng = np.random.default_rng(seed)
    grids = [(2, 2, 2), (3, 2, 2), (2, 3, 3), (3, 3, 2), (2, 2, 4), (4, 2, 2)]
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


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_matches_individual_solves(precision):
    from polyfem_amd import BatchSolver, NewtonSolver, NLAssembler, NLProblem
    probs = make_problems()
    g0 = []
    for p in probs:
        a = NLAssembler(p["material"], p["mesh"], lame=p["lame"], precision="fp64")
        pr = NLProblem(p["mesh"].n_dofs, p["fixed_dofs"], p["fixed_values"])
        x0 = p["x0"].copy()
        x0[pr.fixed] = pr.fixed_values
        g0.append(np.linalg.norm(a.assemble_gradient(x0)[pr.free]))
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
    g0 = []
    for p in probs:
        a = NLAssembler(p["material"], p["mesh"], lame=p["lame"], precision="fp64")
        pr = NLProblem(p["mesh"].n_dofs, p["fixed_dofs"], p["fixed_values"])
        x0 = p["x0"].copy()
        x0[pr.fixed] = pr.fixed_values
        g0.append(np.linalg.norm(a.assemble_gradient(x0)[pr.free]))
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
