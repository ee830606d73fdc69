"""MooneyRivlin on the GPU (plain path) against the FP64 oracle restatement
(oracle/mooney.py: the literal energy of MooneyRivlinElasticity.hpp:25-44,
its closed-form stress pinned by complex step, the stiffness by complex step
of the stress).  Runs on an MI355X."""
import numpy as np
import pytest

from oracle import mooney
from polyfem_amd import NewtonSolver, NLAssembler, NLProblem, PfgpuError
from polyfem_amd.mesh import BASE_SEED, face_nodes, kuhn_box, smooth_displacement

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
# the FP32 tolerances of the other materials (test_gpu_parity.py)
MIXED_TOL = {"energy": 1e-5, "gradient": 2e-5, "hessian": 2e-5, "hvp": 2e-5, "vjp": 5e-5}


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def setup(order, precision, seed=0, grid=(3, 2, 2)):
    m = kuhn_box(*grid, order=order)
    rng = np.random.default_rng(BASE_SEED + 800 + seed)
    c1 = 2e5 * np.exp(rng.normal(0, 0.3, m.n_elements))
    c2 = 1e5 * np.exp(rng.normal(0, 0.3, m.n_elements))
    k = 2e6 * np.exp(rng.normal(0, 0.3, m.n_elements))
    x = smooth_displacement(m.nodes, 0.05, 1e-3, BASE_SEED + 810 + seed)
    a = NLAssembler("MooneyRivlin", m, params=(c1, c2, k), precision=precision)
    return m, x, a, mooney.MooneyOracle(m, c1, c2, k), (c1, c2, k)


@pytest.mark.parametrize("order", [1, 2])
def test_mooney_fp64_parity(order):
    m, x, a, o, _ = setup(order, "fp64")
    print(f"MooneyRivlin fp64 P{order} E={m.n_elements}", flush=True)
    assert a.path == "plain"
    assert rel(a.assemble_energy(x), o.energy(x)) < FP64_TOL
    assert rel(a.assemble_energy_per_element(x), o.energy_per_element(x)) < FP64_TOL
    assert rel(a.assemble_gradient(x), o.gradient(x)) < FP64_TOL
    outer, inner, ref = o.hessian(x)
    go, gi = a.hessian_pattern()
    assert np.array_equal(go, outer) and np.array_equal(gi, inner)
    assert rel(a.hessian_values(x), ref) < FP64_TOL
    e, g, vals = a.assemble_energy_gradient_hessian(x)
    assert rel(g, o.gradient(x)) < FP64_TOL and rel(vals, ref) < FP64_TOL
    rng = np.random.default_rng(BASE_SEED + 820)
    v = rng.normal(size=m.n_dofs)
    assert rel(a.hessian_vector_product(x, v), o.hessian_scipy(x) @ v) < FP64_TOL
    p = rng.normal(size=m.n_dofs)
    assert rel(a.force_shape_derivative(x, p), o.shape_vjp(x, p)) < FP64_TOL


@pytest.mark.parametrize("order", [1, 2])
def test_mooney_mixed_parity(order):
    m, x, a, o, _ = setup(order, "mixed", seed=1)
    print(f"MooneyRivlin mixed P{order}", flush=True)
    assert rel(a.assemble_energy(x), o.energy(x)) < MIXED_TOL["energy"]
    assert rel(a.assemble_gradient(x), o.gradient(x)) < MIXED_TOL["gradient"]
    outer, inner, ref = o.hessian(x)
    assert rel(a.hessian_values(x), ref) < MIXED_TOL["hessian"]
    v = np.random.default_rng(BASE_SEED + 821).normal(size=m.n_dofs)
    assert rel(a.hessian_vector_product(x, v), o.hessian_scipy(x) @ v) < MIXED_TOL["hvp"]
    p = np.random.default_rng(BASE_SEED + 822).normal(size=m.n_dofs)
    assert rel(a.force_shape_derivative(x, p), o.shape_vjp(x, p)) < MIXED_TOL["vjp"]


def test_mooney_parameter_update_and_unsupported_calls():
    m, x, a, o, (c1, c2, k) = setup(1, "fp64", seed=2)
    a.update_mooney_rivlin_params(c1 * 1.5, c2 * 0.5, k * 2.0)
    o2 = mooney.MooneyOracle(m, c1 * 1.5, c2 * 0.5, k * 2.0)
    assert rel(a.assemble_gradient(x), o2.gradient(x)) < FP64_TOL
    p = np.ones(m.n_dofs)
    with pytest.raises(PfgpuError) as ei:
        a.force_material_derivative(x, p)
    assert ei.value.status == 4
    with pytest.raises(PfgpuError) as ei:
        a.set_path("factored")
    assert ei.value.status == 4
    with pytest.raises(PfgpuError):
        a.update_lame_params(np.ones(m.n_elements), np.ones(m.n_elements))


@pytest.mark.parametrize("order", [1, 2])
def test_mooney_newton_solve(order):
    """A stretched bar: Newton-CG converges and the oracle gradient vanishes
    on the free dofs at the solution."""
    m, x, a, o, _ = setup(order, "fp64", seed=3)
    bot, top = face_nodes(m, 2, 0.0), face_nodes(m, 2, 1.0)
    fixed = np.concatenate([3 * bot[:, None] + np.arange(3), 3 * top[:, None] + np.arange(3)]).ravel()
    vals = np.concatenate([np.zeros(3 * bot.size), np.tile([0.0, 0.0, 0.08], top.size)])
    prob = NLProblem(m.n_dofs, fixed, vals)
    x0 = np.zeros(m.n_dofs)
    x0[2::3] = 0.08 * m.nodes[:, 2]
    x0[prob.fixed] = prob.fixed_values
    g0 = np.linalg.norm(o.gradient(x0)[prob.free])
    newton = NewtonSolver(max_iterations=30, grad_norm=1e-10 * g0, cg_max_iterations=5000, cg_tolerance=1e-12)
    print(f"MooneyRivlin Newton P{order}", flush=True)
    u, rep = newton.minimize(a, x0, prob)
    assert rep.converged, rep
    assert np.linalg.norm(o.gradient(u)[prob.free]) <= 1e-8 * g0
