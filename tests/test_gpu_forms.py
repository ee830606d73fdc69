"""GPU tests of the forms around the assembly (runs on an MI355X).

* ``ElasticForm`` (ElasticForm.cpp:287-328, 388-418; Form.hpp:26-56):
  value and derivatives scaled by weight / scale, the linear-elastic
  stiffness assembled once and cached, ``is_step_valid`` false on an
  inverted element.
* The body load (BodyForm.cpp: the gradient of the body form is
  -current_rhs): a gravity load f = M g from the library's mass product,
  balanced at the solution: the oracle gradient at x* equals f on the free
  dofs; the load's total equals rho |V| g (closed form).
* Linear elasticity with a load: one Newton step solves K u = f, checked
  against a sparse direct solve with the oracle's K.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from polyfem_amd import ElasticForm, NewtonSolver, NLAssembler, NLProblem, TransientSimulation
from polyfem_amd.mesh import BASE_SEED, face_nodes, kuhn_box, random_lame, smooth_displacement

pytestmark = pytest.mark.gpu

MAT = {"NeoHookean": po.NEOHOOKEAN, "LinearElasticity": po.LINEAR_ELASTICITY}


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("name", ["NeoHookean", "LinearElasticity"])
def test_elastic_form_weight_scale(name, order):
    m = kuhn_box(3, 2, 2, order=order)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 700)
    x = smooth_displacement(m.nodes, 0.05, 1e-3, BASE_SEED + 701)
    a = NLAssembler(name, m, lame=(lam, mu), precision="fp64")
    o = po.Oracle(m, lam, mu, material=MAT[name])
    print(f"ElasticForm {name} P{order}", flush=True)
    # weight dt^2 as in a transient solve (SolveData.cpp:479-492), scale from normalize_forms
    w, s = 1e-3 ** 2, 4.0
    form = ElasticForm(a, weight=w, scale=s)
    assert form.weight() == w
    c = w / s
    assert rel(form.value(x), c * o.energy(x)) < 1e-10
    assert rel(form.value_unweighted(x), o.energy(x)) < 1e-10
    assert rel(form.value_per_element_unweighted(x), o.energy_per_element(x)) < 1e-10
    assert rel(form.first_derivative(x), c * o.gradient(x)) < 1e-10
    H = o.hessian_scipy(x)
    Hf = form.second_derivative(x)
    assert rel(Hf.toarray(), c * H.toarray()) < 1e-10
    form.set_weight(2.0)
    assert rel(form.first_derivative(x), 0.5 * o.gradient(x)) < 1e-10


def test_elastic_form_caches_linear_stiffness():
    """compute_cached_stiffness (ElasticForm.cpp:411-418): the linear-elastic
    Hessian is assembled once and reused at every x; it is the oracle's K."""
    m = kuhn_box(3, 3, 2, order=2)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 702)
    a = NLAssembler("LinearElasticity", m, lame=(lam, mu), precision="fp64")
    o = po.Oracle(m, lam, mu, material=po.LINEAR_ELASTICITY)
    form = ElasticForm(a)
    x1 = smooth_displacement(m.nodes, 0.05, 1e-3, BASE_SEED + 703)
    x2 = smooth_displacement(m.nodes, 0.02, 1e-3, BASE_SEED + 704)
    K1 = form.second_derivative_unweighted(x1)
    K2 = form.second_derivative_unweighted(x2)
    assert K1 is K2
    assert rel(K1.toarray(), o.hessian_scipy(np.zeros(m.n_dofs)).toarray()) < 1e-10
    # a NeoHookean form is not cached: its Hessian follows x
    nh = ElasticForm(NLAssembler("NeoHookean", m, lame=(lam, mu), precision="fp64"))
    onh = po.Oracle(m, lam, mu)
    assert rel(nh.second_derivative_unweighted(x2).toarray(), onh.hessian_scipy(x2).toarray()) < 1e-10
    assert rel(nh.second_derivative_unweighted(x1).toarray(), onh.hessian_scipy(x1).toarray()) < 1e-10


def test_elastic_form_step_validity():
    """is_step_valid (ElasticForm.cpp:388-405): a step that inverts an
    element gives a non-finite gradient and is rejected."""
    m = kuhn_box(2, 2, 2, order=1)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 705)
    a = NLAssembler("NeoHookean", m, lame=(lam, mu), precision="fp64")
    form = ElasticForm(a)
    x0 = smooth_displacement(m.nodes, 0.02, 1e-4, BASE_SEED + 706)
    assert form.is_step_valid(x0, x0)
    x1 = x0.copy()
    n0 = m.elem_nodes[0]
    # vertex 3 of element 0 reflected through vertex 0: the element inverts
    x1.reshape(-1, 3)[n0[3]] = 2 * (m.nodes[n0[0]] - m.nodes[n0[3]])
    assert not form.is_step_valid(x0, x1)


@pytest.mark.parametrize("order", [1, 2])
def test_gravity_load_equilibrium(order):
    """Cantilever under gravity: the Newton solution balances the load."""
    m = kuhn_box(4, 2, 2, order=order, size=(2.0, 0.5, 0.5))
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 710, E_mean=1e5, E_sigma=0.3)
    a = NLAssembler("NeoHookean", m, lame=(lam, mu), precision="fp64")
    o = po.Oracle(m, lam, mu)
    rho, gz = 1000.0, -9.81
    gvec = np.zeros(m.n_dofs)
    gvec[2::3] = gz
    f = TransientSimulation(a, rho=rho).mass_matvec(gvec)
    # the consistent mass integrates rho phi_i: the load totals rho |V| g
    vol = 2.0 * 0.5 * 0.5
    assert abs(f[2::3].sum() - rho * vol * gz) <= 1e-12 * abs(rho * vol * gz)
    assert np.all(f[0::3] == 0) and np.all(f[1::3] == 0)
    clamp = face_nodes(m, 0, 0.0)
    prob = NLProblem(m.n_dofs, (3 * clamp[:, None] + np.arange(3)).ravel())
    a.set_load(f)
    print(f"gravity P{order} dofs={m.n_dofs}", flush=True)
    newton = NewtonSolver(max_iterations=30, grad_norm=1e-9 * np.abs(f).max(), cg_max_iterations=5000,
                          cg_tolerance=1e-12)
    u, rep = newton.minimize(a, np.zeros(m.n_dofs), prob)
    assert rep.converged, rep
    r = o.gradient(u) - f
    assert np.abs(r[prob.free]).max() <= 1e-8 * np.abs(f).max()
    # the free end sags
    tip = face_nodes(m, 0, 2.0)
    assert np.all(u[3 * tip + 2] < 0)
    # without the load the clamped bar stays at rest
    a.set_load(None)
    u0, rep0 = newton.minimize(a, np.zeros(m.n_dofs), prob)
    assert np.abs(u0).max() == 0.0


def test_linear_elasticity_load_solve():
    """Quadratic energy: the solve of E(u) - f.u is K_ff u_f = f_f - K_fc u_c."""
    import scipy.sparse.linalg as spla
    m = kuhn_box(3, 2, 2, order=2)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 720)
    a = NLAssembler("LinearElasticity", m, lame=(lam, mu), precision="fp64")
    o = po.Oracle(m, lam, mu, material=po.LINEAR_ELASTICITY)
    rng = np.random.default_rng(BASE_SEED + 721)
    f = rng.normal(size=m.n_dofs) * 1e3
    bot, top = face_nodes(m, 2, 0.0), face_nodes(m, 2, 1.0)
    fixed = np.concatenate([3 * bot[:, None] + np.arange(3), 3 * top[:, None] + np.arange(3)]).ravel()
    vals = np.concatenate([np.zeros(3 * bot.size), np.tile([0.0, 0.01, 0.05], top.size)])
    prob = NLProblem(m.n_dofs, fixed, vals)
    a.set_load(f)
    newton = NewtonSolver(max_iterations=5, grad_norm=1e-9 * np.abs(f).max(), cg_max_iterations=5000,
                          cg_tolerance=1e-13)
    u, rep = newton.minimize(a, np.zeros(m.n_dofs), prob)
    assert rep.converged, rep
    K = o.hessian_scipy(np.zeros(m.n_dofs)).tocsr()
    fr, fx = prob.free, prob.fixed
    Kff = K[fr][:, fr].tocsc()
    Kfc = K[fr][:, fx]
    uf = spla.spsolve(Kff, f[fr] - Kfc @ prob.fixed_values)
    assert rel(u[fr], uf) < 1e-8
    assert np.array_equal(u[fx], prob.fixed_values)
