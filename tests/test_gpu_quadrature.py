"""Quadrature orders 3 and 4 (the 6- and 11-point Xiao-Gimbutas rules,
TetQuadrature.cpp:14-36 with autogen/auto_tetrahedron.ipp:13-24) on the plain
path, against the FP64 oracle on the same rule: energy, gradient, Hessian
values, Hv, material and shape VJPs for every material.  The fused and
factored paths take orders 1 and 2 only and refuse the others."""
import numpy as np
import pytest

from oracle import pyoracle as po
from oracle.fcr import FcrOracle
from oracle.mooney import MooneyOracle
from polyfem_amd import NLAssembler, PfgpuError
from polyfem_amd.mesh import BASE_SEED, kuhn_box, random_lame, smooth_displacement

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
MIXED_TOL = 2e-5
MAT = {"NeoHookean": po.NEOHOOKEAN, "LinearElasticity": po.LINEAR_ELASTICITY, "SaintVenant": po.SAINT_VENANT}


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def case(order, name, qo, precision, seed=0):
    m = kuhn_box(3, 2, 2, order=order)
    rng = np.random.default_rng(BASE_SEED + 500 + seed)
    x = smooth_displacement(m.nodes, 0.05, 1e-3, BASE_SEED + 510 + seed)
    if name == "MooneyRivlin":
        c1, c2, k = (rng.uniform(2e5, 4e5, m.n_elements), rng.uniform(5e4, 1e5, m.n_elements),
                     rng.uniform(1e6, 2e6, m.n_elements))
        a = NLAssembler(name, m, params=(c1, c2, k), precision=precision, quadrature_order=qo)
        o = MooneyOracle(m, c1, c2, k, quad_order=qo)
    else:
        lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 520 + seed, E_mean=1e6, E_sigma=0.5)
        a = NLAssembler(name, m, lame=(lam, mu), precision=precision, quadrature_order=qo)
        o = FcrOracle(m, lam, mu, quad_order=qo) if name == "FixedCorotational" else \
            po.Oracle(m, lam, mu, material=MAT[name], quad_order=qo)
    return m, x, a, o


NAMES = ["NeoHookean", "LinearElasticity", "SaintVenant", "FixedCorotational", "MooneyRivlin"]


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("order,qo", [(1, 3), (2, 3), (2, 4)])
def test_fp64_parity_higher_quadrature(order, qo, name):
    m, x, a, o = case(order, name, qo, "fp64")
    print(f"quadrature order {qo}, P{order} {name}", flush=True)
    assert a.info().n_quad == {3: 6, 4: 11}[qo]
    assert rel(a.assemble_energy(x), o.energy(x)) < FP64_TOL
    assert rel(a.assemble_gradient(x), o.gradient(x)) < FP64_TOL
    _, _, ref = o.hessian(x)
    assert rel(a.hessian_values(x), ref) < FP64_TOL
    e, g, v = a.assemble_energy_gradient_hessian(x)
    assert rel(v, ref) < FP64_TOL and rel(g, o.gradient(x)) < FP64_TOL
    rng = np.random.default_rng(BASE_SEED + 530)
    d = rng.normal(size=m.n_dofs)
    assert rel(a.hessian_vector_product(x, d), o.hessian_scipy(x) @ d) < FP64_TOL
    if name in ("NeoHookean", "LinearElasticity", "SaintVenant"):
        p = rng.normal(size=m.n_dofs)
        assert rel(a.force_material_derivative(x, p), o.material_vjp(x, p)) < FP64_TOL
        assert rel(a.force_shape_derivative(x, p), o.shape_vjp(x, p)) < FP64_TOL


@pytest.mark.parametrize("order,qo", [(1, 3), (2, 4)])
def test_mixed_parity_higher_quadrature(order, qo):
    m, x, a, o = case(order, "NeoHookean", qo, "mixed", seed=1)
    assert rel(a.assemble_energy(x), o.energy(x)) < MIXED_TOL
    assert rel(a.assemble_gradient(x), o.gradient(x)) < MIXED_TOL
    _, _, ref = o.hessian(x)
    assert rel(a.hessian_values(x), ref) < MIXED_TOL


def test_higher_quadrature_changes_the_result_and_runs_plain():
    """Orders 3 and 4 give other values than order 2 (the rule is used), the
    default path is plain, and the fused and factored paths refuse them."""
    m, x, a3, o = case(2, "NeoHookean", 3, "fp64")
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 520, E_mean=1e6, E_sigma=0.5)
    a2 = NLAssembler("NeoHookean", m, lame=(lam, mu), precision="fp64")
    assert a3.path == "plain" and a2.path == "factored"
    assert rel(a3.assemble_energy(x), a2.assemble_energy(x)) > 1e-8
    for path in ("factored", "fused"):
        with pytest.raises(PfgpuError):
            a3.set_path(path)


def test_quadrature_order_above_four_is_refused():
    m = kuhn_box(2, 2, 2, order=2)
    with pytest.raises(PfgpuError):
        NLAssembler("NeoHookean", m, young=1e6, poisson=0.3, quadrature_order=5)
