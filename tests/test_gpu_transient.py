"""GPU tests of the transient pieces: consistent mass matrix, implicit Euler
and its checkpointed adjoint (runs on an MI355X).

The mass matrix is checked against an independent numerical integration
(a collapsed-coordinate Gauss rule of degree 8 on each element, with the
oracle's basis values); the adjoint gradients against central finite
differences of the objective through full forward runs, and the
checkpointed backward pass against the store-every-state one (bitwise).
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from polyfem_amd import NewtonSolver, NLAssembler, NLProblem, TransientSimulation
from polyfem_amd.mesh import BASE_SEED, face_nodes, kuhn_box, random_lame

pytestmark = pytest.mark.gpu


def tet_rule(n=6):
    """Collapsed (Duffy) Gauss-Legendre rule on the reference tet, exact to degree 2n-3."""
    g, w = np.polynomial.legendre.leggauss(n)
    g = 0.5 * (g + 1)
    w = 0.5 * w
    pts, wts = [], []
    for i in range(n):
        for j in range(n):
            for k in range(n):
                u, v, t = g[i], g[j], g[k]
                pts.append([u, v * (1 - u), t * (1 - u) * (1 - v)])
                wts.append(w[i] * w[j] * w[k] * (1 - u) ** 2 * (1 - v))
    return np.array(pts), np.array(wts)


def numpy_mass(m, rho):
    pts, wts = tet_rule()
    po.build()
    phi = np.array([po.basis_val(m.order, p) for p in pts])  # (n_points, n_local)
    R = (phi * wts[:, None]).T @ phi  # reference-tet integrals (volume 1/6)
    import scipy.sparse as sp
    X = m.vertices[m.elem_vertices]
    det = np.linalg.det(X[:, 1:] - X[:, :1])
    en = m.elem_nodes
    nb = en.shape[1]
    vals = (rho * det)[:, None, None] * R[None]
    rows = np.repeat(en, nb, axis=1).ravel()
    cols = np.tile(en, (1, nb)).ravel()
    return sp.csr_matrix((vals.ravel(), (rows, cols)), shape=(m.n_nodes, m.n_nodes))


@pytest.mark.parametrize("order", [1, 2])
def test_lumped_mass_is_row_sums(order):
    """lump_mass: diag(row sums of the consistent mass), utils::lump_matrix
    (MatrixUtils.cpp:84-101), against the numerically integrated matrix; the
    total mass is kept; switching back restores the consistent matrix."""
    m = kuhn_box(3, 2, 2, order=order)
    rho = np.random.default_rng(BASE_SEED + 22).uniform(500, 1500, size=m.n_elements)
    a = NLAssembler("NeoHookean", m, young=1e6, poisson=0.3)
    print(f"lumped mass order={order}", flush=True)
    sim = TransientSimulation(a, rho=rho, lump_mass=True)
    v = np.random.default_rng(BASE_SEED + 23).normal(size=m.n_dofs)
    Ms = numpy_mass(m, rho)
    d = np.asarray(Ms.sum(axis=1)).ravel()
    ref = (d[:, None] * v.reshape(-1, 3)).ravel()
    y = sim.mass_matvec(v)
    assert np.max(np.abs(y - ref)) <= 1e-12 * np.max(np.abs(ref))
    ones = np.ones(m.n_dofs)
    X = m.vertices[m.elem_vertices]
    vol = np.abs(np.linalg.det(X[:, 1:] - X[:, :1])) / 6
    assert abs(sim.mass_matvec(ones).reshape(-1, 3)[:, 0].sum() - rho @ vol) <= 1e-12 * (rho @ vol)
    check_lib = __import__("polyfem_amd._ffi", fromlist=["lib"])
    check_lib.check(check_lib.lib().pfgpu_set_mass_lumping(a._h, 0))
    full = (Ms @ v.reshape(-1, 3)).ravel()
    assert np.max(np.abs(sim.mass_matvec(v) - full)) <= 1e-12 * np.max(np.abs(full))


@pytest.mark.parametrize("order", [1, 2])
def test_hrz_lumped_mass(order):
    """HRZ lumping (HRZMass, Mass.cpp:48-86): per element the diagonal of the
    numerically integrated element mass, scaled to keep the element's total;
    positive at P2 (where row sums are negative at the vertices); the total
    mass is kept."""
    m = kuhn_box(3, 2, 2, order=order)
    rho = np.random.default_rng(BASE_SEED + 24).uniform(500, 1500, size=m.n_elements)
    a = NLAssembler("NeoHookean", m, young=1e6, poisson=0.3)
    print(f"HRZ mass order={order}", flush=True)
    sim = TransientSimulation(a, rho=rho, lump_mass="hrz")
    pts, wts = tet_rule()
    phi = np.array([po.basis_val(m.order, p) for p in pts])
    R = (phi * wts[:, None]).T @ phi
    X = m.vertices[m.elem_vertices]
    det = np.linalg.det(X[:, 1:] - X[:, :1])
    d = np.zeros(m.n_nodes)
    for e in range(m.n_elements):
        Me = rho[e] * det[e] * R
        np.add.at(d, m.elem_nodes[e], np.diag(Me) * Me.sum() / np.trace(Me))
    v = np.random.default_rng(BASE_SEED + 25).normal(size=m.n_dofs)
    ref = (d[:, None] * v.reshape(-1, 3)).ravel()
    y = sim.mass_matvec(v)
    assert np.max(np.abs(y - ref)) <= 1e-12 * np.max(np.abs(ref))
    diag = sim.mass_matvec(np.ones(m.n_dofs)).reshape(-1, 3)[:, 0]
    assert np.all(diag > 0)
    vol = np.abs(det) / 6
    assert abs(diag.sum() - rho @ vol) <= 1e-12 * (rho @ vol)


@pytest.mark.parametrize("order", [1, 2])
def test_mass_matches_numerical_integration(order):
    m = kuhn_box(3, 2, 2, order=order)
    rho = np.random.default_rng(BASE_SEED + 20).uniform(500, 1500, size=m.n_elements)
    a = NLAssembler("NeoHookean", m, young=1e6, poisson=0.3)
    sim = TransientSimulation(a, rho=rho)
    v = np.random.default_rng(BASE_SEED + 21).normal(size=m.n_dofs)
    Ms = numpy_mass(m, rho)
    ref = (Ms @ v.reshape(-1, 3)).ravel()
    y = sim.mass_matvec(v)
    assert np.max(np.abs(y - ref)) <= 1e-12 * np.max(np.abs(ref))
    # total mass: 1^T M 1 = sum rho |T|
    X = m.vertices[m.elem_vertices]
    vol = np.linalg.det(X[:, 1:] - X[:, :1]) / 6
    ones = np.tile([1.0, 0, 0], m.n_nodes)
    assert abs(ones @ sim.mass_matvec(ones) - rho @ vol) <= 1e-12 * (rho @ vol)


def setup(every, lam=None, mu=None, material="NeoHookean", lump=False):
    m = kuhn_box(2, 2, 2, order=1)
    if lam is None:
        lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 22)
    bot = face_nodes(m, 2, 0.0)
    prob = NLProblem(m.n_dofs, (3 * bot[:, None] + np.arange(3)).ravel())
    a = NLAssembler(material, m, lame=(lam, mu), precision="fp64")
    newton = NewtonSolver(max_iterations=50, grad_norm=1e-13, cg_max_iterations=5000, cg_tolerance=1e-14)
    sim = TransientSimulation(a, rho=1000.0, dt=2e-3, time_steps=6, checkpoint_every=every, newton=newton,
                              lump_mass=lump)
    # seeded initial velocity (the sine field of the box configurations
    # vanishes at every node of a 2x2x2 grid)
    X = m.nodes
    v0 = 0.5 * np.stack([np.sin(np.pi * X[:, 2]) * np.cos(np.pi * X[:, 1]),
                         np.sin(np.pi * X[:, 2]) * X[:, 0],
                         0.3 * np.sin(np.pi * X[:, 2])], axis=1).ravel()
    v0 += 0.05 * np.random.default_rng(BASE_SEED + 23).normal(size=v0.size)
    v0[prob.fixed] = 0.0
    return m, prob, a, sim, np.zeros(m.n_dofs), v0, lam, mu


def objective(a, sim, xT, vT):
    """J = 1/2 v_T^T M v_T + E_el(x_T) and its partials"""
    MvT = sim.mass_matvec(vT)
    return 0.5 * vT @ MvT + a.assemble_energy(xT), a.assemble_gradient(xT), MvT


def test_checkpointed_adjoint_equals_store_all():
    out = []
    for every in (1, 2, 4):
        m, prob, a, sim, x0, v0, _, _ = setup(every)
        print(f"transient checkpoint_every={every}", flush=True)
        xT, vT = sim.forward(x0, v0, prob)
        J, gx, gv = objective(a, sim, xT, vT)
        assert J > 0 and np.abs(vT).max() > 0.05
        out.append(sim.adjoint(gx, gv))
        assert np.abs(out[-1][0]).max() > 0 and np.abs(out[-1][2]).max() > 0
    for k in range(1, len(out)):
        for u, w in zip(out[0], out[k]):
            assert np.array_equal(u, w)


@pytest.mark.parametrize("lump", [False, True])
def test_transient_adjoint_matches_finite_differences(lump):
    m, prob, a, sim, x0, v0, lam, mu = setup(2, lump=lump)
    print(f"transient adjoint FD lump_mass={lump}", flush=True)
    xT, vT = sim.forward(x0, v0, prob)
    J0, gx, gv = objective(a, sim, xT, vT)
    dth, dx0, dv0 = sim.adjoint(gx, gv)
    assert np.all(dv0[prob.fixed] == 0) and np.all(dx0[prob.fixed] == 0)
    assert J0 > 0 and np.abs(dv0).max() > 0 and np.abs(dth).max() > 0
    rng = np.random.default_rng(BASE_SEED + 24)

    def J_of(lm, mm, vv):
        _, _, a2, s2, _, _, _, _ = setup(2, lm, mm, lump=lump)
        xx, v2 = s2.forward(x0, vv, prob)
        return objective(a2, s2, xx, v2)[0]

    h = 1e-5
    # material direction
    dl = rng.normal(size=m.n_elements) * lam
    dm = rng.normal(size=m.n_elements) * mu
    fd = (J_of(lam + h * dl, mu + h * dm, v0) - J_of(lam - h * dl, mu - h * dm, v0)) / (2 * h)
    # J contains E_el(x_T): its explicit Lame derivative adds to the adjoint terms
    tot = dth + a.energy_lame_gradient(xT)
    lin = tot[:m.n_elements] @ dl + tot[m.n_elements:] @ dm
    assert abs(fd - lin) <= 1e-5 * abs(lin), (fd, lin)
    # initial velocity direction (free dofs)
    dv = rng.normal(size=m.n_dofs)
    dv[prob.fixed] = 0
    fd = (J_of(lam, mu, v0 + h * dv) - J_of(lam, mu, v0 - h * dv)) / (2 * h)
    assert abs(fd - dv0 @ dv) <= 1e-5 * abs(dv0 @ dv), (fd, dv0 @ dv)


@pytest.mark.parametrize("name", ["NeoHookean", "LinearElasticity"])
def test_p2_matrix_free_preconditioner_with_inertia(name):
    """P2 implicit Euler with a matrix-free operator and the block-Jacobi
    preconditioner: on the factored path the diagonal blocks come from the
    factor records with the mass and the dt^2 weight added (no assembly), on
    the plain path from the assembled values.  Both give the same end state."""
    m = kuhn_box(2, 2, 3, order=2)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 80)
    bot = face_nodes(m, 2, 0.0)
    fixed = (3 * bot[:, None] + np.arange(3)).ravel()
    prob = NLProblem(m.n_dofs, fixed, np.zeros(fixed.size))
    X = m.nodes
    v0 = np.stack([0.4 * X[:, 2] ** 2, -0.2 * X[:, 2], 0.1 * np.sin(np.pi * X[:, 0])], axis=1).ravel()
    v0[prob.fixed] = 0.0
    out = {}
    for path in ("plain", "factored"):
        a = NLAssembler(name, m, lame=(lam, mu), precision="fp64", path=path)
        a.set_solver_options("matrix_free", "block_jacobi")
        sim = TransientSimulation(a, rho=1000.0, dt=2e-3, time_steps=4, checkpoint_every=2)
        print(f"P2 transient {name} {path}", flush=True)
        out[path] = sim.forward(np.zeros(m.n_dofs), v0, prob)
    for u, w in zip(out["plain"], out["factored"]):
        assert np.abs(u).max() > 0
        assert np.max(np.abs(u - w)) <= 1e-8 * np.max(np.abs(u))
