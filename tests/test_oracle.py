"""Pins the FP64 oracle (CPU, no GPU needed).

Checks follow the reference's own tests where they exist and the mathematics
where they do not:

* SparseMatrixCache known-answer test (tests/test_matrix.cpp:202-249);
* P1/P2 bases against an independent barycentric implementation at
  order-8-like sample points (tests/test_bases.cpp:799-860, margin 1e-10);
* tet quadrature: weights, exactness for polynomials of degree <= 2
  (tests/test_quadrature.cpp:165-230);
* NeoHookean hand-coded derivatives against an independent derivative of the
  energy (complex-step differentiation, exact to round-off) at margin 1e-12,
  the analogue of "generic_elastic_assembler" (tests/test_assembler.cpp:127-286);
* NH Hessian at x = 0 equals the linear-elastic stiffness ("hessian_lin",
  tests/test_assembler.cpp:17-70);
* linear elasticity: g = K x, E = x^T K x / 2;
* central finite differences of energy/gradient (test_form_derivatives.cpp:330-344);
* rigid motions: zero energy and gradient; frame indifference;
* adjoint terms against central differences (test_diff.cpp:119-157).
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from polyfem_amd.mesh import BASE_SEED, face_nodes, kuhn_box, lame_from_young, smooth_displacement

LAM, MU = lame_from_young(2e4, 0.3)  # test_form_derivatives.cpp:41-165


@pytest.fixture(scope="module", autouse=True)
def _built():
    po.build()


# ------------------------------------------------------------------- KAT

def test_sparse_matrix_cache_known_answer():
    """tests/test_matrix.cpp:202-249: (0,0)=1, (0,1)=2, (9,4)=3+3, (9,9)=4 on a 10x10 matrix."""
    rows = [0, 0, 9, 9, 9]
    cols = [0, 1, 4, 4, 9]
    vals = [1, 2, 3, 3, 4]
    outer, inner, values = po.triplets_to_csc(10, rows, cols, vals)
    import scipy.sparse as sp
    M = sp.csc_matrix((values, inner, outer), shape=(10, 10))
    assert M[0, 0] == 1 and M[0, 1] == 2 and M[9, 4] == 6 and M[9, 9] == 4
    assert M.nnz == 4
    # inner indices sorted within each column
    for c in range(10):
        seg = inner[outer[c]:outer[c + 1]]
        assert np.all(np.diff(seg) > 0)


def test_csc_keeps_explicit_zeros():
    outer, inner, values = po.triplets_to_csc(3, [0, 0, 2], [0, 0, 1], [1.0, -1.0, 0.0])
    assert len(inner) == 2 and values[0] == 0.0 and values[1] == 0.0


# --------------------------------------------------------------- bases

def _bary_basis(order, p):
    x, y, z = p
    lam = np.array([1 - x - y - z, x, y, z])
    dlam = np.array([[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=float)
    if order == 1:
        return lam, dlam
    edges = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
    v = [lam[a] * (2 * lam[a] - 1) for a in range(4)] + [4 * lam[i] * lam[j] for i, j in edges]
    g = [(4 * lam[a] - 1) * dlam[a] for a in range(4)] + [4 * (lam[i] * dlam[j] + lam[j] * dlam[i]) for i, j in edges]
    return np.array(v), np.array(g)


@pytest.mark.parametrize("order", [1, 2])
def test_bases_match_barycentric(order):
    rng = np.random.default_rng(BASE_SEED)
    pts = rng.dirichlet(np.ones(4), size=200)[:, 1:]
    for p in pts:
        v, g = _bary_basis(order, p)
        assert np.allclose(po.basis_val(order, p), v, atol=1e-10, rtol=0)
        assert np.allclose(po.basis_grad(order, p), g, atol=1e-10, rtol=0)


def test_p2_nodal_interpolation():
    nodes = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [.5, 0, 0], [.5, .5, 0], [0, .5, 0], [0, 0, .5],
                      [.5, 0, .5], [0, .5, .5]])  # auto_p_bases.cpp:1437-1448
    V = np.array([po.basis_val(2, n) for n in nodes])
    assert np.allclose(V, np.eye(10), atol=1e-14)


# ----------------------------------------------------------- quadrature

@pytest.mark.parametrize("qorder,deg", [(1, 1), (2, 2), (3, 3), (4, 4)])
def test_quadrature_exactness(qorder, deg):
    """Every monomial of degree <= the order integrates exactly (closed form
    a! b! c! / (a + b + c + 3)!); the rules of orders 3 and 4 have 6 and 11
    points with positive weights (Xiao-Gimbutas)."""
    pts, w = po.quadrature(qorder)
    assert len(w) == {1: 1, 2: 4, 3: 6, 4: 11}[qorder] and np.all(w > 0)
    assert abs(w.sum() - 1 / 6) < 1e-15
    assert np.all(pts >= 0) and np.all(pts.sum(axis=1) <= 1)
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            for c in range(deg + 1 - a - b):
                exact = math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)
                got = np.sum(w * pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c)
                assert abs(got - exact) < 1e-15


def test_quadrature_order_is_sharp():
    """Orders 3 and 4 are not exact one degree higher (a rule of a lower
    order mistyped as these would pass the exactness test above only by
    chance; this one checks the rule is not accidentally of higher order)."""
    for qorder in (3, 4):
        pts, w = po.quadrature(qorder)
        deg = qorder + 1
        worst = 0.0
        for a in range(deg + 1):
            for b in range(deg + 1 - a):
                c = deg - a - b
                exact = math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)
                worst = max(worst, abs(np.sum(w * pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c) - exact))
        assert worst > 1e-9


def test_hms_points_are_the_published_rule():
    pts, _ = po.quadrature(2)
    a, b = (5 - math.sqrt(5)) / 20, (5 + 3 * math.sqrt(5)) / 20
    expect = np.array([[a, a, b], [a, b, a], [b, a, a], [a, a, a]])
    assert np.allclose(pts, expect, atol=1e-15)


# ------------------------------------------------------------- geometry

def test_geometry_volume_and_linear_reproduction():
    m = kuhn_box(2, 3, 2, order=1, size=(1.0, 1.5, 0.7))
    o = po.Oracle(m, LAM, MU)
    V = m.vertices[m.elem_vertices]
    for e in range(m.n_elements):
        jit, det = o.geometry(e)
        vol = abs(np.linalg.det(V[e, 1:] - V[e, 0])) / 6
        assert abs(det[0] / 6 - vol) < 1e-14
        # gradient of the linear field f(X) = c.X is reproduced exactly
        c = np.array([0.3, -1.2, 2.0])
        f = V[e] @ c
        G = po.basis_grad(1, [0.25] * 3) @ jit[0]
        assert np.allclose(f @ G, c, atol=1e-12)
    assert abs(sum(o.geometry(e)[1][0] for e in range(m.n_elements)) / 6 - 1.05) < 1e-12


# ---------------------------------------------------- complex-step energy

def _energy_cs(mesh, e, x, lam, mu, order, material):
    """Energy of one element written from NeoHookeanElasticity.cpp:337-386 /
    LinearElasticity.cpp:105-134 / SaintVenantElasticity.cpp:235-278 (isotropic
    C), analytic in x so complex-step differentiates it."""
    pts, w = po.quadrature(1 if order == 1 else 2)
    V = mesh.vertices[mesh.elem_vertices[e]]
    T = V[1:] - V[0]
    jit = np.linalg.inv(T).T
    det = np.linalg.det(T)
    nodes = mesh.elem_nodes[e]
    U = np.stack([x[3 * n:3 * n + 3] for n in nodes])
    E = 0
    for q in range(len(w)):
        G = po.basis_grad(order, pts[q]) @ jit
        Du = U.T @ G
        if material == po.NEOHOOKEAN:
            F = np.eye(3) + Du
            J = (F[0, 0] * (F[1, 1] * F[2, 2] - F[1, 2] * F[2, 1]) - F[0, 1] * (F[1, 0] * F[2, 2] - F[1, 2] * F[2, 0])
                 + F[0, 2] * (F[1, 0] * F[2, 1] - F[1, 1] * F[2, 0]))
            lj = np.log(J)
            val = mu / 2 * (np.sum(F * F) - 3 - 2 * lj) + lam / 2 * lj * lj
        elif material == po.SAINT_VENANT:
            Eg = (Du.T @ Du + Du + Du.T) / 2
            val = mu * np.sum(Eg * Eg) + lam / 2 * np.trace(Eg) ** 2
        else:
            eps = (Du + Du.T) / 2
            val = mu * np.sum(eps * eps) + lam / 2 * np.trace(eps) ** 2
        E = E + val * det * w[q]
    return E


def _local_x(mesh, e, xloc, x):
    y = x.astype(complex)
    for a, n in enumerate(mesh.elem_nodes[e]):
        y[3 * n:3 * n + 3] = xloc[3 * a:3 * a + 3]
    return y


def _cs_gradient(mesh, e, x, lam, mu, order, material):
    n = mesh.elem_nodes.shape[1]
    x0 = np.concatenate([x[3 * k:3 * k + 3] for k in mesh.elem_nodes[e]]).astype(complex)
    g = np.zeros(3 * n)
    h = 1e-30
    for i in range(3 * n):
        xp = x0.copy()
        xp[i] += 1j * h
        g[i] = _energy_cs(mesh, e, _local_x(mesh, e, xp, x), lam, mu, order, material).imag / h
    return g


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("material", [po.NEOHOOKEAN, po.LINEAR_ELASTICITY, po.SAINT_VENANT])
def test_hand_derivatives_vs_complex_step(order, material):
    """Margin 1e-12 as generic_elastic_assembler (test_assembler.cpp:191,206,222)."""
    m = kuhn_box(1, 1, 1, order=order)
    o = po.Oracle(m, LAM, MU, material=material)
    rng = np.random.default_rng(BASE_SEED + 3)
    for trial in range(3):
        x = rng.uniform(-0.05, 0.05, size=m.n_dofs)
        for e in (0, 3):
            Ecs = _energy_cs(m, e, x, LAM, MU, order, material).real
            assert abs(o.element_energy(e, x) - Ecs) <= 1e-12 * max(1.0, abs(Ecs))
            gcs = _cs_gradient(m, e, x, LAM, MU, order, material)
            g = o.element_gradient(e, x)
            assert np.max(np.abs(g - gcs)) <= 1e-12 * max(1.0, np.max(np.abs(gcs)))


def _cs_hessian(mesh, e, x, order, material, oracle):
    """Complex-step of the closed-form local gradient sum_q da P(F) G (written
    here from NeoHookeanElasticity.cpp:532-538, not from the oracle)."""
    pts, w = po.quadrature(1 if order == 1 else 2)
    V = mesh.vertices[mesh.elem_vertices[e]]
    T = V[1:] - V[0]
    jit = np.linalg.inv(T).T
    det = np.linalg.det(T)
    n = mesh.elem_nodes.shape[1]
    x0 = np.concatenate([x[3 * k:3 * k + 3] for k in mesh.elem_nodes[e]])

    def grad(xl):
        U = xl.reshape(n, 3)
        g = np.zeros((n, 3), dtype=complex)
        for q in range(len(w)):
            G = po.basis_grad(order, pts[q]) @ jit
            Du = U.T @ G
            if material == po.NEOHOOKEAN:
                F = np.eye(3) + Du
                cof = np.array([np.cross(F[:, 1], F[:, 2]), np.cross(F[:, 2], F[:, 0]), np.cross(F[:, 0], F[:, 1])]).T
                J = np.sum(F[:, 0] * cof[:, 0])
                P = MU * F - MU / J * cof + LAM * np.log(J) / J * cof
            elif material == po.SAINT_VENANT:
                Eg = (Du.T @ Du + Du + Du.T) / 2
                P = (np.eye(3) + Du) @ (LAM * np.trace(Eg) * np.eye(3) + 2 * MU * Eg)
            else:
                P = MU * (Du + Du.T) + LAM * np.trace(Du) * np.eye(3)
            g += det * w[q] * (G @ P.T)
        return g.reshape(-1)

    H = np.zeros((3 * n, 3 * n))
    for i in range(3 * n):
        xp = x0.astype(complex)
        xp[i] += 1e-30j
        H[:, i] = grad(xp).imag / 1e-30
    return H


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("material", [po.NEOHOOKEAN, po.LINEAR_ELASTICITY, po.SAINT_VENANT])
def test_hessian_vs_complex_step(order, material):
    m = kuhn_box(1, 1, 1, order=order)
    o = po.Oracle(m, LAM, MU, material=material)
    rng = np.random.default_rng(BASE_SEED + 4)
    x = rng.uniform(-0.05, 0.05, size=m.n_dofs)
    for e in (1, 4):
        Hcs = _cs_hessian(m, e, x, order, material, o)
        H = o.element_hessian(e, x)
        scale = np.max(np.abs(Hcs))
        assert np.max(np.abs(H - Hcs)) <= 1e-12 * scale
        assert np.max(np.abs(H - H.T)) <= 1e-12 * scale


@pytest.mark.parametrize("order", [1, 2])
def test_nh_hessian_at_rest_equals_linear_stiffness(order):
    """hessian_lin (test_assembler.cpp:17-70), margin 1e-8."""
    m = kuhn_box(2, 2, 2, order=order)
    lam = np.linspace(1e3, 3e3, m.n_elements)
    mu = np.linspace(5e2, 9e2, m.n_elements)
    x0 = np.zeros(m.n_dofs)
    Hn = po.Oracle(m, lam, mu, material=po.NEOHOOKEAN).hessian_scipy(x0)
    Hl = po.Oracle(m, lam, mu, material=po.LINEAR_ELASTICITY).hessian_scipy(x0)
    assert abs(Hn - Hl).max() < 1e-8


@pytest.mark.parametrize("order", [1, 2])
def test_sv_hessian_at_rest_equals_linear_stiffness(order):
    """SaintVenant linearises to linear elasticity: equal Hessians at u = 0."""
    m = kuhn_box(2, 2, 2, order=order)
    lam = np.linspace(1e3, 3e3, m.n_elements)
    mu = np.linspace(5e2, 9e2, m.n_elements)
    x0 = np.zeros(m.n_dofs)
    Hs = po.Oracle(m, lam, mu, material=po.SAINT_VENANT).hessian_scipy(x0)
    Hl = po.Oracle(m, lam, mu, material=po.LINEAR_ELASTICITY).hessian_scipy(x0)
    assert abs(Hs - Hl).max() < 1e-8


@pytest.mark.parametrize("order", [1, 2])
def test_sv_uniaxial_stretch_closed_form(order):
    """u = (s - 1) x e_x on the unit cube: E = diag(e, 0, 0), e = (s^2 - 1)/2,
    psi = (lambda/2 + mu) e^2; the gradient at interior nodes vanishes."""
    m = kuhn_box(2, 2, 2, order=order)
    o = po.Oracle(m, LAM, MU, material=po.SAINT_VENANT)
    s = 1.3
    x = np.zeros((m.n_nodes, 3))
    x[:, 0] = (s - 1) * m.nodes[:, 0]
    x = x.reshape(-1)
    e = (s * s - 1) / 2
    assert abs(o.energy(x) - (LAM / 2 + MU) * e * e) < 1e-12 * (LAM / 2 + MU) * e * e
    inner = np.all((m.nodes > 1e-9) & (m.nodes < 1 - 1e-9), axis=1)
    g = o.gradient(x).reshape(-1, 3)
    assert np.abs(g[inner]).max() < 1e-9 * LAM


@pytest.mark.parametrize("order", [1, 2])
def test_linear_elasticity_is_quadratic(order):
    m = kuhn_box(2, 1, 2, order=order)
    o = po.Oracle(m, LAM, MU, material=po.LINEAR_ELASTICITY)
    x = smooth_displacement(m.nodes, 0.02, 1e-3, BASE_SEED + 5)
    K = o.hessian_scipy(np.zeros(m.n_dofs))
    g = o.gradient(x)
    assert np.allclose(K @ x, g, rtol=0, atol=1e-9 * np.abs(g).max())
    assert abs(0.5 * x @ (K @ x) - o.energy(x)) < 1e-10 * o.energy(x)


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("material", [po.NEOHOOKEAN, po.SAINT_VENANT])
def test_rigid_motion_and_frame_indifference(order, material):
    m = kuhn_box(2, 2, 1, order=order)
    o = po.Oracle(m, LAM, MU, material=material)
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    X = m.nodes
    u_rot = (X @ R.T - X + np.array([0.1, -0.2, 0.3])).reshape(-1)
    assert abs(o.energy(u_rot)) < 1e-9
    assert np.abs(o.gradient(u_rot)).max() < 1e-7
    # E(R(X + u) - X) == E(u)
    u = smooth_displacement(X, 0.03, 1e-3, BASE_SEED + 6).reshape(-1, 3)
    u2 = ((X + u) @ R.T - X).reshape(-1)
    assert abs(o.energy(u2) - o.energy(u.reshape(-1))) < 1e-10 * o.energy(u.reshape(-1))


@pytest.mark.parametrize("order", [1, 2])
def test_finite_differences(order):
    """test_form_derivatives.cpp:330-344: FD step 1e-7, tolerance 1e-4."""
    m = kuhn_box(2, 2, 2, order=order)
    o = po.Oracle(m, LAM, MU)
    rng = np.random.default_rng(BASE_SEED + 7)
    x = rng.uniform(0, 1, size=m.n_dofs) / 100  # test_form_derivatives.cpp:226-227
    d = rng.normal(size=m.n_dofs)
    h = 1e-7
    fd = (o.energy(x + h * d) - o.energy(x - h * d)) / (2 * h)
    assert abs(fd - o.gradient(x) @ d) <= 1e-4 * max(1.0, abs(fd))
    H = o.hessian_scipy(x)
    fdg = (o.gradient(x + h * d) - o.gradient(x - h * d)) / (2 * h)
    assert np.abs(H @ d - fdg).max() <= 1e-4 * max(1.0, np.abs(fdg).max())
    pe = o.energy_per_element(x)
    assert abs(pe.sum() - o.energy(x)) < max(1e-10 * o.energy(x), 1e-10)  # Assembler.cpp:565-569


def test_pattern_counts_match_box_closed_form():
    """nnz = 9 (N + 2 #edges) for the Kuhn box; 20^3 P1 gives 1,157,229."""
    m = kuhn_box(20, 20, 20, order=1)
    o = po.Oracle(m, 1.0, 1.0)
    assert po.lib().orc_hessian_nnz(o.p) == 1157229
    m2 = kuhn_box(3, 2, 2, order=2)
    o2 = po.Oracle(m2, 1.0, 1.0)
    outer, inner, _ = o2.hessian(np.zeros(m2.n_dofs))
    # brute force: node pairs sharing an element, times a 3x3 block
    pairs = set()
    for el in m2.elem_nodes:
        for a in el:
            for b in el:
                pairs.add((int(a), int(b)))
    assert len(inner) == 9 * len(pairs)


# -------------------------------------------------------------- adjoint

@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("material", [po.NEOHOOKEAN, po.LINEAR_ELASTICITY, po.SAINT_VENANT])
def test_material_term_vs_fd(order, material):
    """<term, dtheta> = -p^T (dg/dtheta) dtheta (ElasticForceDerivative.cpp:123-124)."""
    m = kuhn_box(2, 2, 1, order=order)
    E = m.n_elements
    rng = np.random.default_rng(BASE_SEED + 8)
    lam = LAM * np.exp(rng.normal(0, 0.3, E))
    mu = MU * np.exp(rng.normal(0, 0.3, E))
    x = smooth_displacement(m.nodes, 0.03, 1e-3, BASE_SEED + 9)
    p = rng.normal(size=m.n_dofs)
    o = po.Oracle(m, lam, mu, material=material)
    term = o.material_vjp(x, p)
    dl, dm = rng.normal(size=E), rng.normal(size=E)
    h = 1e-3
    gp = po.Oracle(m, lam + h * dl, mu + h * dm, material=material).gradient(x)
    gm = po.Oracle(m, lam - h * dl, mu - h * dm, material=material).gradient(x)
    fd = -p @ (gp - gm) / (2 * h)
    assert abs(term[:E] @ dl + term[E:] @ dm - fd) <= 1e-6 * abs(fd)


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("material", [po.NEOHOOKEAN, po.LINEAR_ELASTICITY, po.SAINT_VENANT])
def test_shape_term_vs_fd(order, material):
    """<term, dX> = + d(p^T g)/dX dX with the displacement dofs held fixed."""
    m = kuhn_box(2, 1, 2, order=order)
    rng = np.random.default_rng(BASE_SEED + 10)
    x = smooth_displacement(m.nodes, 0.03, 1e-3, BASE_SEED + 11)
    p = rng.normal(size=m.n_dofs)
    term = po.Oracle(m, LAM, MU, material=material).shape_vjp(x, p)
    dX = rng.normal(size=m.vertices.shape) * 1e-2
    h = 1e-5

    def g_at(s):
        m2 = kuhn_box(2, 1, 2, order=order)
        m2.vertices = m.vertices + s * dX
        return po.Oracle(m2, LAM, MU, material=material).gradient(x)

    fd = p @ (g_at(h) - g_at(-h)) / (2 * h)
    assert abs(term @ dX.reshape(-1) - fd) <= 1e-6 * max(abs(fd), 1e-12)


def test_baseline_matches_single_thread():
    m = kuhn_box(3, 3, 2, order=2)
    rng = np.random.default_rng(BASE_SEED + 12)
    lam = LAM * np.exp(rng.normal(0, 0.3, m.n_elements))
    o = po.Oracle(m, lam, MU)
    x = smooth_displacement(m.nodes, 0.03, 1e-3, BASE_SEED + 13)
    b = po.Baseline(o, 3)
    assert abs(b.energy(x) - o.energy(x)) < 1e-12 * o.energy(x)
    assert np.allclose(b.gradient(x), o.gradient(x), rtol=0, atol=1e-12 * np.abs(o.gradient(x)).max())
    _, _, ref = o.hessian(x)
    v1 = np.zeros(len(ref))
    b.hessian(x, v1)  # first call: triplets + slot map
    v2 = np.zeros(len(ref))
    b.hessian(x, v2)  # later call: slot map + per-thread arrays
    assert np.allclose(v1, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    assert np.allclose(v2, ref, rtol=0, atol=1e-10 * np.abs(ref).max())


@pytest.mark.parametrize("order", [1, 2])
def test_static_shape_gradient_through_solves(order):
    """dJ/dX = dJ/dX|explicit - shape_vjp(u, p) with H p = dJ/du, u the static
    equilibrium (bottom clamped, top pulled) and J = 1/2 sum_S |X + u - X*|^2
    over the vertices of a mid plane: against central differences of J
    through oracle Newton solves (sparse direct) on moved meshes."""
    import scipy.sparse.linalg as spla
    m = kuhn_box(2, 2, 3, order=order)
    bot, top = face_nodes(m, 2, 0.0), face_nodes(m, 2, 1.0)
    fixed = np.concatenate([3 * bot[:, None] + np.arange(3), 3 * top[:, None] + np.arange(3)]).ravel()
    vals = np.concatenate([np.zeros(3 * bot.size), np.tile([0.0, 0.0, 0.05], top.size)])
    free = np.setdiff1d(np.arange(m.n_dofs), fixed)
    rng = np.random.default_rng(BASE_SEED + 14)
    lam = LAM * np.exp(rng.normal(0, 0.3, m.n_elements))
    mu = MU * np.exp(rng.normal(0, 0.3, m.n_elements))
    v2n = np.empty(m.vertices.shape[0], dtype=np.int64)
    v2n[m.elem_vertices.ravel()] = m.elem_nodes[:, :4].ravel()
    S = np.flatnonzero(np.abs(m.vertices[:, 2] - 1.0 / 3.0) < 1e-9)
    Sd = (3 * v2n[S][:, None] + np.arange(3)).ravel()
    Xstar = m.vertices[S] + rng.normal(size=(S.size, 3)) * 1e-3

    def solve(V, u0):
        m2 = kuhn_box(2, 2, 3, order=order)
        m2.vertices = V
        o = po.Oracle(m2, lam, mu)
        u = u0.copy()
        u[fixed] = vals
        for _ in range(30):
            g = o.gradient(u)
            if np.linalg.norm(g[free]) < 1e-10:
                break
            H = o.hessian_scipy(u).tocsr()[free][:, free].tocsc()
            u[free] -= spla.spsolve(H, g[free])
        return u, o

    def J(V, u):
        return 0.5 * np.sum((V[S] + u[Sd].reshape(-1, 3) - Xstar) ** 2)

    V0 = m.vertices.copy()
    u, o = solve(V0, np.zeros(m.n_dofs))
    r = V0[S] + u[Sd].reshape(-1, 3) - Xstar
    dJdu = np.zeros(m.n_dofs)
    dJdu[Sd] = r.ravel()
    p = np.zeros(m.n_dofs)
    p[free] = spla.spsolve(o.hessian_scipy(u).tocsr()[free][:, free].tocsc(), dJdu[free])
    explicit = np.zeros_like(V0)
    explicit[S] = r
    grad = explicit.ravel() - o.shape_vjp(u, p)
    h = 1e-5
    for _ in range(2):
        dV = rng.normal(size=V0.shape) * 0.02
        fd = (J(V0 + h * dV, solve(V0 + h * dV, u)[0]) - J(V0 - h * dV, solve(V0 - h * dV, u)[0])) / (2 * h)
        assert abs(grad @ dV.ravel() - fd) <= 1e-6 * abs(fd)


@pytest.mark.parametrize("order", [1, 2])
def test_psd_projection_oracle(order):
    """The projected Hessian: same pattern as the plain one; every element
    matrix after projection is PSD, so the sum is PSD; PSD element matrices
    (linear elasticity) are left unchanged; projecting a compressed NeoHookean
    state changes it; projection of the projected element matrices is a no-op."""
    import scipy.sparse as sp
    m = kuhn_box(2, 2, 1, order=order)
    lam, mu = np.full(m.n_elements, LAM), np.full(m.n_elements, MU)
    x = smooth_displacement(m.nodes, 0.05, 1e-3, BASE_SEED + 15)
    ol = po.Oracle(m, lam, mu, material=po.LINEAR_ELASTICITY)
    o1, i1, v1 = ol.hessian(x)
    o2, i2, v2 = ol.hessian_psd(x)
    assert np.array_equal(o1, o2) and np.array_equal(i1, i2)
    assert np.abs(v1 - v2).max() <= 1e-9 * np.abs(v1).max()
    # a strongly compressed NeoHookean state has indefinite element Hessians
    xc = x.copy()
    xc[2::3] -= 0.45 * m.nodes[:, 2]
    on = po.Oracle(m, lam, mu, material=po.NEOHOOKEAN)
    H = on.hessian_scipy(xc).toarray()
    outer, inner, vals = on.hessian_psd(xc)
    P = sp.csc_matrix((vals, inner, outer), shape=H.shape).toarray()
    assert np.linalg.eigvalsh(H).min() < -1e-6 * np.abs(H).max()
    assert np.linalg.eigvalsh(P).min() > -1e-9 * np.abs(P).max()
    assert np.abs(P - P.T).max() <= 1e-12 * np.abs(P).max()


# ------------------------------------------------------ FixedCorotational

def test_fcr_stress_and_stiffness_are_derivatives():
    """P = dpsi/dF and A = dP/dF (central FD), A symmetric."""
    from oracle import fcr
    rng = np.random.default_rng(BASE_SEED + 16)
    h = 1e-6
    for _ in range(4):
        F = np.eye(3) + 0.2 * rng.normal(size=(3, 3))
        P = fcr.stress(F, LAM, MU)
        A = fcr.stiffness(F, LAM, MU)
        for i in range(3):
            for j in range(3):
                d = np.zeros((3, 3))
                d[i, j] = h
                fd = (fcr.energy_density(F + d, LAM, MU) - fcr.energy_density(F - d, LAM, MU)) / (2 * h)
                assert abs(fd - P[i, j]) <= 1e-7 * np.abs(P).max()
                dP = (fcr.stress(F + d, LAM, MU) - fcr.stress(F - d, LAM, MU)) / (2 * h)
                assert np.abs(dP.reshape(-1, order="F") - A[:, i + 3 * j]).max() <= 1e-7 * np.abs(A).max()
        assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()


def test_fcr_closed_forms():
    """Rotations cost nothing; a uniaxial stretch diag(s, 1, 1) costs
    mu (s-1)^2 + lambda/2 (s-1)^2; at rest the element Hessian is the linear
    elasticity stiffness."""
    from oracle import fcr
    th = 0.8
    R = np.array([[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]])
    assert abs(fcr.energy_density(R, LAM, MU)) < 1e-9 * MU
    assert np.abs(fcr.stress(R, LAM, MU)).max() < 1e-9 * MU
    s = 1.3
    assert abs(fcr.energy_density(np.diag([s, 1, 1]), LAM, MU) - (MU + LAM / 2) * (s - 1) ** 2) < 1e-9 * MU
    m = kuhn_box(1, 1, 1, order=2)
    o = fcr.FcrOracle(m, LAM, MU)
    ol = po.Oracle(m, LAM, MU, material=po.LINEAR_ELASTICITY)
    x0 = np.zeros(m.n_dofs)
    assert abs(o.hessian_scipy(x0) - ol.hessian_scipy(x0)).max() < 1e-8 * LAM


@pytest.mark.parametrize("order", [1, 2])
def test_fcr_assembly_derivatives_vs_fd(order):
    """Global gradient and Hessian against central FD of the global energy
    and gradient; material and shape terms against FD of the gradient."""
    from oracle import fcr
    m = kuhn_box(2, 1, 1, order=order)
    rng = np.random.default_rng(BASE_SEED + 17)
    lam = LAM * np.exp(rng.normal(0, 0.3, m.n_elements))
    mu = MU * np.exp(rng.normal(0, 0.3, m.n_elements))
    o = fcr.FcrOracle(m, lam, mu)
    x = smooth_displacement(m.nodes, 0.05, 1e-3, BASE_SEED + 18)
    d = rng.normal(size=m.n_dofs)
    h = 1e-6
    fd = (o.energy(x + h * d) - o.energy(x - h * d)) / (2 * h)
    g = o.gradient(x)
    assert abs(fd - g @ d) <= 1e-6 * abs(fd)
    fdg = (o.gradient(x + h * d) - o.gradient(x - h * d)) / (2 * h)
    assert np.abs(o.hessian_scipy(x) @ d - fdg).max() <= 1e-6 * np.abs(fdg).max()
    p = rng.normal(size=m.n_dofs)
    term = o.material_vjp(x, p)
    dl, dm = rng.normal(size=m.n_elements), rng.normal(size=m.n_elements)
    gp = fcr.FcrOracle(m, lam + h * dl * LAM, mu + h * dm * MU).gradient(x)
    gm = fcr.FcrOracle(m, lam - h * dl * LAM, mu - h * dm * MU).gradient(x)
    fdm = -p @ (gp - gm) / (2 * h)
    assert abs(term[:m.n_elements] @ (dl * LAM) + term[m.n_elements:] @ (dm * MU) - fdm) <= 1e-6 * abs(fdm)
    shape = o.shape_vjp(x, p)
    dX = rng.normal(size=m.vertices.shape) * 1e-2

    def g_at(sg):
        m2 = kuhn_box(2, 1, 1, order=order)
        m2.vertices = m.vertices + sg * dX
        return fcr.FcrOracle(m2, lam, mu).gradient(x)

    hs = 1e-5
    fds = p @ (g_at(hs) - g_at(-hs)) / (2 * hs)
    assert abs(shape @ dX.reshape(-1) - fds) <= 1e-5 * max(abs(fds), 1e-12)


# ------------------------------------------------------ MooneyRivlin

MR_C1, MR_C2, MR_K = 2.0e5, 1.0e5, 2.0e6


def test_mooney_stress_is_complex_step_of_the_energy():
    """The closed-form P against d psi / d F by complex step of the literal
    energy (F~ = J^{-1/3} F and the invariants of C~), exact to rounding;
    the stiffness is symmetric and agrees with central FD of P."""
    from oracle import mooney
    rng = np.random.default_rng(BASE_SEED + 30)
    for scale in (0.05, 0.3):
        F = np.eye(3) + scale * rng.normal(size=(3, 3))
        if np.linalg.det(F) <= 0.2:
            F = np.eye(3) + 0.1 * rng.normal(size=(3, 3))
        P = mooney.stress(F, MR_C1, MR_C2, MR_K)
        cs = np.zeros((3, 3))
        for i in range(3):
            for j in range(3):
                Fc = F.astype(np.complex128)
                Fc[i, j] += 1e-30j
                cs[i, j] = mooney.energy_density(Fc, MR_C1, MR_C2, MR_K).imag / 1e-30
        assert np.abs(cs - P).max() <= 1e-12 * np.abs(P).max()
        A = mooney.stiffness(F, MR_C1, MR_C2, MR_K)
        assert np.abs(A - A.T).max() <= 1e-10 * np.abs(A).max()
        h = 1e-6
        for i in range(3):
            for j in range(3):
                d = np.zeros((3, 3))
                d[i, j] = h
                dP = (mooney.stress(F + d, MR_C1, MR_C2, MR_K) - mooney.stress(F - d, MR_C1, MR_C2, MR_K)) / (2 * h)
                assert np.abs(dP.reshape(-1, order="F") - A[:, i + 3 * j]).max() <= 1e-7 * np.abs(A).max()


def test_mooney_closed_forms():
    """Zero energy and stress at rest and under a rotation; a pure volume
    change s I costs k/2 ln^2 s^3 (the isochoric invariants stay 3); at rest
    the element Hessian is the linear-elasticity stiffness with
    mu = 2 (c1 + c2), lambda = k - 4/3 (c1 + c2)."""
    from oracle import mooney
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    for F in (np.eye(3), R):
        assert abs(mooney.energy_density(F, MR_C1, MR_C2, MR_K)) < 1e-9 * MR_C1
        assert np.abs(mooney.stress(F, MR_C1, MR_C2, MR_K)).max() < 1e-9 * MR_C1
    s = 1.1
    assert abs(mooney.energy_density(s * np.eye(3), MR_C1, MR_C2, MR_K) - MR_K / 2 * np.log(s ** 3) ** 2) < 1e-9 * MR_K
    F = np.eye(3) + 0.2 * np.random.default_rng(BASE_SEED + 31).normal(size=(3, 3))
    assert abs(mooney.energy_density(R @ F, MR_C1, MR_C2, MR_K) - mooney.energy_density(F, MR_C1, MR_C2, MR_K)) \
        < 1e-9 * abs(mooney.energy_density(F, MR_C1, MR_C2, MR_K))
    mu = 2 * (MR_C1 + MR_C2)
    lam = MR_K - 2.0 / 3.0 * mu
    for order in (1, 2):
        m = kuhn_box(1, 1, 1, order=order)
        o = mooney.MooneyOracle(m, MR_C1, MR_C2, MR_K)
        ol = po.Oracle(m, lam, mu, material=po.LINEAR_ELASTICITY)
        x0 = np.zeros(m.n_dofs)
        assert abs(o.hessian_scipy(x0) - ol.hessian_scipy(x0)).max() < 1e-9 * MR_K


@pytest.mark.parametrize("order", [1, 2])
def test_mooney_assembly_derivatives_vs_fd(order):
    """Global gradient and Hessian against central FD of the global energy
    and gradient; the shape term against FD of the gradient."""
    from oracle import mooney
    m = kuhn_box(2, 1, 1, order=order)
    rng = np.random.default_rng(BASE_SEED + 32)
    c1 = MR_C1 * np.exp(rng.normal(0, 0.3, m.n_elements))
    c2 = MR_C2 * np.exp(rng.normal(0, 0.3, m.n_elements))
    kk = MR_K * np.exp(rng.normal(0, 0.3, m.n_elements))
    o = mooney.MooneyOracle(m, c1, c2, kk)
    x = smooth_displacement(m.nodes, 0.05, 1e-3, BASE_SEED + 33)
    d = rng.normal(size=m.n_dofs)
    h = 1e-6
    fd = (o.energy(x + h * d) - o.energy(x - h * d)) / (2 * h)
    g = o.gradient(x)
    assert abs(fd - g @ d) <= 1e-6 * abs(fd)
    fdg = (o.gradient(x + h * d) - o.gradient(x - h * d)) / (2 * h)
    assert np.abs(o.hessian_scipy(x) @ d - fdg).max() <= 1e-6 * np.abs(fdg).max()
    p = rng.normal(size=m.n_dofs)
    shape = o.shape_vjp(x, p)
    dX = rng.normal(size=m.vertices.shape) * 1e-2

    def g_at(sg):
        m2 = kuhn_box(2, 1, 1, order=order)
        m2.vertices = m.vertices + sg * dX
        return mooney.MooneyOracle(m2, c1, c2, kk).gradient(x)

    hs = 1e-5
    fds = p @ (g_at(hs) - g_at(-hs)) / (2 * hs)
    assert abs(shape @ dX.reshape(-1) - fds) <= 1e-5 * max(abs(fds), 1e-12)
