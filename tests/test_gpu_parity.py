"""GPU parity of libpfgpu against the FP64 oracle (runs on an MI355X).

Bars (BASELINE targets): sparsity indices bit-exact; FP64 energy, gradient and
Hessian within 1e-10 relative; mixed precision within the stated FP32
tolerances below.  Relative errors are max|a-b| / max|b| over the array.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from polyfem_amd import NLAssembler, PfgpuError
from polyfem_amd.mesh import BASE_SEED, kuhn_box, lame_from_young, random_lame, smooth_displacement, tet_mesh

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
# FP32 element kernels with FP64 sums: the errors measured against the oracle
# over every path, material and order on these states are at most 5.4e-8
# (energy), 3.1e-7 (gradient), 2.2e-7 (Hessian), 1.9e-7 (Hv) and 3.3e-7
# (VJPs); the bars sit about three times above.
MIXED_TOL = {"energy": 2e-7, "gradient": 1e-6, "hessian": 1e-6, "hvp": 1e-6, "vjp": 1e-6}

PATHS = ["fused", "plain", "factored"]
CASES = [(1, "NeoHookean"), (2, "NeoHookean"), (1, "LinearElasticity"), (2, "LinearElasticity")]
# SaintVenant runs on the plain path only
SV_CASES = [(1, "SaintVenant"), (2, "SaintVenant")]
PATH_CASES = [(p, o, n) for p in PATHS for o, n in CASES] + [("plain", o, n) for o, n in SV_CASES]
MAT = {"NeoHookean": po.NEOHOOKEAN, "LinearElasticity": po.LINEAR_ELASTICITY, "SaintVenant": po.SAINT_VENANT}


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def setup_case(order, name, grid=(3, 2, 3), seed=0, precision="fp64", path="plain"):
    m = kuhn_box(*grid, order=order)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 100 + seed, E_mean=1e6, E_sigma=0.5)
    x = smooth_displacement(m.nodes, 0.05, 1e-3, BASE_SEED + 200 + seed)
    a = NLAssembler(name, m, lame=(lam, mu), precision=precision, path=path)
    o = po.Oracle(m, lam, mu, material=MAT[name])
    return m, x, a, o, lam, mu


@pytest.mark.parametrize("path,order,name", PATH_CASES)
def test_fp64_parity(order, name, path):
    m, x, a, o, lam, mu = setup_case(order, name, path=path)
    print(f"case order={order} material={name} path={path} E={m.n_elements}", flush=True)
    assert rel(a.assemble_energy(x), o.energy(x)) < FP64_TOL
    assert rel(a.assemble_energy_per_element(x), o.energy_per_element(x)) < FP64_TOL
    assert rel(a.assemble_gradient(x), o.gradient(x)) < FP64_TOL
    outer, inner, ref = o.hessian(x)
    go, gi = a.hessian_pattern()
    assert np.array_equal(go, outer) and np.array_equal(gi, inner)  # bit-exact sparsity
    assert rel(a.hessian_values(x), ref) < FP64_TOL
    e, g, vals = a.assemble_energy_gradient_hessian(x)
    assert rel(e, o.energy(x)) < FP64_TOL and rel(g, o.gradient(x)) < FP64_TOL and rel(vals, ref) < FP64_TOL
    rng = np.random.default_rng(BASE_SEED + 300)
    v = rng.normal(size=m.n_dofs)
    H = o.hessian_scipy(x)
    assert rel(a.hessian_vector_product(x, v), H @ v) < FP64_TOL
    p = rng.normal(size=m.n_dofs)
    assert rel(a.force_material_derivative(x, p), o.material_vjp(x, p)) < FP64_TOL
    assert rel(a.force_shape_derivative(x, p), o.shape_vjp(x, p)) < FP64_TOL
    # forward mixed derivative is the transpose of the material term
    dl, dm = rng.normal(size=m.n_elements), rng.normal(size=m.n_elements)
    jvp = a.material_jvp(x, dl, dm)
    term = o.material_vjp(x, p)
    lhs = -p @ jvp
    rhs = term[:m.n_elements] @ dl + term[m.n_elements:] @ dm
    assert abs(lhs - rhs) <= 1e-10 * abs(rhs)


@pytest.mark.parametrize("path,order,name", PATH_CASES)
def test_mixed_precision_parity(order, name, path):
    m, x, a, o, lam, mu = setup_case(order, name, precision="mixed", path=path, seed=1)
    print(f"case mixed order={order} material={name} path={path}", flush=True)
    assert rel(a.assemble_energy(x), o.energy(x)) < MIXED_TOL["energy"]
    assert rel(a.assemble_gradient(x), o.gradient(x)) < MIXED_TOL["gradient"]
    outer, inner, ref = o.hessian(x)
    vals = a.hessian_values(x)
    assert vals.dtype == np.float32
    assert rel(vals, ref) < MIXED_TOL["hessian"]
    v = np.random.default_rng(BASE_SEED + 301).normal(size=m.n_dofs)
    assert rel(a.hessian_vector_product(x, v), o.hessian_scipy(x) @ v) < MIXED_TOL["hvp"]
    p = np.random.default_rng(BASE_SEED + 302).normal(size=m.n_dofs)
    assert rel(a.force_material_derivative(x, p), o.material_vjp(x, p)) < MIXED_TOL["vjp"]
    assert rel(a.force_shape_derivative(x, p), o.shape_vjp(x, p)) < MIXED_TOL["vjp"]


@pytest.mark.parametrize("path", PATHS)
def test_config0_full_size_fp64(path):
    """The 20^3 x 6 P1 NeoHookean state (48,000 tets, 27,783 dofs, nnz 1,157,229)."""
    m = kuhn_box(20, 20, 20, order=1)
    lam, mu = lame_from_young(1e6, 0.4)
    x = smooth_displacement(m.nodes, 0.05, 1e-3, BASE_SEED)
    a = NLAssembler("NeoHookean", m, young=1e6, poisson=0.4, path=path)
    o = po.Oracle(m, lam, mu)
    assert a.nnz == 1157229
    e, g, vals = a.assemble_energy_gradient_hessian(x)
    outer, inner, ref = o.hessian(x)
    go, gi = a.hessian_pattern()
    assert np.array_equal(go, outer) and np.array_equal(gi, inner)
    assert rel(e, o.energy(x)) < FP64_TOL
    assert rel(g, o.gradient(x)) < FP64_TOL
    assert rel(vals, ref) < FP64_TOL


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("path", PATHS)
def test_bitwise_deterministic(precision, path):
    m, x, a, o, lam, mu = setup_case(2, "NeoHookean", grid=(4, 3, 3), precision=precision, path=path, seed=2)
    r1 = a.assemble_energy_gradient_hessian(x)
    r2 = a.assemble_energy_gradient_hessian(x)
    assert r1[0] == r2[0]
    assert np.array_equal(r1[1], r2[1]) and np.array_equal(r1[2], r2[2])
    v = np.ones(m.n_dofs)
    assert np.array_equal(a.hessian_vector_product(x, v), a.hessian_vector_product(x, v))


def test_inverted_element_reports_invalid_state():
    """Inverted elements give ln of a negative J -> NaN; the ABI returns
    INVALID_STATE instead of crashing (ElasticForm.cpp:388-405)."""
    m, x, a, o, lam, mu = setup_case(1, "NeoHookean", grid=(2, 2, 2), seed=3)
    y = x.copy()
    n0 = m.elem_nodes[0]
    # collapse element 0 through its opposite face
    y.reshape(-1, 3)[n0[3]] = m.nodes[n0[0]] - m.nodes[n0[3]] + (m.nodes[n0[0]] - m.nodes[n0[3]])
    with pytest.raises(PfgpuError) as ei:
        a.assemble_energy(y)
    assert ei.value.status == 3
    with pytest.raises(PfgpuError) as ei:
        a.assemble_gradient(y)
    assert ei.value.status == 3


def test_update_lame_params():
    m, x, a, o, lam, mu = setup_case(2, "NeoHookean", seed=4)
    lam2, mu2 = lam * 1.3, mu * 0.7
    a.update_lame_params(lam2, mu2)
    o2 = po.Oracle(m, lam2, mu2)
    assert rel(a.assemble_gradient(x), o2.gradient(x)) < FP64_TOL


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("path", ["fused", "factored"])
def test_many_patches_and_chunks(order, precision, path):
    """Meshes large enough for many fused patches and staging chunks and many
    factored-gather chunks, against the oracle."""
    grid = (8, 8, 8) if order == 2 else (16, 14, 12)
    m, x, a, o, lam, mu = setup_case(order, "NeoHookean", grid=grid, precision=precision, path=path, seed=5)
    tol = FP64_TOL if precision == "fp64" else MIXED_TOL["hessian"]
    print(f"case many patches order={order} precision={precision} path={path} E={m.n_elements}", flush=True)
    e, g, vals = a.assemble_energy_gradient_hessian(x)
    outer, inner, ref = o.hessian(x)
    assert rel(vals, ref) < tol
    assert rel(g, o.gradient(x)) < (FP64_TOL if precision == "fp64" else MIXED_TOL["gradient"])
    assert rel(e, o.energy(x)) < (FP64_TOL if precision == "fp64" else MIXED_TOL["energy"])
    v = np.random.default_rng(BASE_SEED + 303).normal(size=m.n_dofs)
    assert rel(a.hessian_vector_product(x, v), o.hessian_scipy(x) @ v) < (FP64_TOL if precision == "fp64" else MIXED_TOL["hvp"])


@pytest.mark.parametrize("order,name", CASES + SV_CASES)
def test_energy_lame_gradient(order, name):
    """dE/d(lambda, mu) per element: the energy is linear in the Lame pair, so
    the derivative equals the per-element energy with unit (1, 0) / (0, 1)."""
    m, x, a, o, lam, mu = setup_case(order, name)
    d = a.energy_lame_gradient(x)
    E = m.n_elements
    o_l = po.Oracle(m, np.ones(E), np.zeros(E), material=MAT[name])
    o_m = po.Oracle(m, np.zeros(E), np.ones(E), material=MAT[name])
    assert rel(d[:E], o_l.energy_per_element(x)) < FP64_TOL
    assert rel(d[E:], o_m.energy_per_element(x)) < FP64_TOL
    assert abs(lam @ d[:E] + mu @ d[E:] - o.energy(x)) <= FP64_TOL * abs(o.energy(x))


def test_autotune_and_tunings_agree():
    """Every plain launch shape gives bit-identical results (the summation
    order does not depend on the workgroup sizes); the autotuner's choice
    matches the oracle."""
    torch = pytest.importorskip("torch")
    m, x, a, o, lam, mu = setup_case(2, "NeoHookean", grid=(4, 3, 3), precision="mixed", path="plain")
    ref = None
    for gt in (128, 256, 512):
        for ht in (64, 128, 256):
            a.set_tuning("plain", gt, ht)
            e, g, vals = a.assemble_energy_gradient_hessian(x)
            if ref is None:
                ref = (e, g, vals)
            assert e == ref[0] and np.array_equal(g, ref[1]) and np.array_equal(vals, ref[2]), (gt, ht)
    # the factored route's row-unit gather (64 to 256 threads): its chunking
    # depends on the size, its per-block summation order does not; at 512
    # threads the per-contribution gather forms each block with another
    # operation order (same values to rounding)
    fref = None
    for gt in (64, 128, 256, 512):
        a.set_tuning("factored", gt, 256)
        e, g, vals = a.assemble_energy_gradient_hessian(x)
        if fref is None:
            fref = vals
        assert e == ref[0] and np.array_equal(g, ref[1]), gt
        if gt <= 256:
            assert np.array_equal(vals, fref), gt
        else:
            assert rel(vals, fref) < 1e-6, gt
    assert rel(fref, ref[2]) < MIXED_TOL["hessian"]
    xd = torch.tensor(x, device="cuda:0")
    rep = a.autotune_dev(xd.data_ptr())
    print("autotune:", rep, flush=True)
    assert rep["candidates"] >= 13 and rep["ms"] > 0 and rep["ms_factored"] > 0
    e, g, vals = a.assemble_energy_gradient_hessian(x)
    _, _, hv = o.hessian(x)
    assert rel(vals, hv) < MIXED_TOL["hessian"] and rel(g, o.gradient(x)) < MIXED_TOL["gradient"]


def test_saint_venant_is_plain_path_only():
    """The fused kernel carries the NeoHookean/LE tangent only: choosing it for
    SaintVenant is an error, and the autotuner keeps the plain path."""
    import torch
    from polyfem_amd import _ffi
    m = kuhn_box(2, 2, 2, order=2)
    a = NLAssembler("SaintVenant", m, young=1e6, poisson=0.35, precision="mixed")
    for path in ("fused", "factored"):
        with pytest.raises(PfgpuError) as err:
            a.set_path(path)
        assert err.value.status == _ffi.ERR_UNSUPPORTED
    x = torch.zeros(m.n_dofs, dtype=torch.float64, device="cuda")
    rep = a.autotune_dev(x.data_ptr())
    assert rep["path"] == "plain" and rep["ms_fused"] is None and rep["ms_factored"] is None


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("order,name", [(1, "NeoHookean"), (2, "NeoHookean"), (2, "SaintVenant")])
def test_project_to_psd(order, name, precision):
    """Element Hessians projected onto the PSD cone (eigenvalues <= 0 clamped)
    against the oracle's projection through LAPACK eigh, on a strongly
    compressed state whose element Hessians are indefinite."""
    m = kuhn_box(2, 2, 2, order=order)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 110)
    x = smooth_displacement(m.nodes, 0.05, 1e-3, BASE_SEED + 111)
    x[2::3] -= 0.45 * m.nodes[:, 2]
    print(f"psd order={order} {name} {precision}", flush=True)
    a = NLAssembler(name, m, lame=(lam, mu), precision=precision)
    o = po.Oracle(m, lam, mu, material=MAT[name])
    outer, inner, ref = o.hessian_psd(x)
    _, _, plain = o.hessian(x)
    assert rel(plain, ref) > 1e-6  # the projection changes this state
    go, gi = a.hessian_pattern()
    assert np.array_equal(go, outer) and np.array_equal(gi, inner)
    tol = FP64_TOL if precision == "fp64" else 1e-4
    assert rel(a.hessian_values(x, project_to_psd=True), ref) < tol
    # the plain values are unaffected by an earlier projected call
    assert rel(a.hessian_values(x), plain) < (FP64_TOL if precision == "fp64" else MIXED_TOL["hessian"])


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("compress", [0.0, 0.45])
def test_project_to_psd_mixed_sign_iteration(order, compress):
    """FP32 element Hessians take the matrix-core projection (Newton-Schulz
    iterations for sign(A); P1 with two elements per product, so the P1 mesh
    has an odd element count to leave the last element alone): against the
    oracle's LAPACK projection on a 4^3 mesh, with elements of zero Lame
    parameters (zero Hessians, which must stay zero), at 1e-5 relative, ten
    times tighter than the FP32 bar of the eigendecomposition test above."""
    m = kuhn_box(4, 4, 4, order=order)
    if order == 1:
        m = tet_mesh(m.vertices, m.elem_vertices[:-1], order=1)
        assert m.n_elements % 2 == 1
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 112)
    lam[::17] = 0.0
    mu[::17] = 0.0
    x = smooth_displacement(m.nodes, 0.05, 1e-3, BASE_SEED + 113)
    x[2::3] -= compress * m.nodes[:, 2]
    print(f"psd sign iteration order={order} compress={compress}", flush=True)
    a = NLAssembler("NeoHookean", m, lame=(lam, mu), precision="mixed")
    o = po.Oracle(m, lam, mu)
    _, _, ref = o.hessian_psd(x)
    _, _, plain = o.hessian(x)
    assert rel(plain, ref) > 1e-4  # the projection changes this state
    got = a.hessian_values(x, project_to_psd=True)
    assert rel(got, ref) < 1e-5
    # a second call gives the same values bit for bit
    assert np.array_equal(a.hessian_values(x, project_to_psd=True), got)


def test_default_paths():
    """Without a path argument the library takes the factored path for
    NeoHookean and linear elasticity and the plain path for the others."""
    m = kuhn_box(2, 2, 2, order=1)
    for name, path in (("NeoHookean", "factored"), ("LinearElasticity", "factored"), ("SaintVenant", "plain"),
                       ("FixedCorotational", "plain")):
        a = NLAssembler(name, m, young=1e6, poisson=0.3)
        assert a.path == path, (name, a.path)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("order", [1, 2])
def test_irregular_mesh_all_paths(order, precision):
    """Shuffled element order, shuffled vertex ids and perturbed vertices (no
    lattice structure left in the numbering or the adjacency): every path
    against the oracle, and the paths against each other."""
    from polyfem_amd.mesh import tet_mesh
    base = kuhn_box(4, 3, 3, order=1)
    rng = np.random.default_rng(BASE_SEED + 150)
    perm_v = rng.permutation(base.vertices.shape[0])
    inv = np.empty_like(perm_v)
    inv[perm_v] = np.arange(perm_v.size)
    V = base.vertices[perm_v] + rng.normal(size=base.vertices.shape) * 0.03
    T = inv[base.elem_vertices][rng.permutation(base.n_elements)]
    m = tet_mesh(V, T, order=order)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 151, E_mean=1e6, E_sigma=0.5)
    x = smooth_displacement(m.nodes, 0.05, 1e-3, BASE_SEED + 152)
    o = po.Oracle(m, lam, mu)
    outer, inner, ref = o.hessian(x)
    g_ref, e_ref = o.gradient(x), o.energy(x)
    tol = FP64_TOL if precision == "fp64" else MIXED_TOL["hessian"]
    print(f"irregular mesh order={order} {precision} E={m.n_elements}", flush=True)
    vals = {}
    for path in PATHS:
        a = NLAssembler("NeoHookean", m, lame=(lam, mu), precision=precision, path=path)
        e, g, v = a.assemble_energy_gradient_hessian(x)
        go, gi = a.hessian_pattern()
        assert np.array_equal(go, outer) and np.array_equal(gi, inner)
        assert rel(v, ref) < tol, path
        assert rel(g, g_ref) < (FP64_TOL if precision == "fp64" else MIXED_TOL["gradient"]), path
        assert rel(e, e_ref) < (FP64_TOL if precision == "fp64" else MIXED_TOL["energy"]), path
        vals[path] = v
    assert rel(vals["factored"], vals["plain"]) < tol


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("order", [1, 2])
def test_high_valence_split_blocks(order, precision):
    """A centre vertex in 60 tets: its blocks are summed in pieces of 8
    contributions by several threads and combined in piece order.  The
    row-unit gather at 64, 128 and 256 threads gives bit-identical values,
    gradients and energies, within tolerance of the oracle and of the plain
    path."""
    from meshes import star_mesh
    m = star_mesh(30, order=order, seed=3)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 161, E_mean=1e6, E_sigma=0.5)
    x = smooth_displacement(m.nodes, 0.05, 1e-3, BASE_SEED + 162)
    o = po.Oracle(m, lam, mu)
    _, _, ref = o.hessian(x)
    tol = FP64_TOL if precision == "fp64" else MIXED_TOL["hessian"]
    gtol = FP64_TOL if precision == "fp64" else MIXED_TOL["gradient"]
    a = NLAssembler("NeoHookean", m, lame=(lam, mu), precision=precision, path="factored")
    print(f"star mesh order={order} {precision} E={m.n_elements}", flush=True)
    first = None
    for gt in (64, 128, 256):
        a.set_tuning("factored", gt, 256)
        e, g, v = a.assemble_energy_gradient_hessian(x)
        # the Hessian alone: the records come from the Hessian-only element
        # pass (same values to rounding), the gather without the gradient
        h = a.hessian_values(x)
        if first is None:
            first = (e, g, v, h)
            assert rel(v, ref) < tol and rel(g, o.gradient(x)) < gtol and rel(h, ref) < tol
        assert e == first[0] and np.array_equal(g, first[1]) and np.array_equal(v, first[2]), gt
        assert np.array_equal(h, first[3]), gt
    p = NLAssembler("NeoHookean", m, lame=(lam, mu), precision=precision, path="plain")
    assert rel(p.assemble_energy_gradient_hessian(x)[2], first[2]) < tol
