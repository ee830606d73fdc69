"""Synthetic tetrahedral box meshes and seeded states for the elasticity path.

The meshes are Kuhn (Freudenthal) splits of a structured box grid: every hex
cell (i, j, k) is cut into the six tetrahedra that share its main diagonal.
Vertex ``v0 = (i, j, k)`` is walked to ``v3 = (i+1, j+1, k+1)`` along the three
axes in the order of one permutation; odd permutations get ``v1``/``v2``
swapped so that every tetrahedron has a positive Jacobian (PolyFEM does not
reorient tets, ``mesh/mesh3D/CMesh3D.cpp:197,405``, so a generator has to emit
positively oriented ones).

Degrees of freedom follow PolyFEM's numbering rule for Lagrange tets
(``basis/LagrangeBasis3d.cpp:238-330`` and ``:853-888``,
``mesh/MeshNodes.cpp:159-180``): nodes get ids in first-encounter order while
looping over the elements, and within an element the order is the four
vertices, then the six edges ``(v0v1, v1v2, v2v0, v0v3, v1v3, v2v3)``.  The
local P2 node order matches ``autogen/auto_p_bases.cpp:1437-1448``.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

#: Base seed of the reference's differentiability tests (``tests/test_diff.cpp:42``).
BASE_SEED = 0x6A09E667F3BCC909

# Local edges of a tet, in PolyFEM's order (LagrangeBasis3d.cpp:283-290).
TET_EDGES = np.array([[0, 1], [1, 2], [2, 0], [0, 3], [1, 3], [2, 3]], dtype=np.int64)

_PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
_PERM_ODD = [False, True, True, False, False, True]


@dataclass
class TetMesh:
    """A conforming P1 or P2 tet mesh with P1 geometry.

    ``vertices``/``elem_vertices`` describe the geometry (four vertices per
    element).  ``nodes``/``elem_nodes`` describe the discretisation: the
    element-to-node table is the input state of every assembly call.
    """

    order: int
    vertices: np.ndarray  # (Nv, 3) float64
    elem_vertices: np.ndarray  # (E, 4) int32, geometric vertex ids
    nodes: np.ndarray  # (N, 3) float64
    elem_nodes: np.ndarray  # (E, n) int32, n = 4 (P1) or 10 (P2)
    grid: tuple = field(default=(0, 0, 0))
    size: tuple = field(default=(1.0, 1.0, 1.0))

    @property
    def n_elements(self) -> int:
        return int(self.elem_nodes.shape[0])

    @property
    def n_nodes(self) -> int:
        return int(self.nodes.shape[0])

    @property
    def n_dofs(self) -> int:
        return 3 * self.n_nodes

    @property
    def n_local(self) -> int:
        return int(self.elem_nodes.shape[1])

    @property
    def n_quad(self) -> int:
        """Default tet quadrature size: order max(2(p-1), 1) (AssemblerUtils.cpp:225-228)."""
        return 1 if self.order == 1 else 4


def _first_encounter(prims: np.ndarray, n_prims: int) -> np.ndarray:
    """Map primitive ids to node ids in first-encounter order over ``prims`` (row-major)."""
    flat = prims.reshape(-1)
    uniq, first = np.unique(flat, return_index=True)
    rank = np.empty(uniq.shape[0], dtype=np.int64)
    rank[np.argsort(first, kind="stable")] = np.arange(uniq.shape[0], dtype=np.int64)
    node_of_prim = np.full(n_prims, -1, dtype=np.int64)
    node_of_prim[uniq] = rank
    return node_of_prim


def kuhn_box(nx: int, ny: int, nz: int, order: int = 1, size=(1.0, 1.0, 1.0),
             numbering: str = "first_encounter") -> TetMesh:
    """Kuhn 6-tet split of an ``nx x ny x nz`` box grid spanning ``[0, size]``.

    ``numbering`` is ``"first_encounter"`` (PolyFEM's rule) or ``"natural"``
    (vertex nodes keep their grid ids, edge nodes follow in edge order).
    """
    if order not in (1, 2):
        raise ValueError("order must be 1 or 2")
    nx, ny, nz = int(nx), int(ny), int(nz)
    hx, hy, hz = size[0] / nx, size[1] / ny, size[2] / nz
    sx, sy = nx + 1, (nx + 1) * (ny + 1)
    nv = (nx + 1) * (ny + 1) * (nz + 1)

    gi, gj, gk = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    vid = (gi + sx * gj + sy * gk).reshape(-1)
    vertices = np.empty((nv, 3), dtype=np.float64)
    vertices[vid, 0] = gi.reshape(-1) * hx
    vertices[vid, 1] = gj.reshape(-1) * hy
    vertices[vid, 2] = gk.reshape(-1) * hz

    # cells, i fastest
    ck, cj, ci = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ci, cj, ck = ci.reshape(-1), cj.reshape(-1), ck.reshape(-1)
    base = ci + sx * cj + sy * ck
    step = np.array([1, sx, sy], dtype=np.int64)
    ev = np.empty((base.shape[0], 6, 4), dtype=np.int64)
    for p, (perm, odd) in enumerate(zip(_PERMS, _PERM_ODD)):
        v0 = base
        v1 = v0 + step[perm[0]]
        v2 = v1 + step[perm[1]]
        v3 = v2 + step[perm[2]]
        if odd:
            v1, v2 = v2, v1
        ev[:, p, 0], ev[:, p, 1], ev[:, p, 2], ev[:, p, 3] = v0, v1, v2, v3
    elem_vertices = ev.reshape(-1, 4)
    m = tet_mesh(vertices, elem_vertices, order=order, numbering=numbering, reorient=False)
    m.grid = (nx, ny, nz)
    m.size = tuple(float(s) for s in size)
    return m


def tet_mesh(vertices, tets, order: int = 1, numbering: str = "first_encounter", reorient: bool = True) -> TetMesh:
    """A P1 or P2 mesh over the tets ``tets`` (E x 4 vertex ids) of ``vertices``.

    Nodes are numbered by PolyFEM's first-encounter rule (or ``"natural"``:
    vertex nodes keep their vertex ids, edge nodes follow in edge order); P2
    edge nodes sit at the edge midpoints (P1 geometry).  With ``reorient``, a
    negatively oriented tet has its vertices 1 and 2 swapped (the reference
    expects positive orientation); a degenerate tet is an error.  Vertices no
    tet uses get no node.
    """
    if order not in (1, 2):
        raise ValueError("order must be 1 or 2")
    vertices = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
    elem_vertices = np.array(tets, dtype=np.int64).reshape(-1, 4)
    nv = vertices.shape[0]
    if elem_vertices.size and (elem_vertices.min() < 0 or elem_vertices.max() >= nv):
        raise ValueError("a tet references a vertex out of range")
    if reorient:
        X = vertices[elem_vertices]
        det = np.einsum("ij,ij->i", np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]), X[:, 3] - X[:, 0])
        if np.any(det == 0) or not np.all(np.isfinite(det)):
            raise ValueError("degenerate tet")
        neg = det < 0
        elem_vertices[neg, 1], elem_vertices[neg, 2] = elem_vertices[neg, 2].copy(), elem_vertices[neg, 1].copy()

    if order == 1:
        prims = elem_vertices
        n_prims = nv
        prim_coords = vertices
    else:
        a = elem_vertices[:, TET_EDGES[:, 0]]
        b = elem_vertices[:, TET_EDGES[:, 1]]
        keys = np.minimum(a, b) * nv + np.maximum(a, b)
        ukeys, edge_idx = np.unique(keys.reshape(-1), return_inverse=True)
        edge_idx = edge_idx.reshape(keys.shape)
        prims = np.concatenate([elem_vertices, nv + edge_idx], axis=1)
        n_prims = nv + ukeys.shape[0]
        ea, eb = ukeys // nv, ukeys % nv
        prim_coords = np.concatenate([vertices, 0.5 * (vertices[ea] + vertices[eb])], axis=0)

    if numbering == "first_encounter":
        node_of_prim = _first_encounter(prims, n_prims)
        used = node_of_prim >= 0
        nodes = np.empty((int(used.sum()), 3), dtype=np.float64)
        nodes[node_of_prim[used]] = prim_coords[used]
    elif numbering == "natural":
        node_of_prim = np.arange(n_prims, dtype=np.int64)
        nodes = np.array(prim_coords, dtype=np.float64)
    else:
        raise ValueError(f"unknown numbering {numbering!r}")
    elem_nodes = node_of_prim[prims]
    return TetMesh(order=order, vertices=vertices, elem_vertices=elem_vertices.astype(np.int32),
                   nodes=nodes, elem_nodes=elem_nodes.astype(np.int32))


def lame_from_young(E, nu):
    """Lamé parameters from Young's modulus and Poisson ratio in 3D (MatParams.cpp:11-22)."""
    E = np.asarray(E, dtype=np.float64)
    nu = np.asarray(nu, dtype=np.float64)
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    return lam, mu


def smooth_displacement(nodes: np.ndarray, amplitude: float = 0.05, noise: float = 1e-3,
                        seed: int = BASE_SEED) -> np.ndarray:
    """u_d(X) = A sin(2 pi x) sin(2 pi y) sin(2 pi z) for d = 0, 1, 2, plus N(0, noise^2) per dof.

    Returned in PolyFEM's interleaved dof layout ``dof = node * 3 + d``.
    """
    rng = np.random.default_rng(seed)
    s = amplitude * np.sin(2 * np.pi * nodes[:, 0]) * np.sin(2 * np.pi * nodes[:, 1]) * np.sin(2 * np.pi * nodes[:, 2])
    u = np.repeat(s[:, None], 3, axis=1)
    if noise > 0:
        u = u + rng.normal(0.0, noise, size=u.shape)
    return np.ascontiguousarray(u.reshape(-1))


def random_smooth_field(nodes: np.ndarray, amplitude: float = 0.5, seed: int = BASE_SEED, modes: int = 6,
                        max_wavenumber: int = 2) -> np.ndarray:
    """A seeded smooth random vector field: per component, a sum of ``modes``
    sines sin(2 pi k.X + phase) with integer wave vectors |k_i| <= max_wavenumber
    and N(0, 1) weights, scaled so that max |u| = amplitude.  Dof layout
    ``node * 3 + d``."""
    rng = np.random.default_rng(seed)
    u = np.zeros((nodes.shape[0], 3))
    for d in range(3):
        for _ in range(modes):
            k = rng.integers(-max_wavenumber, max_wavenumber + 1, size=3)
            if not k.any():
                k[rng.integers(3)] = 1
            u[:, d] += rng.normal() * np.sin(2 * np.pi * nodes @ k + rng.uniform(0, 2 * np.pi))
    u *= amplitude / max(np.abs(u).max(), 1e-300)
    return np.ascontiguousarray(u.reshape(-1))


def random_lame(n_elements: int, seed: int = BASE_SEED + 1, E_mean: float = 1e6, E_sigma: float = 0.5,
                nu_range=(0.3, 0.45)):
    """Per-element E ~ logNormal(ln E_mean, E_sigma), nu ~ U(nu_range); returns (lambda, mu)."""
    rng = np.random.default_rng(seed)
    E = np.exp(rng.normal(np.log(E_mean), E_sigma, size=n_elements))
    nu = rng.uniform(nu_range[0], nu_range[1], size=n_elements)
    return lame_from_young(E, nu)


def face_nodes(mesh: TetMesh, axis: int, value: float, tol: float = 1e-12) -> np.ndarray:
    """Ids of nodes lying on the plane ``X[axis] == value``."""
    return np.nonzero(np.abs(mesh.nodes[:, axis] - value) <= tol)[0].astype(np.int32)
