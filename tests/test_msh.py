"""Gmsh .msh input (polyfem_amd.msh) and general tet meshes (mesh.tet_mesh).

Hand-written files pin the format details (node-tag mapping, skipped
element types, body ids from physical tags, tet types with more than four
nodes); binary files are written per the gmsh format description in both
byte orders; a mesh written out and read back must give the same P1/P2 mesh
as the generator, node numbering included; a two-tet P2 mesh pins the node
numbering against ids derived by hand from the reference's rule.
"""
import numpy as np
import pytest

from polyfem_amd.mesh import kuhn_box, tet_mesh
from polyfem_amd.msh import load_msh, read_msh

# two tets sharing the face (1, 2, 3), plus a surface triangle gmsh writes alongside
MSH22 = """$MeshFormat
2.2 0 8
$EndMeshFormat
$Nodes
5
1 0 0 0
2 1 0 0
3 0 1 0
4 0 0 1
5 1 1 1
$EndNodes
$Elements
3
1 2 2 3 1 1 2 3
2 4 2 5 1 1 2 3 4
3 4 2 6 1 2 4 3 5
$EndElements
"""

MSH41 = """$MeshFormat
4.1 0 8
$EndMeshFormat
$Entities
0 0 1 1
1 0 0 0 1 1 0 0 0
1 0 0 0 1 1 1 1 7 1 1
$EndEntities
$Nodes
2 5 10 50
2 1 0 2
10
20
0 0 0
1 0 0
3 1 0 3
30
40
50
0 1 0
0 0 1
1 1 1
$EndNodes
$Elements
2 3 1 3
2 1 2 1
1 10 20 30
3 1 11 2
2 10 20 30 40 91 92 93 94 95 96
3 20 40 30 50 91 92 93 94 95 96
$EndElements
"""


def test_msh22(tmp_path):
    p = tmp_path / "two.msh"
    p.write_text(MSH22)
    V, T, body = read_msh(str(p))
    assert np.array_equal(V, [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]])
    assert np.array_equal(T, [[0, 1, 2, 3], [1, 3, 2, 4]])
    assert np.array_equal(body, [5, 6])


def test_msh41_tags_entities_and_tet10(tmp_path):
    """Node tags 10..50 (not 1..n) map in file order; tet10 elements keep
    their first four nodes; the volume's physical tag is the body id."""
    p = tmp_path / "two41.msh"
    p.write_text(MSH41)
    V, T, body = read_msh(str(p))
    assert np.array_equal(V, [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]])
    assert np.array_equal(T, [[0, 1, 2, 3], [1, 3, 2, 4]])
    assert np.array_equal(body, [7, 7])


def write_msh41_binary(path, V, T, order="<", phys=7, with_surface=True):
    """Binary MSH 4.1 per the gmsh format description: size_t = 8 bytes, an
    $Entities section with one volume (physical tag phys), one node block,
    a triangle block (skipped by the reader) and a tet block."""
    import struct
    st, i4, f8 = order + "Q", order + "i", order + "d"
    b = bytearray(b"$MeshFormat\n4.1 1 8\n")
    b += struct.pack(i4, 1) + b"\n$EndMeshFormat\n$Entities\n"
    b += struct.pack(order + "4Q", 1, 0, 1, 1)
    b += struct.pack(i4, 1) + struct.pack(order + "3d", 0, 0, 0) + struct.pack(st, 0)
    b += struct.pack(i4, 1) + struct.pack(order + "6d", 0, 0, 0, 1, 1, 1) + struct.pack(st, 0) + struct.pack(st, 0)
    b += struct.pack(i4, 1) + struct.pack(order + "6d", 0, 0, 0, 1, 1, 1) + struct.pack(st, 1) + struct.pack(i4, phys)
    b += struct.pack(st, 1) + struct.pack(i4, 1)
    b += b"\n$EndEntities\n$Nodes\n"
    n = len(V)
    b += struct.pack(order + "4Q", 1, n, 1, n) + struct.pack(order + "3i", 3, 1, 0) + struct.pack(st, n)
    b += struct.pack(order + f"{n}Q", *range(1, n + 1))
    b += struct.pack(order + f"{3 * n}d", *np.asarray(V, dtype=np.float64).ravel())
    b += b"\n$EndNodes\n$Elements\n"
    nt = len(T)
    blocks = 2 if with_surface else 1
    b += struct.pack(order + "4Q", blocks, nt + (1 if with_surface else 0), 1, nt + 1)
    if with_surface:
        b += struct.pack(order + "3i", 2, 1, 2) + struct.pack(st, 1) + struct.pack(order + "4Q", nt + 1, 1, 2, 3)
    b += struct.pack(order + "3i", 3, 1, 4) + struct.pack(st, nt)
    for e, t in enumerate(T):
        b += struct.pack(order + "5Q", e + 1, *(int(v) + 1 for v in t))
    b += b"\n$EndElements\n"
    with open(path, "wb") as f:
        f.write(bytes(b))


def write_msh22_binary(path, V, T, order="<"):
    """Binary MSH 2.2: ASCII counts, int tags and doubles, element blocks of
    (type, count, ntags) headers; two tags (physical 5, elementary 1)."""
    import struct
    i4 = order + "i"
    b = bytearray(b"$MeshFormat\n2.2 1 8\n") + struct.pack(i4, 1) + b"\n$EndMeshFormat\n"
    b += f"$Nodes\n{len(V)}\n".encode()
    for k, (x, y, z) in enumerate(np.asarray(V, dtype=np.float64)):
        b += struct.pack(order + "i3d", k + 1, x, y, z)
    b += b"\n$EndNodes\n"
    b += f"$Elements\n{len(T) + 1}\n".encode()
    b += struct.pack(order + "3i", 2, 1, 2) + struct.pack(order + "6i", 1, 5, 1, 1, 2, 3)
    b += struct.pack(order + "3i", 4, len(T), 2)
    for e, t in enumerate(T):
        b += struct.pack(order + "7i", e + 2, 5, 1, *(int(v) + 1 for v in t))
    b += b"\n$EndElements\n"
    with open(path, "wb") as f:
        f.write(bytes(b))


@pytest.mark.parametrize("order", ["<", ">"])
def test_binary_msh41(tmp_path, order):
    m = kuhn_box(3, 2, 2, order=2)
    p = tmp_path / "bin41.msh"
    write_msh41_binary(p, m.vertices, m.elem_vertices, order=order)
    V, T, body = read_msh(str(p))
    assert np.array_equal(V, m.vertices) and np.array_equal(T, m.elem_vertices)
    assert np.all(body == 7)
    r = load_msh(str(p), order=2)
    assert np.array_equal(r.elem_nodes, m.elem_nodes) and np.array_equal(r.nodes, m.nodes)


@pytest.mark.parametrize("order", ["<", ">"])
def test_binary_msh22(tmp_path, order):
    m = kuhn_box(2, 3, 2, order=1)
    p = tmp_path / "bin22.msh"
    write_msh22_binary(p, m.vertices, m.elem_vertices, order=order)
    V, T, body = read_msh(str(p))
    assert np.array_equal(V, m.vertices) and np.array_equal(T, m.elem_vertices)
    assert np.all(body == 5)


def test_truncated_binary_is_an_error(tmp_path):
    m = kuhn_box(1, 1, 1, order=1)
    p = tmp_path / "cut.msh"
    write_msh41_binary(p, m.vertices, m.elem_vertices)
    raw = p.read_bytes()
    cut = raw[:raw.find(b"$Elements") + 40]
    p.write_bytes(cut)
    with pytest.raises(ValueError):
        read_msh(str(p))


def test_p2_numbering_known_answer():
    """Node ids of a two-tet P2 mesh derived by hand from the reference's
    rule: cells in order; a cell's vertices, then its edges (v0v1, v1v2,
    v2v0, v0v3, v1v3, v2v3) get the next id at their first encounter
    (LagrangeBasis3d.cpp:238-330 with MeshNodes::node_id_from_primitive,
    MeshNodes.cpp:159-180; the cell loop LagrangeBasis3d.cpp:853-888)."""
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1.0]])
    T = np.array([[0, 1, 2, 3], [1, 2, 3, 4]])
    # both tets positively oriented, so no vertex swap applies
    for t in T:
        X = V[t]
        assert np.linalg.det(X[1:] - X[0]) > 0
    m = tet_mesh(V, T, order=2)
    # cell 0: vertices 0..3 -> 0..3; edges {0,1}->4 {1,2}->5 {2,0}->6 {0,3}->7 {1,3}->8 {2,3}->9
    # cell 1 (1,2,3,4): vertex 4 -> 10; edges {1,2}=5 {2,3}=9 {3,1}=8, then new {1,4}->11 {2,4}->12 {3,4}->13
    assert np.array_equal(m.elem_nodes, [[0, 1, 2, 3, 4, 5, 6, 7, 8, 9], [1, 2, 3, 10, 5, 9, 8, 11, 12, 13]])
    assert m.n_nodes == 14
    mid = lambda i, j: (V[i] + V[j]) / 2
    expect = np.array([V[0], V[1], V[2], V[3], mid(0, 1), mid(1, 2), mid(2, 0), mid(0, 3), mid(1, 3), mid(2, 3),
                       V[4], mid(1, 4), mid(2, 4), mid(3, 4)])
    assert np.array_equal(m.nodes, expect)
    # P1: the vertices in first-encounter order
    m1 = tet_mesh(V, T, order=1)
    assert np.array_equal(m1.elem_nodes, [[0, 1, 2, 3], [1, 2, 3, 4]])
    # first encounter differs from the vertex ids when the cells list them out of order
    m2 = tet_mesh(V, [[1, 2, 3, 4], [0, 1, 2, 3]], order=1)
    assert np.array_equal(m2.elem_nodes, [[0, 1, 2, 3], [4, 0, 1, 2]])


def write_msh41(path, V, T):
    with open(path, "w") as f:
        f.write("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n")
        f.write(f"$Nodes\n1 {len(V)} 1 {len(V)}\n3 1 0 {len(V)}\n")
        f.write("".join(f"{i + 1}\n" for i in range(len(V))))
        f.write("".join(f"{float(x)!r} {float(y)!r} {float(z)!r}\n" for x, y, z in V))
        f.write(f"$EndNodes\n$Elements\n1 {len(T)} 1 {len(T)}\n3 1 4 {len(T)}\n")
        f.write("".join(f"{e + 1} {a + 1} {b + 1} {c + 1} {d + 1}\n" for e, (a, b, c, d) in enumerate(T)))
        f.write("$EndElements\n")


@pytest.mark.parametrize("order", [1, 2])
def test_round_trip_matches_generator(tmp_path, order):
    m = kuhn_box(3, 2, 2, order=order)
    p = tmp_path / "box.msh"
    write_msh41(p, m.vertices, m.elem_vertices)
    r = load_msh(str(p), order=order)
    assert np.array_equal(r.vertices, m.vertices)
    assert np.array_equal(r.elem_vertices, m.elem_vertices)
    assert np.array_equal(r.elem_nodes, m.elem_nodes)
    assert np.array_equal(r.nodes, m.nodes)


def test_tet_mesh_reorients_and_skips_unused_vertices():
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [5, 5, 5.0]])
    m = tet_mesh(V, [[0, 2, 1, 3]], order=2)  # negative orientation; vertex 4 unused
    assert np.array_equal(m.elem_vertices, [[0, 1, 2, 3]])
    assert m.n_nodes == 10 and m.vertices.shape[0] == 5
    # P2 edge nodes at midpoints, in the local edge order (v0v1, v1v2, v2v0, v0v3, v1v3, v2v3)
    X = m.nodes[m.elem_nodes[0]]
    assert np.allclose(X[4:], [[0.5, 0, 0], [0.5, 0.5, 0], [0, 0.5, 0], [0, 0, 0.5], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    with pytest.raises(ValueError, match="degenerate"):
        tet_mesh(V, [[0, 1, 2, 2]])


@pytest.mark.gpu
@pytest.mark.parametrize("order", [1, 2])
def test_assembly_on_a_loaded_mesh_with_an_unused_vertex(tmp_path, order):
    """A file with a vertex that no tet uses: the assembly matches the mesh
    without it, and the shape VJP is zero at the unused vertex."""
    from polyfem_amd import NLAssembler
    from polyfem_amd.mesh import BASE_SEED, random_lame, smooth_displacement
    m = kuhn_box(2, 2, 2, order=order)
    V = np.concatenate([m.vertices, [[9.0, 9.0, 9.0]]])
    p = tmp_path / "extra.msh"
    write_msh41(p, V, m.elem_vertices)
    r = load_msh(str(p), order=order)
    assert r.vertices.shape[0] == m.vertices.shape[0] + 1 and r.n_nodes == m.n_nodes
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 80)
    x = smooth_displacement(m.nodes, 0.02, 1e-3, BASE_SEED + 81)
    print(f"loaded mesh order={order}", flush=True)
    a = NLAssembler("NeoHookean", m, lame=(lam, mu), precision="fp64")
    b = NLAssembler("NeoHookean", r, lame=(lam, mu), precision="fp64")
    assert a.assemble_energy(x) == b.assemble_energy(x)
    assert np.array_equal(a.assemble_gradient(x), b.assemble_gradient(x))
    pv = np.random.default_rng(BASE_SEED + 82).normal(size=m.n_dofs)
    sa, sb = a.force_shape_derivative(x, pv), b.force_shape_derivative(x, pv)
    assert np.array_equal(sb[:-3], sa) and np.all(sb[-3:] == 0.0)
