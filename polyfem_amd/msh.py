"""Gmsh ``.msh`` input for tetrahedral meshes (formats 2.2 and 4.1, ASCII or binary).

Follows the reference's reader (``io/MshReader.cpp:39-230``, which parses with
mshio and keeps the volume elements):

* node tags map to vertex indices as ``tag - 1`` when the tags are exactly
  ``1..n``, otherwise in file order (the reference warns that this condenses
  the numbering);
* every tetrahedron type (4, 11, 29, 30, 31 nodes 4/10/20/35/56) contributes
  its first four nodes, the vertices, as the element's cell;
* the body id of an element is the first physical tag of its volume entity
  (format 4.1) or its first tag (format 2.2), 0 when there is none.

Binary files (either byte order, detected from the format's endianness
word) follow the same rules; 2-D meshes are rejected with an error.
"""
from __future__ import annotations

import numpy as np

from .mesh import TetMesh, tet_mesh

#: gmsh element types that are tetrahedra, with their node counts
TET_TYPES = {4: 4, 11: 10, 29: 20, 30: 35, 31: 56}
#: node counts of the other gmsh element types a volume file may hold (skipped)
OTHER_TYPES = {1: 2, 2: 3, 3: 4, 5: 8, 6: 6, 7: 5, 8: 3, 9: 6, 10: 9, 12: 27, 13: 18, 14: 14, 15: 1,
               16: 8, 17: 20, 18: 15, 19: 13, 20: 9, 21: 10, 22: 12, 23: 15, 24: 15, 25: 21, 26: 4,
               27: 5, 28: 6, 92: 64, 93: 125}


class _Lines:
    def __init__(self, text: str):
        self.lines = [ln.strip() for ln in text.splitlines()]
        self.i = 0

    def next(self) -> str:
        while self.i < len(self.lines) and not self.lines[self.i]:
            self.i += 1
        if self.i >= len(self.lines):
            raise ValueError("unexpected end of .msh file")
        ln = self.lines[self.i]
        self.i += 1
        return ln

    def ints(self):
        return [int(t) for t in self.next().split()]

    def section(self, name: str):
        """Advance to the line after ``$name``; False if it does not exist."""
        for j in range(self.i, len(self.lines)):
            if self.lines[j] == "$" + name:
                self.i = j + 1
                return True
        return False


def _node_index(tags, n_nodes):
    """tag -> vertex index: tag - 1 when the tags are exactly 1..n, else file order."""
    tags = np.asarray(tags, dtype=np.int64)
    max_tag = int(tags.max()) if tags.size else 0
    lut = np.full(max_tag + 1, -1, dtype=np.int64)
    if max_tag == n_nodes:
        lut[tags] = tags - 1
    else:
        lut[tags] = np.arange(tags.size)
    return lut


def _read_22(L: _Lines):
    if not L.section("Nodes"):
        raise ValueError("no $Nodes section")
    n = int(L.next())
    tags = np.empty(n, dtype=np.int64)
    xyz = np.empty((n, 3))
    for k in range(n):
        t = L.next().split()
        tags[k] = int(t[0])
        xyz[k] = [float(t[1]), float(t[2]), float(t[3])]
    lut = _node_index(tags, n)
    V = np.empty((n, 3))
    V[lut[tags]] = xyz
    if not L.section("Elements"):
        raise ValueError("no $Elements section")
    ne = int(L.next())
    tets, body = [], []
    for _ in range(ne):
        t = [int(v) for v in L.next().split()]
        etype, ntags = t[1], t[2]
        if etype in TET_TYPES:
            nodes = t[3 + ntags:3 + ntags + TET_TYPES[etype]]
            tets.append(lut[nodes[:4]])
            body.append(t[3] if ntags > 0 else 0)
        elif etype not in OTHER_TYPES:
            raise ValueError(f"unknown gmsh element type {etype}")
    return V, tets, body


def _read_41(L: _Lines):
    volume_phys = {}
    if L.section("Entities"):
        npt, ncu, nsu, nvo = L.ints()
        for _ in range(npt):
            L.next()
        for _ in range(ncu + nsu):
            L.next()
        for _ in range(nvo):
            t = L.next().split()
            nphys = int(t[7])
            volume_phys[int(t[0])] = int(t[8]) if nphys > 0 else 0
    if not L.section("Nodes"):
        raise ValueError("no $Nodes section")
    nblocks, n, _, _ = L.ints()
    tags = np.empty(n, dtype=np.int64)
    xyz = np.empty((n, 3))
    k = 0
    for _ in range(nblocks):
        _, _, parametric, nb = L.ints()
        if parametric:
            raise ValueError("parametric node coordinates are not supported")
        for j in range(nb):
            tags[k + j] = int(L.next())
        for j in range(nb):
            t = L.next().split()
            xyz[k + j] = [float(t[0]), float(t[1]), float(t[2])]
        k += nb
    lut = _node_index(tags, n)
    V = np.empty((n, 3))
    V[lut[tags]] = xyz
    if not L.section("Elements"):
        raise ValueError("no $Elements section")
    nblocks, _, _, _ = L.ints()
    tets, body = [], []
    for _ in range(nblocks):
        dim, etag, etype, nb = L.ints()
        if etype in TET_TYPES and dim == 3:
            for _ in range(nb):
                t = L.ints()
                tets.append(lut[t[1:5]])
                body.append(volume_phys.get(etag, 0))
        elif etype in TET_TYPES or etype in OTHER_TYPES:
            for _ in range(nb):
                L.next()
        else:
            raise ValueError(f"unknown gmsh element type {etype}")
    return V, tets, body


class _Bytes:
    """Cursor over a binary .msh file: section headers and counts are text
    lines, the payload is native binary data in the file's byte order."""

    def __init__(self, raw: bytes, order: str, size_t: int):
        self.raw = raw
        self.i = 0
        self.o = order
        self.st = {4: "u4", 8: "u8"}[size_t]

    def section(self, name: str) -> bool:
        key = b"$" + name.encode() + b"\n"
        j = self.raw.find(key, self.i)
        if j < 0:
            return False
        self.i = j + len(key)
        return True

    def line(self) -> str:
        j = self.raw.find(b"\n", self.i)
        if j < 0:
            raise ValueError("unexpected end of .msh file")
        ln = self.raw[self.i:j].decode("ascii").strip()
        self.i = j + 1
        return ln if ln else self.line()

    def read(self, kind: str, n: int = 1) -> np.ndarray:
        dt = np.dtype(self.o + (self.st if kind == "size_t" else kind))
        if self.i + dt.itemsize * n > len(self.raw):
            raise ValueError("unexpected end of .msh file")
        a = np.frombuffer(self.raw, dtype=dt, count=n, offset=self.i)
        self.i += dt.itemsize * n
        return a.astype(np.float64 if kind == "f8" else np.int64)

    def one(self, kind: str) -> int:
        return int(self.read(kind, 1)[0])


def _read_22_binary(B: _Bytes):
    if not B.section("Nodes"):
        raise ValueError("no $Nodes section")
    n = int(B.line())
    rec = np.dtype([("tag", B.o + "i4"), ("xyz", B.o + "f8", 3)])
    if B.i + rec.itemsize * n > len(B.raw):
        raise ValueError("unexpected end of .msh file")
    a = np.frombuffer(B.raw, dtype=rec, count=n, offset=B.i)
    B.i += rec.itemsize * n
    tags = a["tag"].astype(np.int64)
    lut = _node_index(tags, n)
    V = np.empty((n, 3))
    V[lut[tags]] = a["xyz"]
    if not B.section("Elements"):
        raise ValueError("no $Elements section")
    ne = int(B.line())
    tets, body = [], []
    done = 0
    while done < ne:
        etype, nfollow, ntags = (int(v) for v in B.read("i4", 3))
        if etype in TET_TYPES:
            nn = TET_TYPES[etype]
        elif etype in OTHER_TYPES:
            nn = OTHER_TYPES[etype]
        else:
            raise ValueError(f"unknown gmsh element type {etype}")
        blk = B.read("i4", nfollow * (1 + ntags + nn)).reshape(nfollow, 1 + ntags + nn)
        if etype in TET_TYPES:
            tets.extend(lut[blk[:, 1 + ntags:5 + ntags]])
            body.extend(blk[:, 1].tolist() if ntags > 0 else [0] * nfollow)
        done += nfollow
    return V, tets, body


def _read_41_binary(B: _Bytes):
    volume_phys = {}
    if B.section("Entities"):
        npt, ncu, nsu, nvo = (int(v) for v in B.read("size_t", 4))
        for _ in range(npt):
            B.one("i4")
            B.read("f8", 3)
            B.read("i4", B.one("size_t"))
        for kind in range(3):
            for _ in range((ncu, nsu, nvo)[kind]):
                tag = B.one("i4")
                B.read("f8", 6)
                phys = B.read("i4", B.one("size_t"))
                B.read("i4", B.one("size_t"))
                if kind == 2:
                    volume_phys[tag] = int(phys[0]) if phys.size else 0
    if not B.section("Nodes"):
        raise ValueError("no $Nodes section")
    nblocks, n, _, _ = (int(v) for v in B.read("size_t", 4))
    tags = np.empty(n, dtype=np.int64)
    xyz = np.empty((n, 3))
    k = 0
    for _ in range(nblocks):
        _, _, parametric = B.read("i4", 3)
        nb = B.one("size_t")
        if parametric:
            raise ValueError("parametric node coordinates are not supported")
        tags[k:k + nb] = B.read("size_t", nb)
        xyz[k:k + nb] = B.read("f8", 3 * nb).reshape(nb, 3)
        k += nb
    lut = _node_index(tags, n)
    V = np.empty((n, 3))
    V[lut[tags]] = xyz
    if not B.section("Elements"):
        raise ValueError("no $Elements section")
    nblocks = int(B.read("size_t", 4)[0])
    tets, body = [], []
    for _ in range(nblocks):
        dim, etag, etype = (int(v) for v in B.read("i4", 3))
        nb = B.one("size_t")
        if etype in TET_TYPES:
            nn = TET_TYPES[etype]
        elif etype in OTHER_TYPES:
            nn = OTHER_TYPES[etype]
        else:
            raise ValueError(f"unknown gmsh element type {etype}")
        blk = B.read("size_t", nb * (1 + nn)).reshape(nb, 1 + nn)
        if etype in TET_TYPES and dim == 3:
            tets.extend(lut[blk[:, 1:5]])
            body.extend([volume_phys.get(etag, 0)] * nb)
    return V, tets, body


def read_msh(path: str):
    """(vertices (N, 3), tets (E, 4) vertex indices, body_ids (E,)) of a tet mesh."""
    with open(path, "rb") as f:
        raw = f.read()
    head = raw.find(b"$MeshFormat\n")
    if head < 0:
        raise ValueError(f"{path}: no $MeshFormat section")
    j = raw.find(b"\n", head + 12)
    fmt = raw[head + 12:j].decode("ascii", errors="replace").split()
    if len(fmt) < 3:
        raise ValueError(f"{path}: malformed $MeshFormat")
    version, file_type, data_size = fmt[0], int(fmt[1]), int(fmt[2])
    if file_type == 1:
        # the endianness word: the integer 1 in the writer's byte order
        word = raw[j + 1:j + 5]
        if len(word) < 4:
            raise ValueError(f"{path}: truncated binary header")
        order = "<" if int.from_bytes(word, "little") == 1 else ">"
        B = _Bytes(raw, order, data_size)
        B.i = j + 5
        if version.startswith("2"):
            V, tets, body = _read_22_binary(B)
        elif version.startswith("4"):
            V, tets, body = _read_41_binary(B)
        else:
            raise ValueError(f"{path}: unsupported .msh version {version}")
    else:
        try:
            text = raw.decode("ascii")
        except UnicodeDecodeError:
            raise ValueError(f"{path}: non-ASCII bytes in an ASCII .msh file") from None
        L = _Lines(text)
        L.section("MeshFormat")
        L.next()
        if version.startswith("2"):
            V, tets, body = _read_22(L)
        elif version.startswith("4"):
            V, tets, body = _read_41(L)
        else:
            raise ValueError(f"{path}: unsupported .msh version {version}")
    if not tets:
        raise ValueError(f"{path}: no tetrahedra (only 3-D tet meshes are supported)")
    tets = np.array(tets, dtype=np.int64)
    if tets.min() < 0:
        raise ValueError(f"{path}: an element references an unknown node tag")
    return V, tets, np.array(body, dtype=np.int32)


def load_msh(path: str, order: int = 1, numbering: str = "first_encounter") -> TetMesh:
    """A P1 or P2 :class:`TetMesh` over the tets of a ``.msh`` file (P1
    geometry; P2 edge nodes at the edge midpoints).  The body ids are kept in
    ``mesh.body_ids``."""
    V, tets, body = read_msh(path)
    m = tet_mesh(V, tets, order=order, numbering=numbering)
    m.body_ids = body
    return m
