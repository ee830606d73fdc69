"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the mesh generator and the FP64 oracle still reproduce the stored
fixtures (which were cross-checked by central finite differences when they
were written).  GPU: libpfgpu reproduces them through the C-ABI, bit-exact
sparsity, 1e-10 relative in FP64 and the FP32 tolerances in mixed precision.
"""
import glob
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from polyfem_amd.mesh import kuhn_box

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))
MAT = {"NeoHookean": po.NEOHOOKEAN, "LinearElasticity": po.LINEAR_ELASTICITY}
MIXED_TOL = {"energy": 2e-7, "gradient": 1e-6, "hessian": 1e-6, "hvp": 1e-6, "vjp": 1e-6}


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def load(path):
    with np.load(path, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def mesh_of(d):
    m = kuhn_box(2, 2, 2, order=int(d["order"]))
    assert np.array_equal(m.elem_nodes, d["elem_nodes"]) and np.array_equal(m.elem_vertices, d["elem_vertices"])
    assert np.array_equal(m.vertices, d["vertices"]) and np.array_equal(m.nodes, d["nodes"])
    return m


def test_fixtures_present():
    assert len(FIXTURES) == 4


@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_oracle_reproduces_golden(path):
    d = load(path)
    m = mesh_of(d)
    o = po.Oracle(m, d["lam"], d["mu"], material=MAT[str(d["material"])])
    x = d["x"]
    assert rel(o.energy(x), d["energy"]) < 1e-12
    assert rel(o.energy_per_element(x), d["energy_per_element"]) < 1e-12
    assert rel(o.gradient(x), d["gradient"]) < 1e-12
    outer, inner, values = o.hessian(x)
    assert np.array_equal(outer, d["outer"]) and np.array_equal(inner, d["inner"])
    assert rel(values, d["values"]) < 1e-12
    assert rel(o.material_vjp(x, d["p"]), d["material_vjp"]) < 1e-12
    assert rel(o.shape_vjp(x, d["p"]), d["shape_vjp"]) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("path_kind", ["fused", "plain", "factored"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_library_reproduces_golden(path, precision, path_kind):
    from polyfem_amd import NLAssembler
    d = load(path)
    m = mesh_of(d)
    print(f"golden {os.path.basename(path)} {precision} {path_kind}", flush=True)
    a = NLAssembler(str(d["material"]), m, lame=(d["lam"], d["mu"]), precision=precision, path=path_kind)
    x = d["x"]
    tol = (lambda k: 1e-10) if precision == "fp64" else (lambda k: MIXED_TOL[k])
    e, g, vals = a.assemble_energy_gradient_hessian(x)
    outer, inner = a.hessian_pattern()
    assert np.array_equal(outer, d["outer"]) and np.array_equal(inner, d["inner"])
    assert rel(e, d["energy"]) < tol("energy")
    assert rel(a.assemble_energy_per_element(x), d["energy_per_element"]) < tol("energy")
    assert rel(g, d["gradient"]) < tol("gradient")
    assert rel(vals, d["values"]) < tol("hessian")
    assert rel(a.hessian_vector_product(x, d["v"]), d["hv"]) < tol("hvp")
    assert rel(a.force_material_derivative(x, d["p"]), d["material_vjp"]) < tol("vjp")
    assert rel(a.force_shape_derivative(x, d["p"]), d["shape_vjp"]) < tol("vjp")
