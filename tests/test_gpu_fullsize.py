"""The 40^3 x 6 P2 NeoHookean state at full size (384,000 tets, 1,594,323 dofs,
134,475,849 Hessian non-zeros; the benchmarked workload) against the FP64
oracle (runs on an MI355X).

The energy and the whole gradient come from the oracle's threaded assembly
(the reference-algorithm restatement, checked against the single-thread
oracle in test_oracle.py); Hessian values and H v are checked on 1,200
random dof rows, each rebuilt from the oracle's element Hessians of the
elements that touch its node.  Bars: FP64 1e-10 relative; mixed precision
the stated tolerances (relative to the array's largest entry)."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from polyfem_amd import NLAssembler
from polyfem_amd.mesh import BASE_SEED, kuhn_box, random_lame, smooth_displacement

pytestmark = pytest.mark.gpu

TOL = {"fp64": {"energy": 1e-10, "gradient": 1e-10, "hessian": 1e-10, "hvp": 1e-10},
       "mixed": {"energy": 1e-6, "gradient": 2e-6, "hessian": 2e-5, "hvp": 2e-5}}


@pytest.fixture(scope="module")
def state():
    m = kuhn_box(40, 40, 40, order=2)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 1)
    x = smooth_displacement(m.nodes, 0.05, 1e-3, BASE_SEED)
    o = po.Oracle(m, lam, mu)
    # a few GB of host memory at most: the threaded oracle keeps per-thread
    # gradient vectors only (no Hessian here)
    b = po.Baseline(o, min(16, os.cpu_count() or 1))
    return m, lam, mu, x, o, b.energy(x), b.gradient(x)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_full_size_state(state, precision):
    m, lam, mu, x, o, e_ref, g_ref = state
    assert m.n_elements == 384000 and m.n_dofs == 1594323
    a = NLAssembler("NeoHookean", m, lame=(lam, mu), precision=precision)
    assert a.nnz == 134475849
    print(f"full size {precision}", flush=True)
    tol = TOL[precision]
    e, g, vals = a.assemble_energy_gradient_hessian(x)
    assert abs(e - e_ref) <= tol["energy"] * abs(e_ref)
    assert np.abs(g - g_ref).max() <= tol["gradient"] * np.abs(g_ref).max()
    # Hessian rows and H v on 1,200 random dofs
    rng = np.random.default_rng(BASE_SEED + 1000)
    rows = np.sort(rng.choice(m.n_dofs, size=1200, replace=False))
    outer, inner = a.hessian_pattern()
    v = rng.normal(size=m.n_dofs)
    y = a.hessian_vector_product(x, v)
    nodes = np.unique(rows // 3)
    # elements touching each node
    en = m.elem_nodes
    hit = np.isin(en, nodes)
    els = np.flatnonzero(hit.any(axis=1))
    dense = {}
    for e in els:
        He = o.element_hessian(int(e), x)
        dofs = (3 * en[e][:, None] + np.arange(3)).ravel()
        for a_loc in np.flatnonzero(hit[e]):
            for d in range(3):
                r = 3 * en[e][a_loc] + d
                row = dense.setdefault(r, {})
                for c, val in zip(dofs, He[3 * a_loc + d]):
                    row[c] = row.get(c, 0.0) + val
    hmax = max(abs(val) for row in dense.values() for val in row.values())
    herr, yerr, ymax = 0.0, 0.0, 0.0
    for r in rows:
        ref = dense[r]
        cols = inner[outer[r]:outer[r + 1]]
        got = np.asarray(vals[outer[r]:outer[r + 1]], dtype=np.float64)
        assert sorted(ref) == list(cols)  # the row's pattern, bit-exact
        herr = max(herr, np.abs(got - np.array([ref[c] for c in cols])).max())
        yr = sum(val * v[c] for c, val in ref.items())
        yerr = max(yerr, abs(y[r] - yr))
        ymax = max(ymax, abs(yr))
    assert herr <= tol["hessian"] * hmax, herr / hmax
    assert yerr <= tol["hvp"] * ymax, yerr / ymax
