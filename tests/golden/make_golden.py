"""Writes the golden fixtures of the assembly path (test data, not product code).

Each fixture is a small Kuhn box (2x2x2 cells, 48 tets) with seeded per-element
Lame parameters, a seeded displacement and seeded direction vectors, for P1/P2
and NeoHookean/LinearElasticity.  The expected outputs come from the FP64
oracle (oracle/oracle.c, a restatement of the reference assembler):
energy, per-element energies, gradient, the CSC Hessian (outer, inner,
values), H v, the material VJP and the shape VJP.  Before writing, every
fixture is cross-checked by central finite differences (gradient of the
energy, Hessian of the gradient, material VJP of the gradient), so the stored
numbers are pinned independently of the oracle's own formulas.

Run:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyoracle as po  # noqa: E402
from polyfem_amd.mesh import BASE_SEED, kuhn_box, random_lame, smooth_displacement  # noqa: E402

CASES = [(1, "NeoHookean"), (2, "NeoHookean"), (1, "LinearElasticity"), (2, "LinearElasticity")]
MAT = {"NeoHookean": po.NEOHOOKEAN, "LinearElasticity": po.LINEAR_ELASTICITY}


def fixture_name(order, material):
    return f"box2_p{order}_{'nh' if material == 'NeoHookean' else 'le'}.npz"


def fd_check(o, m, x, lam, mu, material, rng):
    """Central finite differences of the oracle on random directions."""
    h = 1e-6
    d = rng.normal(size=m.n_dofs)
    g = o.gradient(x)
    fd = (o.energy(x + h * d) - o.energy(x - h * d)) / (2 * h)
    assert abs(fd - g @ d) <= 1e-5 * max(1.0, abs(g @ d)), ("gradient FD", fd, g @ d)
    H = o.hessian_scipy(x)
    fdv = (o.gradient(x + h * d) - o.gradient(x - h * d)) / (2 * h)
    assert np.max(np.abs(fdv - H @ d)) <= 1e-5 * np.max(np.abs(H @ d)), "Hessian FD"
    p = rng.normal(size=m.n_dofs)
    dl = rng.normal(size=m.n_elements) * 0.1 * lam
    dm = rng.normal(size=m.n_elements) * 0.1 * mu
    gp = po.Oracle(m, lam + h * dl, mu + h * dm, material=MAT[material]).gradient(x)
    gm = po.Oracle(m, lam - h * dl, mu - h * dm, material=MAT[material]).gradient(x)
    fdm = -p @ (gp - gm) / (2 * h)  # material term = -p^T dg/dtheta
    vjp = o.material_vjp(x, p)
    lin = vjp[:m.n_elements] @ dl + vjp[m.n_elements:] @ dm
    assert abs(fdm - lin) <= 1e-5 * max(1.0, abs(lin)), ("material VJP FD", fdm, lin)


def make(order, material):
    m = kuhn_box(2, 2, 2, order=order)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 11 * order + (material == "NeoHookean"))
    x = smooth_displacement(m.nodes, 0.05, 1e-3, BASE_SEED + 7 * order)
    rng = np.random.default_rng(BASE_SEED + 5 + order)
    v = rng.normal(size=m.n_dofs)
    p = rng.normal(size=m.n_dofs)
    o = po.Oracle(m, lam, mu, material=MAT[material])
    fd_check(o, m, x, lam, mu, material, np.random.default_rng(BASE_SEED + 99))
    outer, inner, values = o.hessian(x)
    H = o.hessian_scipy(x)
    return dict(order=np.int32(order), material=np.array(material),
                elem_nodes=m.elem_nodes.astype(np.int32), nodes=m.nodes, vertices=m.vertices,
                elem_vertices=m.elem_vertices.astype(np.int32), lam=lam, mu=mu, x=x, v=v, p=p,
                energy=np.float64(o.energy(x)), energy_per_element=o.energy_per_element(x),
                gradient=o.gradient(x), outer=outer, inner=inner, values=values, hv=H @ v,
                material_vjp=o.material_vjp(x, p), shape_vjp=o.shape_vjp(x, p))


def main():
    po.build()
    for order, material in CASES:
        path = os.path.join(HERE, fixture_name(order, material))
        np.savez_compressed(path, **make(order, material))
        print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
