"""Test meshes beyond the Kuhn box."""
import numpy as np

from polyfem_amd.mesh import tet_mesh


def star_mesh(k=30, order=1, seed=0):
    """A double fan of 2k tets around the axis through a centre vertex: the
    centre lies in all 2k tets, the two apexes in k each, so the Hessian
    blocks of the centre and of the axis edges have far more contributions
    than on a lattice (vertices perturbed, numbering shuffled)."""
    rng = np.random.default_rng(seed)
    th = np.linspace(0.0, 2.0 * np.pi, k, endpoint=False) + rng.uniform(-0.2, 0.2, k) * np.pi / k
    ring = np.stack([np.cos(th), np.sin(th), rng.uniform(-0.1, 0.1, k)], axis=1)
    V = np.concatenate([[[0.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.0, 0.0, -1.0]], ring])
    r = 3 + np.arange(k)
    rn = 3 + (np.arange(k) + 1) % k
    T = np.concatenate([np.stack([np.zeros(k, int), r, rn, np.ones(k, int)], 1),
                        np.stack([np.zeros(k, int), rn, r, np.full(k, 2)], 1)])
    perm = rng.permutation(V.shape[0])
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    return tet_mesh(V[perm], inv[T][rng.permutation(T.shape[0])], order=order)
