"""FP64 restatement of the reference's FixedCorotational material (test
infrastructure only: the checker for the HIP library, never the product).

Follows assembler/FixedCorotational.cpp:
  * psi = mu sum_i (sigma_i - 1)^2 + lambda/2 (prod sigma - 1)^2 over the
    singular values of F with the rotation-variant sign convention
    (utils/svd.hpp AutoFlipSVD: U, V rotations, the last singular value
    carries the sign of det F) (:589-596);
  * P = lambda (J - 1) cof F + 2 mu (F - U V^T) (:683-706);
  * the 9x9 stiffness from the singular-value derivatives and the 2x2 "B"
    blocks in the (U, V) frame (:708-826), entries indexed i + 3 j for F_ij;
  * element gradient  sum_q da G P^T and Hessian sum_q da D^T A D with
    D(3 k + j, 3 a + j) = G(a, k) (:322-436);
  * the adjoint terms through compute_dstress_dmu_dlambda (:493-546) and
    compute_stress_grad_multiply_mat (:438-464) in the generic formulas of
    ElasticForceDerivative.cpp.
The SVD is numpy's (LAPACK); the rest is written out here.
"""
from __future__ import annotations

import numpy as np

from . import pyoracle as po


def autoflip_svd(F):
    U, s, Vt = np.linalg.svd(F)
    V = Vt.T.copy()
    U = U.copy()
    s = s.copy()
    if np.linalg.det(U) < 0:
        U[:, 2] *= -1
        s[2] *= -1
    if np.linalg.det(V) < 0:
        V[:, 2] *= -1
        s[2] *= -1
    return U, s, V


def energy_density(F, lam, mu):
    s = np.linalg.svd(F, compute_uv=False)
    s = s.copy()
    if np.linalg.det(F) < 0:
        s[2] *= -1
    return mu * np.sum((s - 1) ** 2) + lam / 2 * (np.prod(s) - 1) ** 2


def cofactor(F):
    return np.stack([np.cross(F[:, 1], F[:, 2]), np.cross(F[:, 2], F[:, 0]), np.cross(F[:, 0], F[:, 1])], axis=1)


def stress(F, lam, mu):
    U, s, V = autoflip_svd(F)
    return lam * (np.prod(s) - 1) * cofactor(F) + 2 * mu * (F - U @ V.T)


def stiffness(F, lam, mu):
    """9x9 d P / d F, entries [i + 3 j, r + 3 s] = d P_ij / d F_rs (FixedCorotational.cpp:708-826)."""
    U, s, V = autoflip_svd(F)
    prod = np.prod(s)
    noI = np.array([s[1] * s[2], s[2] * s[0], s[0] * s[1]])
    dE = 2 * mu * (s - 1) + noI * lam * (prod - 1)
    d2 = np.diag(2 * mu + lam * noI * noI)
    d2[0, 1] = d2[1, 0] = lam * (s[2] * (prod - 1) + noI[0] * noI[1])
    d2[0, 2] = d2[2, 0] = lam * (s[1] * (prod - 1) + noI[0] * noI[2])
    d2[1, 2] = d2[2, 1] = lam * (s[0] * (prod - 1) + noI[2] * noI[1])
    tmp = lam / 2 * (prod - 1)
    left = [mu - tmp * s[2], mu - tmp * s[0], mu - tmp * s[1]]
    B = []
    for c in range(3):
        cp = (c + 1) % 3
        right = (dE[c] + dE[cp]) / (2 * max(s[c] + s[cp], 1e-12))
        B.append(np.array([[left[c] + right, left[c] - right], [left[c] - right, left[c] + right]]))
    M = np.zeros((9, 9))
    M[np.ix_([0, 4, 8], [0, 4, 8])] = d2
    M[1, 1], M[1, 3], M[3, 1], M[3, 3] = B[0][0, 0], B[0][0, 1], B[0][1, 0], B[0][1, 1]
    M[5, 5], M[5, 7], M[7, 5], M[7, 7] = B[1][0, 0], B[1][0, 1], B[1][1, 0], B[1][1, 1]
    M[2, 2], M[2, 6], M[6, 2], M[6, 6] = B[2][1, 1], B[2][1, 0], B[2][0, 1], B[2][0, 0]
    # M in the (U, V) frame: entry (3 a + b, 3 c + d) multiplies U[i, a] V[j, b] U[r, c] V[t, d]
    H4 = np.einsum("abcd,ia,jb,rc,td->ijrt", M.reshape(3, 3, 3, 3), U, V, U, V)
    return H4.transpose(1, 0, 3, 2).reshape(9, 9)


class FcrOracle:
    """Element loops of the FixedCorotational form on one mesh (FP64)."""

    # the material's energy density, stress and 9x9 stiffness of (F, e)
    def _psi(self, F, e):
        return energy_density(F, self.lam[e], self.mu[e])

    def _P(self, F, e):
        return stress(F, self.lam[e], self.mu[e])

    def _A(self, F, e):
        return stiffness(F, self.lam[e], self.mu[e])

    def __init__(self, mesh, lam, mu, quad_order=0):
        self.mesh = mesh
        self.lam = np.broadcast_to(np.asarray(lam, dtype=np.float64), (mesh.n_elements,))
        self.mu = np.broadcast_to(np.asarray(mu, dtype=np.float64), (mesh.n_elements,))
        self.order = mesh.order
        qo = quad_order if quad_order > 0 else (1 if mesh.order == 1 else 2)
        self.pts, self.w = po.quadrature(qo)
        self.n_local = mesh.elem_nodes.shape[1]
        self.n_dofs = 3 * mesh.n_nodes
        self.n_elements = mesh.n_elements

    def _geo(self, e):
        V = self.mesh.vertices[self.mesh.elem_vertices[e]]
        T = V[1:] - V[0]
        return np.linalg.inv(T).T, np.linalg.det(T)

    def _qp(self, e, x):
        """Per point: (G (n x 3), F, da)."""
        jit, det = self._geo(e)
        U = np.stack([x[3 * k:3 * k + 3] for k in self.mesh.elem_nodes[e]])
        for q in range(len(self.w)):
            G = po.basis_grad(self.order, self.pts[q]) @ jit
            yield G, np.eye(3) + U.T @ G, det * self.w[q]

    def element_energy(self, e, x):
        return sum(self._psi(F, e) * da for _, F, da in self._qp(e, x))

    def energy_per_element(self, x):
        return np.array([self.element_energy(e, x) for e in range(self.n_elements)])

    def energy(self, x):
        return float(np.sum(self.energy_per_element(x)))

    def element_gradient(self, e, x):
        g = np.zeros((self.n_local, 3))
        for G, F, da in self._qp(e, x):
            g += G @ self._P(F, e).T * da
        return g.reshape(-1)

    def gradient(self, x):
        out = np.zeros(self.n_dofs)
        for e in range(self.n_elements):
            g = self.element_gradient(e, x).reshape(-1, 3)
            for a, k in enumerate(self.mesh.elem_nodes[e]):
                out[3 * k:3 * k + 3] += g[a]
        return out

    def element_hessian(self, e, x):
        n = self.n_local
        Hl = np.zeros((3 * n, 3 * n))
        for G, F, da in self._qp(e, x):
            A = self._A(F, e)
            D = np.zeros((9, 3 * n))
            for a in range(n):
                for j in range(3):
                    for k in range(3):
                        D[3 * k + j, 3 * a + j] = G[a, k]
            Hl += D.T @ A @ D * da
        return Hl

    def _assemble(self, local):
        n = self.n_local
        rows, cols, vals = [], [], []
        for e in range(self.n_elements):
            H = local(e)
            g = (3 * self.mesh.elem_nodes[e][:, None] + np.arange(3)).reshape(-1)
            rows.append(np.repeat(g, 3 * n))
            cols.append(np.tile(g, 3 * n))
            vals.append(H.reshape(-1))
        return po.triplets_to_csc(self.n_dofs, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))

    def hessian(self, x):
        return self._assemble(lambda e: self.element_hessian(e, x))

    def hessian_scipy(self, x):
        import scipy.sparse as sp
        outer, inner, values = self.hessian(x)
        return sp.csc_matrix((values, inner, outer), shape=(self.n_dofs, self.n_dofs))

    def _grad_field(self, e, G, f):
        U = np.stack([f[3 * k:3 * k + 3] for k in self.mesh.elem_nodes[e]])
        return U.T @ G

    def material_vjp(self, x, p):
        """term[e] = -sum da <dP/dlambda, grad p>, term[E+e] with dP/dmu
        (FixedCorotational.cpp:493-546)."""
        E = self.n_elements
        out = np.zeros(2 * E)
        for e in range(E):
            for G, F, da in self._qp(e, x):
                gp = self._grad_field(e, G, p)
                U, s, V = autoflip_svd(F)
                dmu = 2 * (F - U @ V.T)
                dla = (np.prod(s) - 1) * cofactor(F)
                out[e] -= np.sum(dla * gp) * da
                out[E + e] -= np.sum(dmu * gp) * da
        return out

    def shape_vjp(self, x, p):
        """force_shape_derivative, elastic branch (ElasticForceDerivative.cpp:208-259),
        with compute_stress_grad_multiply_mat (FixedCorotational.cpp:438-464)."""
        nv = self.mesh.vertices.shape[0]
        term = np.zeros(3 * nv)
        for e in range(self.n_elements):
            jit, _ = self._geo(e)
            for q, (G, F, da) in enumerate(self._qp(e, x)):
                gu = F - np.eye(3)
                gp = self._grad_field(e, G, p)
                A = self._A(F, e)
                P = self._P(F, e)
                gv_all = po.basis_grad(1, self.pts[q]) @ jit
                for v in range(4):
                    gv = gv_all[v]
                    vid = self.mesh.elem_vertices[e][v]
                    for d in range(3):
                        GV = np.zeros((3, 3))
                        GV[d] = gv
                        mat = gu[:, d][:, None] * GV[d][None, :]
                        fp = (A @ mat.reshape(-1, order="F")).reshape(3, 3, order="F")
                        tmp = P @ GV.T - np.trace(GV) * P
                        term[3 * vid + d] -= np.sum((fp + tmp) * gp) * da
        return term
