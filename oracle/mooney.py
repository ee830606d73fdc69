"""FP64 restatement of the reference's MooneyRivlin material (test
infrastructure only: the checker for the HIP library, never the product).

Follows assembler/MooneyRivlinElasticity.hpp:25-44 (a GenericElastic
material, GenericElastic.hpp:93-348):
  * psi = c1 (I1~ - 3) + c2 (I2~ - 3) + k/2 ln^2 J with F~ = F / J^(1/3),
    C~ = F~ F~^T, I1~ = tr C~, I2~ = (tr^2 C~ - tr C~^2) / 2
    (utils/ElasticityUtils.hpp:139-149), written literally;
  * the first Piola-Kirchhoff stress in closed form,
      P = c1 J^{-2/3} (2 F - 2/3 I1 F^{-T})
        + c2 J^{-4/3} (2 I1 F - 2 B F - 4/3 I2 F^{-T}) + k ln J F^{-T},
    B = F F^T, I1 = tr B, I2 = (I1^2 - tr B^2) / 2 (the reference obtains it
    by automatic differentiation of psi; the tests check it against a
    complex-step derivative of psi);
  * the 9x9 stiffness d P / d F by complex-step differentiation of P (exact
    to rounding, no cancellation), in the layout of fcr.stiffness;
  * element loops, assembly and the shape derivative of fcr.FcrOracle
    (GenericElastic::compute_stress_grad_multiply_mat, GenericElastic.cpp:495-532).
The reference has no Lame-parameter (material) derivative for this material.
"""
from __future__ import annotations

import numpy as np

from .fcr import FcrOracle

H_CS = 1e-30


def energy_density(F, c1, c2, k):
    J = np.linalg.det(F)
    Ft = F / J ** (1.0 / 3.0)
    Ct = Ft @ Ft.T
    I1 = np.trace(Ct)
    I2 = 0.5 * (I1 * I1 - np.trace(Ct @ Ct))
    return c1 * (I1 - 3) + c2 * (I2 - 3) + k / 2 * np.log(J) ** 2


def stress(F, c1, c2, k):
    J = np.linalg.det(F)
    A = np.linalg.inv(F).T
    B = F @ F.T
    I1 = np.trace(B)
    I2 = 0.5 * (I1 * I1 - np.trace(B @ B))
    return (c1 * J ** (-2.0 / 3.0) * (2 * F - 2.0 / 3.0 * I1 * A)
            + c2 * J ** (-4.0 / 3.0) * (2 * I1 * F - 2 * B @ F - 4.0 / 3.0 * I2 * A) + k * np.log(J) * A)


def stiffness(F, c1, c2, k):
    """Entries [i + 3 j, r + 3 s] = d P_ij / d F_rs."""
    M = np.zeros((9, 9))
    for r in range(3):
        for s_ in range(3):
            Fc = F.astype(np.complex128)
            Fc[r, s_] += 1j * H_CS
            M[:, r + 3 * s_] = (stress(Fc, c1, c2, k).imag / H_CS).reshape(-1, order="F")
    return M


class MooneyOracle(FcrOracle):
    """Element loops of the MooneyRivlin form on one mesh (FP64)."""

    def __init__(self, mesh, c1, c2, k, quad_order=0):
        super().__init__(mesh, 0.0, 0.0, quad_order)
        E = mesh.n_elements
        self.c1 = np.broadcast_to(np.asarray(c1, dtype=np.float64), (E,))
        self.c2 = np.broadcast_to(np.asarray(c2, dtype=np.float64), (E,))
        self.k = np.broadcast_to(np.asarray(k, dtype=np.float64), (E,))

    def _psi(self, F, e):
        return energy_density(F, self.c1[e], self.c2[e], self.k[e])

    def _P(self, F, e):
        return stress(F, self.c1[e], self.c2[e], self.k[e])

    def _A(self, F, e):
        return stiffness(F, self.c1[e], self.c2[e], self.k[e])

    def material_vjp(self, x, p):
        raise NotImplementedError("the reference has no material derivative for MooneyRivlin")
