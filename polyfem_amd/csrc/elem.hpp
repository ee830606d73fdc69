// Per-element arithmetic of the elasticity path, shared by every kernel.
//
// Conventions (reference: assembler/NeoHookeanElasticity.cpp, LinearElasticity.cpp):
//   U[a][d]   displacement of local node a, component d
//   jit       J^{-T} of the P1 geometry, row-major: jit[k*3+c]
//             (ElementAssemblyValues.cpp:96-102)
//   G_a(c)    = sum_k RG(q,a,k) jit[k][c], the physical basis gradient
//   H = grad u, H[d][c] = sum_a U[a][d] G_a(c);  F = I + H
//   da        = det J * w_q (Assembler.cpp:519), w already divided by 6
//
// All loops over local nodes, quadrature points and 3x3 indices are fully
// unrolled; the reference-element gradients RG(q,a,k) are then compile-time
// constants and the zero ones vanish from the code.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace pfgpu {

enum {
    MAT_NH = PFGPU_NEOHOOKEAN,
    MAT_LE = PFGPU_LINEAR_ELASTICITY,
    MAT_SV = PFGPU_SAINT_VENANT,
    MAT_FC = PFGPU_FIXED_COROTATIONAL,
    MAT_MR = PFGPU_MOONEY_RIVLIN
};

template <int ORDER> struct NLocal { static constexpr int value = ORDER == 1 ? 4 : 10; };

// Hammer-Marlowe-Stroud points a = (5 - sqrt 5)/20, b = (5 + 3 sqrt 5)/20,
// in the reference's order (a,a,b), (a,b,a), (b,a,a), (a,a,a)
// (autogen/auto_tetrahedron.ipp:6-12).
constexpr double HMS_A = 0.13819660112501053195188660538406111299991607666015625;
constexpr double HMS_B = 0.58541019662496840414434018384781666100025177001953125;

// Quadrature orders 3 and 4: the Xiao-Gimbutas rules of 6 and 11 points the
// reference tabulates (autogen/auto_tetrahedron.ipp:13-24), rows (x, y, z,
// w) in its order, weights summing to 1.
constexpr double XG6[6][4] = {
    {0.16200149169852446, 0.18385035049209775, 0.01271836631368145, 0.12232200275734507},
    {0.010905212211189244, 0.2815238021235462, 0.3621268299455338, 0.12806641271074692},
    {0.19011700243928392, 0.011403329444557169, 0.35862072046688387, 0.13256802714444524},
    {0.170816925164989, 0.15281814309092734, 0.6384932999617267, 0.14062440966040324},
    {0.15868516322744058, 0.5856628056552158, 0.13084716895209647, 0.22441516691755742},
    {0.5712260521491151, 0.1469183900871696, 0.14037280579421071, 0.2520039808095023},
};
constexpr double XG11[11][4] = {
    {0.04049050672759043, 0.013560701879802881, 0.7712547326953761, 0.03925109092483996},
    {0.752508507009655, 0.06809937093820666, 0.09798720364927911, 0.0552733691559369},
    {0.0672232948933834, 0.035183929773598716, 0.15636389323939529, 0.05539379887157637},
    {0.419266313879513, 0.04778143555908666, 0.47961101102565507, 0.05993318514655953},
    {0.45076587609127683, 0.05945661629943383, 0.05682401712793367, 0.06946996593763537},
    {0.1294113737889104, 0.3301904148374645, 0.0023910074574393647, 0.07616271524555836},
    {0.1215419913339278, 0.3064939884296903, 0.5629727601430461, 0.07942668006802531},
    {0.09720464458758327, 0.68439041545304, 0.1118007673973831, 0.1064680341554901},
    {0.029569495206479612, 0.3179035602133946, 0.32329398483747895, 0.11023423242849766},
    {0.43271023904776856, 0.3538232392092971, 0.10962240533194123, 0.15497611601624608},
    {0.24027666492807262, 0.12680172591539202, 0.32847320672203845, 0.19341081204963445},
};

// points of the rule with nq points (1, 4, 6, 11)
__host__ __device__ constexpr double quad_point(int nq, int q, int k)
{
    return nq == 1    ? 0.25
           : nq == 4  ? (q == 3 ? HMS_A : (k == 2 - q ? HMS_B : HMS_A))
           : nq == 6  ? XG6[q][k]
                      : XG11[q][k];
}

// weights sum to 1 and are divided by 6 (TetQuadrature.cpp:53)
__host__ __device__ constexpr double quad_weight(int nq, int q)
{
    return nq == 1 ? 1.0 / 6.0 : nq == 4 ? 0.25 / 6.0 : nq == 6 ? XG6[q][3] / 6.0 : XG11[q][3] / 6.0;
}
// the rules of 1 and 4 points weigh every point alike: the element kernels of
// the factored and fused paths (and the batch) fold the weight into
// per-element factors and take these rules only
__host__ __device__ constexpr bool quad_uniform(int nq) { return nq == 1 || nq == 4; }
// points of the rule of the given order (orders 1 to 4)
constexpr int quad_points_of_order(int order) { return order <= 1 ? 1 : order == 2 ? 4 : order == 3 ? 6 : 11; }

// d phi_a / d xi_k of the P1/P2 Lagrange tet (autogen/auto_p_bases.cpp:1100-1123, 1260-1330)
template <int ORDER> __host__ __device__ constexpr double ref_grad(double x, double y, double z, int a, int k)
{
    if (ORDER == 1) return a == 0 ? -1.0 : (k == a - 1 ? 1.0 : 0.0);
    switch (a) {
    case 0: return 4 * x + 4 * y + 4 * z - 3;
    case 1: return k == 0 ? 4 * x - 1 : 0.0;
    case 2: return k == 1 ? 4 * y - 1 : 0.0;
    case 3: return k == 2 ? 4 * z - 1 : 0.0;
    case 4: return k == 0 ? 4 * (-2 * x - y - z + 1) : -4 * x;
    case 5: return k == 0 ? 4 * y : (k == 1 ? 4 * x : 0.0);
    case 6: return k == 1 ? 4 * (-x - 2 * y - z + 1) : -4 * y;
    case 7: return k == 2 ? 4 * (-x - y - 2 * z + 1) : -4 * z;
    case 8: return k == 0 ? 4 * z : (k == 2 ? 4 * x : 0.0);
    default: return k == 1 ? 4 * z : (k == 2 ? 4 * y : 0.0);
    }
}

template <int ORDER, int NQ> __host__ __device__ constexpr double RG(int q, int a, int k)
{
    return ref_grad<ORDER>(quad_point(NQ, q, 0), quad_point(NQ, q, 1), quad_point(NQ, q, 2), a, k);
}

// sum_q RG(q,a,k) RG(q,b,l): the quadrature of the basis-gradient products
template <int ORDER, int NQ> __host__ __device__ constexpr double RGG(int a, int b, int k, int l)
{
    double s = 0;
    for (int q = 0; q < NQ; ++q) s += RG<ORDER, NQ>(q, a, k) * RG<ORDER, NQ>(q, b, l);
    return s;
}

// The same reference quantities as tables, for loops whose node indices are
// runtime values uniform over a wavefront (then the loads are scalar):
//   rg[(q NB + a) 3 + k] = RG(q, a, k)
//   rgg[(a NB + b) 6 + i] = coefficient of metric entry i (00 11 22 01 02 12)
//                           in sum_q G_a(q).G_b(q)
template <typename T, int ORDER, int NQ> struct RefTab {
    static constexpr int NB = NLocal<ORDER>::value;
    T rg[NQ * NB * 3];
    T rgg[NB * NB * 6];
    constexpr RefTab() : rg(), rgg()
    {
        for (int q = 0; q < NQ; ++q)
            for (int a = 0; a < NB; ++a)
                for (int k = 0; k < 3; ++k) rg[(q * NB + a) * 3 + k] = T(RG<ORDER, NQ>(q, a, k));
        const int kk[6] = {0, 1, 2, 0, 0, 1}, ll[6] = {0, 1, 2, 1, 2, 2};
        for (int a = 0; a < NB; ++a)
            for (int b = 0; b < NB; ++b)
                for (int i = 0; i < 6; ++i)
                    rgg[(a * NB + b) * 6 + i] =
                        T(i < 3 ? RGG<ORDER, NQ>(a, b, kk[i], ll[i])
                                : RGG<ORDER, NQ>(a, b, kk[i], ll[i]) + RGG<ORDER, NQ>(a, b, ll[i], kk[i]));
    }
};

// ------------------------------------------------------------ 3x3 helpers

template <typename T> __device__ __forceinline__ T det3(const T *m)
{
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
}

__device__ __forceinline__ float pf_log1p(float v) { return log1pf(v); }
__device__ __forceinline__ double pf_log1p(double v) { return log1p(v); }
__device__ __forceinline__ float pf_expm1(float v) { return expm1f(v); }
__device__ __forceinline__ double pf_expm1(double v) { return expm1(v); }

// delta - log1p(delta) without cancellation: series below a threshold
__device__ __forceinline__ float xmlog1p(float d)
{
    if (fabsf(d) < 0.1f) {
        float s = 1.0f / 12.0f;
#pragma unroll
        for (int n = 11; n >= 2; --n) s = 1.0f / float(n) - d * s;
        return d * d * s;
    }
    return d - log1pf(d);
}
__device__ __forceinline__ double xmlog1p(double d)
{
    if (fabs(d) < 1e-3) {
        double s = 1.0 / 10.0;
#pragma unroll
        for (int n = 9; n >= 2; --n) s = 1.0 / double(n) - d * s;
        return d * d * s;
    }
    return d - log1p(d);
}

// Kinematics of F = I + H, written so that nothing cancels when H is small:
// J - 1 = tr H + (tr^2 H - tr H^2)/2 + det H, and
// I - F^{-T} = (det H I + (1 + tr H) H^T - (H^T)^2) / J.
template <typename T> struct Kin {
    T H[9];
    T J, lnJ, delta, s2, dH;
    T FmT[9];  // F^{-T}
    T D[9];    // I - F^{-T}
};

template <typename T> __device__ __forceinline__ void kinematics(Kin<T> &k)
{
    const T *H = k.H;
    const T t = H[0] + H[4] + H[8];
    k.s2 = H[0] * H[4] + H[4] * H[8] + H[0] * H[8] - H[1] * H[3] - H[5] * H[7] - H[2] * H[6];
    k.dH = det3(H);
    k.delta = t + k.s2 + k.dH;
    k.J = T(1) + k.delta;
    k.lnJ = pf_log1p(k.delta);
    const T iJ = T(1) / k.J;
    // cofactor of F: columns f1 x f2, f2 x f0, f0 x f1 (NeoHookeanElasticity.cpp:513-527)
    const T F0 = T(1) + H[0], F4 = T(1) + H[4], F8 = T(1) + H[8];
    const T F[9] = {F0, H[1], H[2], H[3], F4, H[5], H[6], H[7], F8};
    k.FmT[0] = (F[4] * F[8] - F[5] * F[7]) * iJ;
    k.FmT[1] = (F[5] * F[6] - F[3] * F[8]) * iJ;
    k.FmT[2] = (F[3] * F[7] - F[4] * F[6]) * iJ;
    k.FmT[3] = (F[2] * F[7] - F[1] * F[8]) * iJ;
    k.FmT[4] = (F[0] * F[8] - F[2] * F[6]) * iJ;
    k.FmT[5] = (F[1] * F[6] - F[0] * F[7]) * iJ;
    k.FmT[6] = (F[1] * F[5] - F[2] * F[4]) * iJ;
    k.FmT[7] = (F[2] * F[3] - F[0] * F[5]) * iJ;
    k.FmT[8] = (F[0] * F[4] - F[1] * F[3]) * iJ;
    const T tp1 = T(1) + t;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            // (H^T)^2 [r][c] = sum_k H[k][r] H[c][k]
            const T h2 = H[0 * 3 + r] * H[c * 3 + 0] + H[1 * 3 + r] * H[c * 3 + 1] + H[2 * 3 + r] * H[c * 3 + 2];
            k.D[r * 3 + c] = ((r == c ? k.dH : T(0)) + tp1 * H[c * 3 + r] - h2) * iJ;
        }
}

// ------------------------------------------------------------- element data

template <typename T, int ORDER, int NQ> struct Elem {
    static constexpr int NB = NLocal<ORDER>::value;
    T U[NB][3];
    T jit[9];
    // material parameters: (lambda, mu), or (c1, c2, k) for MooneyRivlin
    T det, lam, mu, p3;
};

template <typename T, int ORDER, int NQ>
__device__ __forceinline__ void load_geometry(const DevMesh<T> &m, int e, Elem<T, ORDER, NQ> &el)
{
    const int E = m.n_elements;
#pragma unroll
    for (int k = 0; k < 9; ++k) el.jit[k] = m.geom[k * E + e];
    el.det = m.geom[9 * E + e];
    el.lam = m.lame[e];
    el.mu = m.lame[E + e];
    el.p3 = m.n_params > 2 ? m.lame[2 * int64_t(E) + e] : T(0);
}

// gather of the element's nodal values from an interleaved dof vector,
// relative to local node 0: U[a] = x_a - x_0, formed in FP64 before the
// conversion to T.  The field gradient is unchanged (the reference
// gradients of the basis sum to zero), but in FP32 the rounding of the
// stored values is relative to the small differences across the element
// instead of to the displacements themselves, which the gradient sum would
// otherwise amplify by |u| / (h |grad u|).
template <typename T, int NB>
__device__ __forceinline__ void gather(const DevMesh<T> &m, int e, const double *__restrict__ x, T (&U)[NB][3])
{
    const int E = m.n_elements;
    const int64_t n0 = 3 * int64_t(m.conn[e]);
    const double x0 = x[n0], x1 = x[n0 + 1], x2 = x[n0 + 2];
    U[0][0] = U[0][1] = U[0][2] = T(0);
#pragma unroll
    for (int a = 1; a < NB; ++a) {
        const int64_t n3 = 3 * int64_t(m.conn[a * E + e]);
        U[a][0] = T(x[n3] - x0);
        U[a][1] = T(x[n3 + 1] - x1);
        U[a][2] = T(x[n3 + 2] - x2);
    }
}

// grad of a nodal field at quadrature point q: out[d][c] = sum_a V[a][d] G_a(c)
template <typename T, int ORDER, int NQ, int NB>
__device__ __forceinline__ void field_grad(int q, const T (&V)[NB][3], const T *jit, T *out)
{
    T A[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) A[i] = T(0);
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double r = RG<ORDER, NQ>(q, a, k);
            if (r != 0.0) {
#pragma unroll
                for (int d = 0; d < 3; ++d) A[d * 3 + k] += V[a][d] * T(r);
            }
        }
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int c = 0; c < 3; ++c) out[d * 3 + c] = A[d * 3 + 0] * jit[0 * 3 + c] + A[d * 3 + 1] * jit[1 * 3 + c] + A[d * 3 + 2] * jit[2 * 3 + c];
}

// acc[a][d] += sum_c G_a(c) S[d][c] * da, as sum_k RG(q,a,k) (da S jit^T)[d][k]
template <typename T, int ORDER, int NQ, int NB>
__device__ __forceinline__ void contract_add(int q, const T *S, const T *jit, T da, T (&acc)[NB][3])
{
    T Q[9];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int k = 0; k < 3; ++k) Q[d * 3 + k] = da * (S[d * 3 + 0] * jit[k * 3 + 0] + S[d * 3 + 1] * jit[k * 3 + 1] + S[d * 3 + 2] * jit[k * 3 + 2]);
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double r = RG<ORDER, NQ>(q, a, k);
            if (r != 0.0) {
#pragma unroll
                for (int d = 0; d < 3; ++d) acc[a][d] += T(r) * Q[d * 3 + k];
            }
        }
}

// ---------------------------------------------------------- constitutive

// SaintVenant kinematics from H = grad u: F = I + H, the Green strain
// E = (H + H^T + H^T H) / 2 (no cancellation for small H) and the second
// Piola-Kirchhoff stress S = lambda tr(E) I + 2 mu E
template <typename T> struct SvKin {
    T F[9], E[9], trE;
};
template <typename T> __device__ __forceinline__ void sv_kinematics(const T *H, SvKin<T> &s)
{
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const T hh = H[0 * 3 + r] * H[0 * 3 + c] + H[1 * 3 + r] * H[1 * 3 + c] + H[2 * 3 + r] * H[2 * 3 + c];
            s.E[r * 3 + c] = T(0.5) * (H[r * 3 + c] + H[c * 3 + r] + hh);
            s.F[r * 3 + c] = H[r * 3 + c] + (r == c ? T(1) : T(0));
        }
    s.trE = s.E[0] + s.E[4] + s.E[8];
}
template <typename T> __device__ __forceinline__ void sv_stress(const SvKin<T> &s, T lam, T mu, T *S)
{
#pragma unroll
    for (int i = 0; i < 9; ++i) S[i] = T(2) * mu * s.E[i] + (i % 4 == 0 ? lam * s.trE : T(0));
}
// FixedCorotational kinematics (det F > 0): the rotation R = I + Q of the polar
// decomposition F = R S, by the scaled Newton iteration
// X <- (g X + X^{-T} / g) / 2 from X = F (Higham's Frobenius-norm scaling
// g = (|X^{-1}|_F / |X|_F)^{1/2}, which converges in a few steps for any
// condition number; g = 1, the plain iteration, once g is within 1e-3 of
// 1), written for Q = X - I:
//   Q <- (g - 1)^2 / (2 g) I + (g Q - X^{-T} Q^T / g) / 2,
// so that F - R = H - Q keeps its precision for small strains; and K^{-1}
// for K = tr(S) I - S, which gives the rotation's derivative
// dR[W] = R [K^{-1} a(W)]x with a(W) the axial vector of R^T W - W^T R.  A
// rotation that has not converged (|R^T R - I|_F above a few rounding
// units) poisons the outputs with NaN.
constexpr int FC_POLAR_ITERATIONS = 12;
template <typename T> struct FcKin {
    T Q[9], Ki6[6]; // Ki6: K^{-1}, entries 00 11 22 01 02 12
};
template <typename T> __device__ __forceinline__ void fc_kinematics(const T *H, FcKin<T> &f)
{
#pragma unroll
    for (int i = 0; i < 9; ++i) f.Q[i] = H[i];
#pragma unroll 1
    for (int it = 0; it < FC_POLAR_ITERATIONS; ++it) {
        const T X[9] = {T(1) + f.Q[0], f.Q[1], f.Q[2], f.Q[3], T(1) + f.Q[4], f.Q[5], f.Q[6], f.Q[7], T(1) + f.Q[8]};
        const T C[9] = {X[4] * X[8] - X[5] * X[7], X[5] * X[6] - X[3] * X[8], X[3] * X[7] - X[4] * X[6],
                        X[2] * X[7] - X[1] * X[8], X[0] * X[8] - X[2] * X[6], X[1] * X[6] - X[0] * X[7],
                        X[1] * X[5] - X[2] * X[4], X[2] * X[3] - X[0] * X[5], X[0] * X[4] - X[1] * X[3]};
        const T id = T(1) / (X[0] * C[0] + X[1] * C[1] + X[2] * C[2]);
        T nx = T(0), nc = T(0);
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            nx += X[i] * X[i];
            nc += C[i] * C[i];
        }
        // |X^{-1}|_F / |X|_F = |C|_F |1/det| / |X|_F
        T g = sqrt(sqrt(nc / nx) * fabs(id));
        if (!(fabs(g - T(1)) >= T(1e-3))) g = T(1);
        const T shift = (g - T(1)) * (g - T(1)) / (T(2) * g), ig = id / g;
        T N[9];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                N[r * 3 + c] = (r == c ? shift : T(0)) +
                               T(0.5) * (g * f.Q[r * 3 + c] - ig * (C[r * 3 + 0] * f.Q[c * 3 + 0] + C[r * 3 + 1] * f.Q[c * 3 + 1] +
                                                                    C[r * 3 + 2] * f.Q[c * 3 + 2]));
#pragma unroll
        for (int i = 0; i < 9; ++i) f.Q[i] = N[i];
    }
    // orthogonality of R = I + Q: R^T R - I = Q + Q^T + Q^T Q
    T orth = T(0);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const T d = f.Q[r * 3 + c] + f.Q[c * 3 + r] + f.Q[0 * 3 + r] * f.Q[0 * 3 + c] + f.Q[1 * 3 + r] * f.Q[1 * 3 + c] +
                        f.Q[2 * 3 + r] * f.Q[2 * 3 + c];
            orth += d * d;
        }
    const bool converged = orth <= (sizeof(T) == 4 ? T(1e-8) : T(1e-20));
    // S - I = R^T F - I = H + Q^T + Q^T H (symmetrised); K = tr(S) I - S
    T Sm[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            Sm[r * 3 + c] = H[r * 3 + c] + f.Q[c * 3 + r] + f.Q[0 * 3 + r] * H[0 * 3 + c] + f.Q[1 * 3 + r] * H[1 * 3 + c] +
                            f.Q[2 * 3 + r] * H[2 * 3 + c];
    const T tr = T(3) + Sm[0] + Sm[4] + Sm[8];
    const T k00 = tr - T(1) - Sm[0], k11 = tr - T(1) - Sm[4], k22 = tr - T(1) - Sm[8];
    const T k01 = -T(0.5) * (Sm[1] + Sm[3]), k02 = -T(0.5) * (Sm[2] + Sm[6]), k12 = -T(0.5) * (Sm[5] + Sm[7]);
    const T c00 = k11 * k22 - k12 * k12, c01 = k02 * k12 - k01 * k22, c02 = k01 * k12 - k02 * k11;
    const T c11 = k00 * k22 - k02 * k02, c12 = k01 * k02 - k00 * k12, c22 = k00 * k11 - k01 * k01;
    const T idet = T(1) / (k00 * c00 + k01 * c01 + k02 * c02);
    f.Ki6[0] = c00 * idet, f.Ki6[1] = c11 * idet, f.Ki6[2] = c22 * idet;
    f.Ki6[3] = c01 * idet, f.Ki6[4] = c02 * idet, f.Ki6[5] = c12 * idet;
    // an inverted or degenerate F (det F <= 0) poisons every output: the
    // orthogonal polar factor is then a reflection, not the reference's
    // sign-flipped rotation, so the stress, tangent and adjoint terms report
    // non-finite values (the energy is NaN there as well)
    const T detF = (T(1) + H[0]) * ((T(1) + H[4]) * (T(1) + H[8]) - H[5] * H[7]) -
                   H[1] * (H[3] * (T(1) + H[8]) - H[5] * H[6]) + H[2] * (H[3] * H[7] - (T(1) + H[4]) * H[6]);
    const T poison = detF > T(0) && converged ? T(0) : T(NAN);
#pragma unroll
    for (int i = 0; i < 9; ++i) f.Q[i] += poison;
#pragma unroll
    for (int i = 0; i < 6; ++i) f.Ki6[i] += poison;
}

// out = A B (3x3)
template <typename T> __device__ __forceinline__ void mul3(const T *A, const T *B, T *out)
{
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) out[r * 3 + c] = A[r * 3 + 0] * B[0 * 3 + c] + A[r * 3 + 1] * B[1 * 3 + c] + A[r * 3 + 2] * B[2 * 3 + c];
}

// MooneyRivlin (MooneyRivlinElasticity.hpp:25-44, the GenericElastic energy
// c1 (I1~ - 3) + c2 (I2~ - 3) + k/2 ln^2 J of F~ = J^{-1/3} F) in forms that
// do not cancel for small H (kinematics() first): with B = F F^T = I + S,
// S = H + H^T + H H^T,
//   I1 - 3 = tr S,  I2 - 3 = 2 tr S + (tr^2 S - tr S^2) / 2,
//   J^{-2/3} - 1 = expm1(-2/3 ln J),  J^{-4/3} - 1 = expm1(-4/3 ln J),
// and the stress P = c1 J^{-2/3} R1 + c2 J^{-4/3} Q2 + k ln J F^{-T} with
//   R1 = 2 F - 2/3 I1 F^{-T}           = 2 (H + D) - 2/3 (I1 - 3) F^{-T},
//   Q2 = 2 I1 F - 2 B F - 4/3 I2 F^{-T} = 4 (H + D) + 2 ((I1 - 3) I - S) F - 4/3 (I2 - 3) F^{-T},
// D = I - F^{-T}.
template <typename T> struct MrKin {
    T S[9];
    T i1, i2, ea, eb;
};
template <typename T> __device__ __forceinline__ void mr_kinematics(const Kin<T> &k, MrKin<T> &r)
{
    const T *H = k.H;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            r.S[a * 3 + c] = H[a * 3 + c] + H[c * 3 + a] + H[a * 3 + 0] * H[c * 3 + 0] + H[a * 3 + 1] * H[c * 3 + 1] +
                             H[a * 3 + 2] * H[c * 3 + 2];
    r.i1 = r.S[0] + r.S[4] + r.S[8];
    T ss = T(0);
#pragma unroll
    for (int i = 0; i < 9; ++i) ss += r.S[i] * r.S[i];
    r.i2 = T(2) * r.i1 + T(0.5) * (r.i1 * r.i1 - ss);
    r.ea = pf_expm1(T(-2.0 / 3.0) * k.lnJ);
    r.eb = pf_expm1(T(-4.0 / 3.0) * k.lnJ);
}
template <typename T> __device__ __forceinline__ void mr_brackets(const Kin<T> &k, const MrKin<T> &r, T *R1, T *Q2)
{
    const T *H = k.H, *A = k.FmT;
    // ((I1 - 3) I - S) F
    T M[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) M[i] = (i % 4 == 0 ? r.i1 : T(0)) - r.S[i];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const T mf = M[a * 3 + c] + M[a * 3 + 0] * H[0 * 3 + c] + M[a * 3 + 1] * H[1 * 3 + c] + M[a * 3 + 2] * H[2 * 3 + c];
            const T hd = H[a * 3 + c] + k.D[a * 3 + c];
            R1[a * 3 + c] = T(2) * hd - T(2.0 / 3.0) * r.i1 * A[a * 3 + c];
            Q2[a * 3 + c] = T(4) * hd + T(2) * mf - T(4.0 / 3.0) * r.i2 * A[a * 3 + c];
        }
}
// dP[W] = c1 J^{-2/3} (dR1 - 2/3 (A:W) R1) + c2 J^{-4/3} (dQ2 - 4/3 (A:W) Q2)
//         + k ((A:W) A - ln J A W^T A),  A = F^{-T}, with
//   dR1 = 2 W - 4/3 (F:W) A + 2/3 I1 A W^T A,
//   dQ2 = 4 (F:W) F + 2 I1 W - 2 W F^T F - 2 F W^T F - 2 B W
//         - 4/3 (2 I1 (F:W) - 2 (B F):W) A + 4/3 I2 A W^T A
template <typename T>
__device__ __forceinline__ void mr_dstress(const Kin<T> &k, T c1, T c2, T kk, const T *W, T *out)
{
    MrKin<T> r;
    mr_kinematics(k, r);
    T R1[9], Q2[9];
    mr_brackets(k, r, R1, Q2);
    const T *H = k.H, *A = k.FmT;
    const T F[9] = {T(1) + H[0], H[1], H[2], H[3], T(1) + H[4], H[5], H[6], H[7], T(1) + H[8]};
    const T Bm[9] = {T(1) + r.S[0], r.S[1], r.S[2], r.S[3], T(1) + r.S[4], r.S[5], r.S[6], r.S[7], T(1) + r.S[8]};
    const T I1 = T(3) + r.i1, I2 = T(3) + r.i2;
    T BF[9], C[9], AWt[9], AWA[9], WC[9], FWt[9], FWF[9], BW[9];
    mul3(Bm, F, BF);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            C[a * 3 + c] = F[0 * 3 + a] * F[0 * 3 + c] + F[1 * 3 + a] * F[1 * 3 + c] + F[2 * 3 + a] * F[2 * 3 + c];
            AWt[a * 3 + c] = A[a * 3 + 0] * W[c * 3 + 0] + A[a * 3 + 1] * W[c * 3 + 1] + A[a * 3 + 2] * W[c * 3 + 2];
            FWt[a * 3 + c] = F[a * 3 + 0] * W[c * 3 + 0] + F[a * 3 + 1] * W[c * 3 + 1] + F[a * 3 + 2] * W[c * 3 + 2];
        }
    mul3(AWt, A, AWA);
    mul3(W, C, WC);
    mul3(FWt, F, FWF);
    mul3(Bm, W, BW);
    T aw = T(0), fw = T(0), gw = T(0);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        aw += A[i] * W[i];
        fw += F[i] * W[i];
        gw += BF[i] * W[i];
    }
    const T a = c1 * (T(1) + r.ea), b = c2 * (T(1) + r.eb);
    const T g = T(4.0 / 3.0) * (T(2) * I1 * fw - T(2) * gw);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const T dR1 = T(2) * W[i] - T(4.0 / 3.0) * fw * A[i] + T(2.0 / 3.0) * I1 * AWA[i];
        const T dQ2 = T(4) * fw * F[i] + T(2) * I1 * W[i] - T(2) * WC[i] - T(2) * FWF[i] - T(2) * BW[i] - g * A[i] +
                      T(4.0 / 3.0) * I2 * AWA[i];
        out[i] = a * (dR1 - T(2.0 / 3.0) * aw * R1[i]) + b * (dQ2 - T(4.0 / 3.0) * aw * Q2[i]) +
                 kk * (aw * A[i] - k.lnJ * AWA[i]);
    }
}


// energy density (NeoHookeanElasticity.cpp:384; LinearElasticity.cpp:129)
template <typename T, int MAT> __device__ __forceinline__ T energy_density(const Kin<T> &k, T lam, T mu, T p3)
{
    const T *H = k.H;
    if (MAT == MAT_MR) {
        // c1 (J^{-2/3} I1 - 3) + c2 (J^{-4/3} I2 - 3) + k/2 ln^2 J with (c1, c2, k) = (lam, mu, p3)
        MrKin<T> r;
        mr_kinematics(k, r);
        return lam * (T(3) * r.ea + r.i1 + r.ea * r.i1) + mu * (T(3) * r.eb + r.i2 + r.eb * r.i2) +
               T(0.5) * p3 * k.lnJ * k.lnJ;
    } else if (MAT == MAT_NH) {
        // mu/2 (|F|^2 - 3 - 2 ln J) + lambda/2 ln^2 J with |F|^2 - 3 = 2 tr H + |H|^2
        // and tr H - ln J = -(s2 + det H) + (delta - ln(1 + delta))
        T hh = T(0);
#pragma unroll
        for (int i = 0; i < 9; ++i) hh += H[i] * H[i];
        return mu * (xmlog1p(k.delta) - k.s2 - k.dH) + T(0.5) * mu * hh + T(0.5) * lam * k.lnJ * k.lnJ;
    } else if (MAT == MAT_FC) {
        // mu |F - R|^2 + lambda/2 (J - 1)^2 (FixedCorotational.cpp:589-596: the
        // sum of (sigma_i - 1)^2 is |F - R|^2); NaN for det F <= 0
        FcKin<T> f;
        fc_kinematics(H, f);
        T dd = T(0);
#pragma unroll
        for (int i = 0; i < 9; ++i) dd += (H[i] - f.Q[i]) * (H[i] - f.Q[i]);
        const T v = mu * dd + T(0.5) * lam * k.delta * k.delta;
        return k.J > T(0) ? v : T(NAN);
    } else if (MAT == MAT_SV) {
        // (SaintVenantElasticity.cpp:235-278: 1/2 sigma : E with sigma = C : E, C isotropic)
        SvKin<T> s;
        sv_kinematics(H, s);
        T ee = T(0);
#pragma unroll
        for (int i = 0; i < 9; ++i) ee += s.E[i] * s.E[i];
        return mu * ee + T(0.5) * lam * s.trE * s.trE;
    } else {
        T ee = T(0);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T eps = T(0.5) * (H[r * 3 + c] + H[c * 3 + r]);
                ee += eps * eps;
            }
        const T tr = H[0] + H[4] + H[8];
        return mu * ee + T(0.5) * lam * tr * tr;
    }
}

// first Piola-Kirchhoff stress (NeoHookeanElasticity.cpp:532: mu F - mu/J cof F + lambda ln J / J cof F)
template <typename T, int MAT> __device__ __forceinline__ void pk1(const Kin<T> &k, T lam, T mu, T p3, T *P)
{
    const T *H = k.H;
    if (MAT == MAT_MR) {
        MrKin<T> r;
        mr_kinematics(k, r);
        T R1[9], Q2[9];
        mr_brackets(k, r, R1, Q2);
        const T a = lam * (T(1) + r.ea), b = mu * (T(1) + r.eb), l = p3 * k.lnJ;
#pragma unroll
        for (int i = 0; i < 9; ++i) P[i] = a * R1[i] + b * Q2[i] + l * k.FmT[i];
    } else if (MAT == MAT_NH) {
        const T ll = lam * k.lnJ;
#pragma unroll
        for (int i = 0; i < 9; ++i) P[i] = mu * (H[i] + k.D[i]) + ll * k.FmT[i];
    } else if (MAT == MAT_FC) {
        // P = 2 mu (F - R) + lambda (J - 1) cof F (FixedCorotational.cpp:683-706)
        FcKin<T> f;
        fc_kinematics(H, f);
        const T lj = lam * k.delta * k.J;
#pragma unroll
        for (int i = 0; i < 9; ++i) P[i] = T(2) * mu * (H[i] - f.Q[i]) + lj * k.FmT[i];
    } else if (MAT == MAT_SV) {
        // P = F S
        SvKin<T> s;
        sv_kinematics(H, s);
        T S[9];
        sv_stress(s, lam, mu, S);
        mul3(s.F, S, P);
    } else {
        const T tr = lam * (H[0] + H[4] + H[8]);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) P[r * 3 + c] = mu * (H[r * 3 + c] + H[c * 3 + r]) + (r == c ? tr : T(0));
    }
}

// linearised stress C : W (NeoHookeanElasticity.cpp:660-680 'result';
// LinearElasticity.cpp:253-270)
template <typename T, int MAT>
__device__ __forceinline__ void dstress(const Kin<T> &k, T lam, T mu, T p3, const T *W, T *out)
{
    if (MAT == MAT_MR) {
        mr_dstress(k, lam, mu, p3, W, out);
    } else if (MAT == MAT_NH) {
        const T *A = k.FmT;
        T fw = T(0);
#pragma unroll
        for (int i = 0; i < 9; ++i) fw += A[i] * W[i];
        T t1[9]; // A W^T
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) t1[r * 3 + c] = A[r * 3 + 0] * W[c * 3 + 0] + A[r * 3 + 1] * W[c * 3 + 1] + A[r * 3 + 2] * W[c * 3 + 2];
        const T cc = mu - lam * k.lnJ, lf = lam * fw;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T t2 = t1[r * 3 + 0] * A[0 * 3 + c] + t1[r * 3 + 1] * A[1 * 3 + c] + t1[r * 3 + 2] * A[2 * 3 + c];
                out[r * 3 + c] = mu * W[r * 3 + c] + cc * t2 + lf * A[r * 3 + c];
            }
    } else if (MAT == MAT_FC) {
        // dP = 2 mu (W - R [K^{-1} a(W)]x) + lambda J (2J - 1) (F^{-T}:W) F^{-T}
        //      - lambda J (J - 1) F^{-T} W^T F^{-T}
        FcKin<T> f;
        fc_kinematics(k.H, f);
        const T *A = k.FmT;
        T fw = T(0);
#pragma unroll
        for (int i = 0; i < 9; ++i) fw += A[i] * W[i];
        T t1[9];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) t1[r * 3 + c] = A[r * 3 + 0] * W[c * 3 + 0] + A[r * 3 + 1] * W[c * 3 + 1] + A[r * 3 + 2] * W[c * 3 + 2];
        const T R[9] = {T(1) + f.Q[0], f.Q[1], f.Q[2], f.Q[3], T(1) + f.Q[4], f.Q[5], f.Q[6], f.Q[7], T(1) + f.Q[8]};
        T B[9]; // R^T W
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) B[r * 3 + c] = R[0 * 3 + r] * W[0 * 3 + c] + R[1 * 3 + r] * W[1 * 3 + c] + R[2 * 3 + r] * W[2 * 3 + c];
        const T a0 = B[7] - B[5], a1 = B[2] - B[6], a2 = B[3] - B[1];
        const T *K = f.Ki6;
        const T w0 = K[0] * a0 + K[3] * a1 + K[4] * a2, w1 = K[3] * a0 + K[1] * a1 + K[5] * a2, w2 = K[4] * a0 + K[5] * a1 + K[2] * a2;
        const T Om[9] = {T(0), -w2, w1, w2, T(0), -w0, -w1, w0, T(0)};
        T RO[9];
        mul3(R, Om, RO);
        const T cl = lam * k.J * (T(1) + T(2) * k.delta), cq = lam * k.J * k.delta;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T t2 = t1[r * 3 + 0] * A[0 * 3 + c] + t1[r * 3 + 1] * A[1 * 3 + c] + t1[r * 3 + 2] * A[2 * 3 + c];
                out[r * 3 + c] = T(2) * mu * (W[r * 3 + c] - RO[r * 3 + c]) + cl * fw * A[r * 3 + c] - cq * t2;
            }
    } else if (MAT == MAT_SV) {
        // dP = W S + F dS, dS = lambda tr(dE) I + 2 mu dE, dE = (W^T F + F^T W) / 2
        SvKin<T> s;
        sv_kinematics(k.H, s);
        T S[9], FtW[9], dS[9], WS[9], FdS[9];
        sv_stress(s, lam, mu, S);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                FtW[r * 3 + c] = s.F[0 * 3 + r] * W[0 * 3 + c] + s.F[1 * 3 + r] * W[1 * 3 + c] + s.F[2 * 3 + r] * W[2 * 3 + c];
        const T trd = FtW[0] + FtW[4] + FtW[8];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) dS[r * 3 + c] = mu * (FtW[r * 3 + c] + FtW[c * 3 + r]) + (r == c ? lam * trd : T(0));
        mul3(W, S, WS);
        mul3(s.F, dS, FdS);
#pragma unroll
        for (int i = 0; i < 9; ++i) out[i] = WS[i] + FdS[i];
    } else {
        const T tr = lam * (W[0] + W[4] + W[8]);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) out[r * 3 + c] = mu * (W[r * 3 + c] + W[c * 3 + r]) + (r == c ? tr : T(0));
    }
}

// SaintVenant local Hessian (derived from dP = W S + F dS): with k_a = F G_a,
//   H_ab = sum_q da [ (G_a^T S G_b) I + mu (G_a . G_b) F F^T + lambda k_a k_b^T + mu k_b k_a^T ].
// The last two terms have the NeoHookean form (hess_qp); sv_extra gives the
// per-point data of the first two: the S-weighted reference metric
// da jit S jit^T and mu da F F^T (symmetric, entries 00 11 22 01 02 12).
template <typename T>
__device__ __forceinline__ void sv_extra(const T *H, const T *jit, T lam, T mu, T da, T *sg6, T *ff6)
{
    SvKin<T> s;
    sv_kinematics(H, s);
    T S[9], JS[9];
    sv_stress(s, lam, mu, S);
    // JS[k][j] = sum_i jit[k][i] S[i][j]; sg[k][l] = sum_j JS[k][j] jit[l][j]
#pragma unroll
    for (int kk = 0; kk < 3; ++kk)
#pragma unroll
        for (int j = 0; j < 3; ++j) JS[kk * 3 + j] = jit[kk * 3 + 0] * S[0 * 3 + j] + jit[kk * 3 + 1] * S[1 * 3 + j] + jit[kk * 3 + 2] * S[2 * 3 + j];
    auto sg = [&](int kk, int l) { return da * (JS[kk * 3 + 0] * jit[l * 3 + 0] + JS[kk * 3 + 1] * jit[l * 3 + 1] + JS[kk * 3 + 2] * jit[l * 3 + 2]); };
    auto ff = [&](int m, int n) { return mu * da * (s.F[m * 3 + 0] * s.F[n * 3 + 0] + s.F[m * 3 + 1] * s.F[n * 3 + 1] + s.F[m * 3 + 2] * s.F[n * 3 + 2]); };
    sg6[0] = sg(0, 0), sg6[1] = sg(1, 1), sg6[2] = sg(2, 2), sg6[3] = sg(0, 1), sg6[4] = sg(0, 2), sg6[5] = sg(1, 2);
    ff6[0] = ff(0, 0), ff6[1] = ff(1, 1), ff6[2] = ff(2, 2), ff6[3] = ff(0, 1), ff6[4] = ff(0, 2), ff6[5] = ff(1, 2);
}

// u^T A v for a symmetric A in the 6-entry form
template <typename T> __device__ __forceinline__ T sym_form(const T *u, const T *a6, const T *v)
{
    return u[0] * (a6[0] * v[0] + a6[3] * v[1] + a6[4] * v[2]) + u[1] * (a6[3] * v[0] + a6[1] * v[1] + a6[5] * v[2]) +
           u[2] * (a6[4] * v[0] + a6[5] * v[1] + a6[2] * v[2]);
}

// ------------------------------------------------------------- Hessian data
//
// With k_a = F^{-T} G_a the NeoHookean local Hessian block is
//   H_ab = sum_q da [ mu (G_a . G_b) I + lambda k_a k_b^T + (mu - lambda ln J) k_b k_a^T ],
// which equals D^T H9 D of NeoHookeanElasticity.cpp:636-652 (the d2J/dF2 term
// contracts to (k_a k_b^T - k_b k_a^T) J / J).  Linear elasticity is the same
// form with k_a = G_a, lambda and mu (LinearElasticity.cpp:30-63).
// M maps reference gradients to k: k_a[m] = sum_k M[m][k] RG(q,a,k).
template <typename T, int MAT>
__device__ __forceinline__ void hess_qp(const Kin<T> &k, const T *jit, T lam, T mu, T da, T *M, T &cl, T &cc)
{
    if (MAT == MAT_NH) {
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int kk = 0; kk < 3; ++kk)
                M[m * 3 + kk] = k.FmT[m * 3 + 0] * jit[kk * 3 + 0] + k.FmT[m * 3 + 1] * jit[kk * 3 + 1] + k.FmT[m * 3 + 2] * jit[kk * 3 + 2];
        cl = da * lam;
        cc = da * (mu - lam * k.lnJ);
    } else if (MAT == MAT_FC) {
        // k_a = F^{-T} G_a: lambda J (2J - 1) k_a k_b^T - lambda J (J - 1) k_b k_a^T; the
        // 2 mu metric term and the rotation term are added by the caller
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int kk = 0; kk < 3; ++kk)
                M[m * 3 + kk] = k.FmT[m * 3 + 0] * jit[kk * 3 + 0] + k.FmT[m * 3 + 1] * jit[kk * 3 + 1] + k.FmT[m * 3 + 2] * jit[kk * 3 + 2];
        cl = da * lam * k.J * (T(1) + T(2) * k.delta);
        cc = -da * lam * k.J * k.delta;
    } else if (MAT == MAT_SV) {
        // k_a = F G_a; the two further SaintVenant terms are added by the caller
        const T F[9] = {T(1) + k.H[0], k.H[1], k.H[2], k.H[3], T(1) + k.H[4], k.H[5], k.H[6], k.H[7], T(1) + k.H[8]};
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int kk = 0; kk < 3; ++kk)
                M[m * 3 + kk] = F[m * 3 + 0] * jit[kk * 3 + 0] + F[m * 3 + 1] * jit[kk * 3 + 1] + F[m * 3 + 2] * jit[kk * 3 + 2];
        cl = da * lam;
        cc = da * mu;
    } else {
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) M[m * 3 + kk] = jit[kk * 3 + m];
        cl = da * lam;
        cc = da * mu;
    }
}

template <typename T, int ORDER, int NQ>
__device__ __forceinline__ void khat(int q, int a, const T *M, T *ka)
{
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        T s = T(0);
#pragma unroll
        for (int kk = 0; kk < 3; ++kk) {
            const double r = RG<ORDER, NQ>(q, a, kk);
            if (r != 0.0) s += M[m * 3 + kk] * T(r);
        }
        ka[m] = s;
    }
}

// metric jit jit^T (symmetric, 6 entries: 00 11 22 01 02 12)
template <typename T> __device__ __forceinline__ void metric(const T *jit, T *g6)
{
    auto dot = [&](int k, int l) { return jit[k * 3 + 0] * jit[l * 3 + 0] + jit[k * 3 + 1] * jit[l * 3 + 1] + jit[k * 3 + 2] * jit[l * 3 + 2]; };
    g6[0] = dot(0, 0); g6[1] = dot(1, 1); g6[2] = dot(2, 2);
    g6[3] = dot(0, 1); g6[4] = dot(0, 2); g6[5] = dot(1, 2);
}

// sum_q G_a(q).G_b(q) = sum_{k,l} metric[k][l] RGG(a,b,k,l)
template <int ORDER, int NQ, typename T> __host__ __device__ __forceinline__ T grad_dot(int a, int b, const T *g6)
{
    T s = T(0);
    const int kk[6] = {0, 1, 2, 0, 0, 1}, ll[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const double c = i < 3 ? RGG<ORDER, NQ>(a, b, kk[i], ll[i])
                               : RGG<ORDER, NQ>(a, b, kk[i], ll[i]) + RGG<ORDER, NQ>(a, b, ll[i], kk[i]);
        if (c != 0.0) s += T(c) * g6[i];
    }
    return s;
}

} // namespace pfgpu
