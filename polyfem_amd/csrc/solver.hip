// Newton-CG on the free dofs, device-resident.
//
// Follows the reference's nonlinear solve of the elastic form: Dirichlet dofs
// are removed from the unknowns (NLProblem::full_to_reduced /
// reduced_to_full, NLProblem.cpp:583-629; the gradient and Hessian rows and
// columns of fixed dofs are dropped, BCLagrangianForm.cpp:149-200), the
// Newton direction solves H_ff d = -g_f (polysolve Newton with a CG linear
// solver, here matrix-free through the Hessian-vector product) and the step
// length comes from a backtracking Armijo line search on the energy.
//
// Fixed dofs are handled with a mask instead of compacted vectors: the
// reduced system is the full system with fixed rows and columns zeroed,
// which keeps every vector in the mesh's dof layout.
//
// Precision: whatever the assembler's element precision, every solver
// vector, dot product, norm and energy difference is FP64, and every
// reduction has a fixed order (per-block partials over a fixed grid, then
// one fixed tree), so a solve is bitwise reproducible.
//
// The objective is w E(x) + 1/2 (x - xt)^T M (x - xt) - w f^T x: the elastic
// form alone (w = 1, no xt, no f), with a body load f, or the implicit-Euler
// step objective (w = dt^2, xt = x_prev + dt v_prev; InertiaForm.cpp:17-35,
// ImplicitEuler.cpp:13-31, the elastic form weighted by dt^2).
//
// One CG iteration is a fixed sequence of kernels whose scalars (alpha,
// beta, r.r) live in device memory, so it can be captured once in a hipGraph
// and replayed; the host reads the residual only every CHECK iterations.
#include <cmath>
#include <cstring>

#include "dispatch.hpp"

namespace pfgpu {

namespace {

constexpr int DOT_BLOCKS = 1024; // fixed grid of the dot-product kernels
constexpr int CHECK = 8;         // CG iterations between host convergence checks
#ifndef PF_SPMV_LANES
#define PF_SPMV_LANES 16
#endif
constexpr int SPMV_LANES = PF_SPMV_LANES; // lanes per block row of the block-sparse product

// scalar slots of the CG workspace (r.z: preconditioned residual product)
// S_FLAG: 0 while the CG runs, 1 stopped on non-positive curvature, 2 stopped
// at the tolerance (r.r <= S_TOL); S_ITER counts the updates taken
enum { S_RR, S_PQ, S_ALPHA, S_BETA, S_RZ, S_FLAG, S_B2, S_GD, S_E0, S_E1, S_M, S_F, S_TOL, S_ITER, S_COUNT };
enum { CG_RUNNING = 0, CG_CURVATURE = 1, CG_TOLERANCE = 2 };

__device__ double block_reduce(double v, double *sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (t < w) sh[t] += sh[t + w];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// per-block partial of sum_i a_i b_i over free dofs (mask: 1 = fixed)
__global__ __launch_bounds__(BLOCK) void k_dot_partials(const double *__restrict__ a, const double *__restrict__ b,
                                                        const uint8_t *__restrict__ fixed, int n, double *__restrict__ part)
{
    __shared__ double s[BLOCK];
    double v = 0;
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += DOT_BLOCKS * BLOCK)
        if (!fixed || !fixed[i]) v = fma(a[i], b[i], v);
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// one block: fixed-order sum of the DOT_BLOCKS partials into scal[slot], then
// the scalar bookkeeping of the step that follows
enum { AFTER_NONE, AFTER_ALPHA };
__device__ double sum_partials(const double *__restrict__ part, double *sh)
{
    double a = 0;
    for (int k = threadIdx.x; k < DOT_BLOCKS; k += BLOCK) a += part[k];
    return block_reduce(a, sh);
}

__global__ __launch_bounds__(BLOCK) void k_finish_dot(const double *__restrict__ part, double *__restrict__ scal, int slot,
                                                      int after)
{
    __shared__ double s[BLOCK];
    const double v = sum_partials(part, s);
    if (threadIdx.x != 0) return;
    scal[slot] = v;
    if (after == AFTER_ALPHA) {
        // alpha = r.z / p.Hp; a non-positive curvature stops the iteration
        const double rz = scal[S_RZ];
        const bool ok = v > 0.0 && rz > 0.0 && scal[S_FLAG] == 0.0;
        if (!(v > 0.0) && rz > 0.0 && scal[S_FLAG] == 0.0) scal[S_FLAG] = CG_CURVATURE;
        scal[S_ALPHA] = ok ? rz / v : 0.0;
    }
}

// z = P r for one node: P the inverse of the node's 3x3 diagonal block (symmetric,
// stored as its upper triangle P00 P01 P02 P11 P12 P22 in FP32), or the identity
constexpr int PINV = 6;
__device__ inline void precondition(const float *__restrict__ pinv, int64_t i, const double r[3], double z[3])
{
    if (!pinv) {
        z[0] = r[0], z[1] = r[1], z[2] = r[2];
        return;
    }
    const float *P = pinv + PINV * i;
    const double a = P[0], b = P[1], c = P[2], d = P[3], e = P[4], f = P[5];
    z[0] = fma(a, r[0], fma(b, r[1], c * r[2]));
    z[1] = fma(b, r[0], fma(d, r[1], e * r[2]));
    z[2] = fma(c, r[0], fma(e, r[1], f * r[2]));
}

// PCG start: y = 0, r = b on free dofs, z = P r, p = z; partials of r.z and r.r
__global__ __launch_bounds__(BLOCK) void k_pcg_start(int n_nodes, const double *__restrict__ b,
                                                     const uint8_t *__restrict__ fixed, const float *__restrict__ pinv,
                                                     double *__restrict__ y, double *__restrict__ r,
                                                     double *__restrict__ z, double *__restrict__ p,
                                                     double *__restrict__ part_rz, double *__restrict__ part_rr)
{
    __shared__ double sh[BLOCK];
    double rz = 0, rr = 0;
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < n_nodes; i += DOT_BLOCKS * BLOCK) {
        const int64_t o = 3 * int64_t(i);
        double ri[3], zi[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) ri[d] = fixed && fixed[o + d] ? 0.0 : b[o + d];
        precondition(pinv, i, ri, zi);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            y[o + d] = 0.0;
            r[o + d] = ri[d];
            z[o + d] = zi[d];
            p[o + d] = zi[d];
            rz = fma(ri[d], zi[d], rz);
            rr = fma(ri[d], ri[d], rr);
        }
    }
    const double a = block_reduce(rz, sh), c = block_reduce(rr, sh);
    if (threadIdx.x == 0) {
        part_rz[blockIdx.x] = a;
        part_rr[blockIdx.x] = c;
    }
}

// y += alpha p, r -= alpha q on free dofs; z = P r; partials of r.z and r.r
__global__ __launch_bounds__(BLOCK) void k_pcg_update(int n_nodes, double *__restrict__ y, double *__restrict__ r,
                                                      double *__restrict__ z, const double *__restrict__ p,
                                                      const double *__restrict__ q, const uint8_t *__restrict__ fixed,
                                                      const float *__restrict__ pinv, const double *__restrict__ scal,
                                                      double *__restrict__ part_rz, double *__restrict__ part_rr)
{
    __shared__ double sh[BLOCK];
    const double a = scal[S_ALPHA];
    double rz = 0, rr = 0;
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < n_nodes; i += DOT_BLOCKS * BLOCK) {
        const int64_t o = 3 * int64_t(i);
        // every load of the node first, then the masked update
        double ri[3], zi[3], pi[3], qi[3], yi[3];
        bool upd[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            ri[d] = r[o + d];
            pi[d] = p[o + d];
            qi[d] = q[o + d];
            yi[d] = y[o + d];
            upd[d] = a != 0.0 && !(fixed && fixed[o + d]);
        }
#pragma unroll
        for (int d = 0; d < 3; ++d)
            if (upd[d]) {
                y[o + d] = fma(a, pi[d], yi[d]);
                ri[d] = fma(-a, qi[d], ri[d]);
                r[o + d] = ri[d];
            }
        precondition(pinv, i, ri, zi);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            z[o + d] = zi[d];
            rz = fma(ri[d], zi[d], rz);
            rr = fma(ri[d], ri[d], rr);
        }
    }
    const double c = block_reduce(rz, sh), e = block_reduce(rr, sh);
    if (threadIdx.x == 0) {
        part_rz[blockIdx.x] = c;
        part_rr[blockIdx.x] = e;
    }
}

// sums of the r.z and r.r partials; at the start they set r.z and r.r, after an
// update they also give beta = r'.z' / r.z (zero once the iteration has stopped)
__global__ __launch_bounds__(BLOCK) void k_finish_pcg(const double *__restrict__ part_rz, const double *__restrict__ part_rr,
                                                      double *__restrict__ scal, int start)
{
    __shared__ double s[BLOCK];
    const double rz = sum_partials(part_rz, s), rr = sum_partials(part_rr, s);
    if (threadIdx.x != 0) return;
    if (start) {
        scal[S_RZ] = rz;
        scal[S_RR] = rr;
        scal[S_FLAG] = CG_RUNNING;
        scal[S_ITER] = 0.0;
        return;
    }
    const double old = scal[S_RZ];
    const bool live = scal[S_FLAG] == 0.0 && old > 0.0;
    scal[S_BETA] = live ? rz / old : 0.0;
    if (scal[S_FLAG] == 0.0) {
        scal[S_RZ] = rz;
        scal[S_RR] = rr;
        if (live) scal[S_ITER] += 1.0;
        // the iteration stops where it meets the tolerance
        if (rr <= scal[S_TOL]) scal[S_FLAG] = CG_TOLERANCE;
    }
}

// y = w H v + M v over the block-sparse pattern: H in the assembled value
// layout (block (i, pos) entry (m, n) at 9 brow_ptr[i] + 3 m deg + 3 pos + n),
// M one value per block (block = m I_3) or null.  SPMV_LANES lanes per block
// row; each row's sums are reduced by a fixed butterfly, so y is reproducible.
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_bsr_spmv(int n_nodes, const int32_t *__restrict__ brow_ptr,
                                                    const int32_t *__restrict__ bcol, const T *__restrict__ vals,
                                                    const double *__restrict__ mass, double w,
                                                    const double *__restrict__ v, double *__restrict__ y, SegSkip skip)
{
    const int lane = threadIdx.x % SPMV_LANES;
    const int i = blockIdx.x * (BLOCK / SPMV_LANES) + threadIdx.x / SPMV_LANES;
    double h0 = 0, h1 = 0, h2 = 0, m0 = 0, m1 = 0, m2 = 0;
    if (i < n_nodes && !skip.node(i)) {
        const int b0 = brow_ptr[i], deg = brow_ptr[i + 1] - b0;
        const T *row = vals + 9 * int64_t(b0);
        for (int pos = lane; pos < deg; pos += SPMV_LANES) {
            const int64_t c = 3 * int64_t(bcol[b0 + pos]);
            const double v0 = v[c], v1 = v[c + 1], v2 = v[c + 2];
            const T *r0 = row + 3 * pos, *r1 = r0 + 3 * deg, *r2 = r1 + 3 * deg;
            h0 = fma(double(r0[0]), v0, fma(double(r0[1]), v1, fma(double(r0[2]), v2, h0)));
            h1 = fma(double(r1[0]), v0, fma(double(r1[1]), v1, fma(double(r1[2]), v2, h1)));
            h2 = fma(double(r2[0]), v0, fma(double(r2[1]), v1, fma(double(r2[2]), v2, h2)));
            if (mass) {
                const double mm = mass[b0 + pos];
                m0 = fma(mm, v0, m0);
                m1 = fma(mm, v1, m1);
                m2 = fma(mm, v2, m2);
            }
        }
    }
#pragma unroll
    for (int o = SPMV_LANES / 2; o > 0; o >>= 1) {
        h0 += __shfl_xor(h0, o, SPMV_LANES);
        h1 += __shfl_xor(h1, o, SPMV_LANES);
        h2 += __shfl_xor(h2, o, SPMV_LANES);
        m0 += __shfl_xor(m0, o, SPMV_LANES);
        m1 += __shfl_xor(m1, o, SPMV_LANES);
        m2 += __shfl_xor(m2, o, SPMV_LANES);
    }
    if (i < n_nodes && lane < 3) {
        const double hv = lane == 0 ? h0 : lane == 1 ? h1 : h2, mv = lane == 0 ? m0 : lane == 1 ? m1 : m2;
        y[3 * int64_t(i) + lane] = fma(w, hv, mv);
    }
}

// inverse of every node's 3x3 diagonal block of scale (w H + M), with the
// rows and columns of fixed dofs replaced by the identity; a block that is
// not positive definite (an indefinite tangent) falls back to 1 / |diagonal|
// The diagonal blocks come from the assembled values (vals, pattern layout)
// or, when ndiag is given, from a node-diagonal array (9 values per node).
// dblk, when given, receives the unscaled block of w H + M with the rows and
// columns of fixed dofs zeroed (the regularization's D).
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_block_jacobi(int n_nodes, const int32_t *__restrict__ brow_ptr,
                                                        const int32_t *__restrict__ diag_pos, const T *__restrict__ vals,
                                                        const T *__restrict__ ndiag, const double *__restrict__ mass,
                                                        double w, const uint8_t *__restrict__ fixed, double scale,
                                                        float *__restrict__ pinv, float *__restrict__ dblk)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n_nodes) return;
    const int b0 = brow_ptr[i], deg = brow_ptr[i + 1] - b0, dp = diag_pos[i];
    const T *row = ndiag ? ndiag + 9 * int64_t(i) : vals + 9 * int64_t(b0) + 3 * dp;
    const int rs = ndiag ? 3 : 3 * deg;
    double B[3][3];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int n = 0; n < 3; ++n) B[m][n] = w * double(row[m * rs + n]) + (mass && m == n ? mass[b0 + dp] : 0.0);
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (fixed && fixed[3 * int64_t(i) + d])
#pragma unroll
            for (int k = 0; k < 3; ++k) B[d][k] = B[k][d] = 0.0;
    if (dblk)
#pragma unroll
        for (int m = 0; m < 9; ++m) dblk[9 * int64_t(i) + m] = float(B[m / 3][m % 3]);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
#pragma unroll
        for (int k = 0; k < 3; ++k) B[d][k] *= scale;
        if (fixed && fixed[3 * int64_t(i) + d]) B[d][d] = 1.0;
    }
    // symmetric part (the exact block is symmetric; rounding is not)
    const double a = B[0][0], b = 0.5 * (B[0][1] + B[1][0]), c = 0.5 * (B[0][2] + B[2][0]);
    const double d = B[1][1], e = 0.5 * (B[1][2] + B[2][1]), f = B[2][2];
    const double c00 = d * f - e * e, c01 = c * e - b * f, c02 = b * e - c * d;
    const double det = a * c00 + b * c01 + c * c02;
    float *P = pinv + PINV * int64_t(i);
    if (a > 0.0 && a * d - b * b > 0.0 && det > 0.0 && isfinite(det)) {
        const double id = 1.0 / det;
        const double c11 = a * f - c * c, c12 = b * c - a * e, c22 = a * d - b * b;
        P[0] = float(c00 * id), P[1] = float(c01 * id), P[2] = float(c02 * id);
        P[3] = float(c11 * id), P[4] = float(c12 * id), P[5] = float(c22 * id);
    } else {
        auto inv = [](double v) { return std::fabs(v) > 0.0 && isfinite(v) ? 1.0 / std::fabs(v) : 1.0; };
        P[0] = float(inv(a)), P[1] = 0.0f, P[2] = 0.0f, P[3] = float(inv(d)), P[4] = 0.0f, P[5] = float(inv(f));
    }
}

// y += delta D v, D the 3x3 diagonal blocks (9 floats per node)
__global__ __launch_bounds__(BLOCK) void k_block_diag_add(int n_nodes, const float *__restrict__ dblk, double delta,
                                                          const double *__restrict__ v, double *__restrict__ y)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n_nodes) return;
    const float *D = dblk + 9 * int64_t(i);
    const double v0 = v[3 * int64_t(i)], v1 = v[3 * int64_t(i) + 1], v2 = v[3 * int64_t(i) + 2];
#pragma unroll
    for (int m = 0; m < 3; ++m)
        y[3 * int64_t(i) + m] += delta * (double(D[3 * m]) * v0 + double(D[3 * m + 1]) * v1 + double(D[3 * m + 2]) * v2);
}

// q := 0 on fixed dofs
__global__ __launch_bounds__(BLOCK) void k_mask(double *__restrict__ q, const uint8_t *__restrict__ fixed, int n)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n && fixed && fixed[i]) q[i] = 0.0;
}

// p = z + beta p
__global__ __launch_bounds__(BLOCK) void k_cg_direction(double *__restrict__ p, const double *__restrict__ z,
                                                        const double *__restrict__ scal, int n)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    p[i] = fma(scal[S_BETA], p[i], z[i]);
}

__global__ void k_set_scalar(double *__restrict__ scal, int slot, double v) { scal[slot] = v; }

// out = x + a d (a read from the host argument), or -v on free dofs
__global__ __launch_bounds__(BLOCK) void k_axpy_to(double *__restrict__ out, const double *__restrict__ x, double a,
                                                   const double *__restrict__ d, int n)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) out[i] = fma(a, d[i], x[i]);
}
__global__ __launch_bounds__(BLOCK) void k_neg_free(double *__restrict__ out, const double *__restrict__ v,
                                                    const uint8_t *__restrict__ fixed, int n)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) out[i] = fixed && fixed[i] ? 0.0 : -v[i];
}

// y = a y + M v (M: one value per 3x3 block of the node pattern, block = m I_3);
// SPMV_LANES lanes per node row, reduced by a fixed butterfly
__global__ __launch_bounds__(BLOCK) void k_mass_apply(int n_nodes, const int32_t *__restrict__ brow_ptr,
                                                      const int32_t *__restrict__ bcol, const double *__restrict__ m,
                                                      const double *__restrict__ v, double a, double *__restrict__ y)
{
    const int lane = threadIdx.x % SPMV_LANES;
    const int i = blockIdx.x * (BLOCK / SPMV_LANES) + threadIdx.x / SPMV_LANES;
    double s0 = 0, s1 = 0, s2 = 0;
    if (i < n_nodes)
        for (int k = brow_ptr[i] + lane; k < brow_ptr[i + 1]; k += SPMV_LANES) {
            const double w = m[k];
            const int64_t j = 3 * int64_t(bcol[k]);
            s0 = fma(w, v[j], s0);
            s1 = fma(w, v[j + 1], s1);
            s2 = fma(w, v[j + 2], s2);
        }
#pragma unroll
    for (int o = SPMV_LANES / 2; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o, SPMV_LANES);
        s1 += __shfl_xor(s1, o, SPMV_LANES);
        s2 += __shfl_xor(s2, o, SPMV_LANES);
    }
    if (i < n_nodes && lane < 3) {
        const int64_t o = 3 * int64_t(i) + lane;
        const double sv = lane == 0 ? s0 : lane == 1 ? s1 : s2;
        y[o] = a == 0.0 ? sv : fma(a, y[o], sv);
    }
}

// vals = w vals + M over the block-sparse layout (the objective's Hessian w H + M
// in place of H); SPMV_LANES lanes per node row
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_fold_operator(int n_nodes, const int32_t *__restrict__ brow_ptr,
                                                         const double *__restrict__ mass, double w, T *__restrict__ vals)
{
    const int lane = threadIdx.x % SPMV_LANES;
    const int i = blockIdx.x * (BLOCK / SPMV_LANES) + threadIdx.x / SPMV_LANES;
    if (i >= n_nodes) return;
    const int b0 = brow_ptr[i], deg = brow_ptr[i + 1] - b0;
    T *row = vals + 9 * int64_t(b0);
    for (int pos = lane; pos < deg; pos += SPMV_LANES) {
        const double mm = mass ? mass[b0 + pos] : 0.0;
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                T &e = row[m * 3 * deg + 3 * pos + n];
                e = T(fma(w, double(e), m == n ? mm : 0.0));
            }
    }
}

// out = x - xt
__global__ __launch_bounds__(BLOCK) void k_sub(double *__restrict__ out, const double *__restrict__ x,
                                               const double *__restrict__ xt, int n)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) out[i] = x[i] - xt[i];
}

// g = w g + mt - w f   (mt, f may be null)
__global__ __launch_bounds__(BLOCK) void k_combine_gradient(double *__restrict__ g, double w, const double *__restrict__ mt,
                                                            const double *__restrict__ f, int n)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    double v = w * g[i];
    if (mt) v += mt[i];
    if (f) v = fma(-w, f[i], v);
    g[i] = v;
}

// y = (x - x_prev) / dt  and  xt = x + dt y (the next step's predictor)
__global__ __launch_bounds__(BLOCK) void k_euler_update(const double *__restrict__ x, const double *__restrict__ xp,
                                                        double dt, double *__restrict__ v, int n)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) v[i] = (x[i] - xp[i]) / dt;
}

// out = a u + b v (either may be null)
__global__ __launch_bounds__(BLOCK) void k_lincomb(double *__restrict__ out, double a, const double *__restrict__ u, double b,
                                                   const double *__restrict__ v, int n)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    double s = 0;
    if (u) s = a * u[i];
    if (v) s = fma(b, v[i], s);
    out[i] = s;
}

// out += a v
__global__ __launch_bounds__(BLOCK) void k_axpy(double *__restrict__ out, double a, const double *__restrict__ v, int n)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) out[i] = fma(a, v[i], out[i]);
}

struct Work {
    double *r, *p, *q, *y, *g, *xt, *d, *t, *mt, *z, *scal, *part, *part2;
};

Work workspace(Handle &h)
{
    const size_t n = size_t(h.n_dofs);
    h.sv.reserve(sizeof(double) * (10 * n + S_COUNT + 2 * DOT_BLOCKS));
    double *b = h.sv.as<double>();
    return Work{b,         b + n,     b + 2 * n,           b + 3 * n, b + 4 * n, b + 5 * n, b + 6 * n, b + 7 * n, b + 8 * n,
                b + 9 * n, b + 10 * n, b + 10 * n + S_COUNT, b + 10 * n + S_COUNT + DOT_BLOCKS};
}

void dot(Handle &h, const Work &w, const double *a, const double *b, const uint8_t *fixed, int slot, int after)
{
    k_dot_partials<<<DOT_BLOCKS, BLOCK, 0, h.stream>>>(a, b, fixed, h.n_dofs, w.part);
    k_finish_dot<<<1, BLOCK, 0, h.stream>>>(w.part, w.scal, slot, after);
}

double read_scalar(Handle &h, const Work &w, int slot)
{
    double v = 0;
    PF_HIP(hipMemcpyAsync(&v, w.scal + slot, sizeof(double), hipMemcpyDeviceToHost, h.stream));
    PF_HIP(hipStreamSynchronize(h.stream));
    return v;
}

bool inertia(const Handle &h) { return h.obj_xt != nullptr; }

bool assembled_operator(const Handle &h)
{
    if (!h.groups.empty()) return false; // mixed batches: matrix-free
    return h.solver_hessian == PFGPU_SOLVER_HESSIAN_ASSEMBLED || h.solver_hessian == PFGPU_SOLVER_HESSIAN_ASSEMBLED_PSD ||
           (h.solver_hessian == PFGPU_SOLVER_HESSIAN_AUTO && h.order == 1);
}

// the regularized operator's term: y += delta D v
void add_regularization(Handle &h, const double *v, double *y)
{
    if (h.op_valid && h.op_reg > 0.0)
        k_block_diag_add<<<grid_for(h.n_nodes), BLOCK, 0, h.stream>>>(h.n_nodes, h.reg_diag.as<float>(), h.op_reg, v, y);
}

// y = (w H(x) + [M] + [delta D]) v: the block-sparse product with the values
// assembled by prepare_operator at x, or the matrix-free product
void hvp_any(Handle &h, const double *x, const double *v, double *y)
{
    if (h.op_valid && h.op_assembled) {
        // the values hold w H + M (folded by prepare_operator)
        const unsigned g = grid_for(h.n_nodes, BLOCK / SPMV_LANES);
        auto spmv = [&]<typename T>() {
            k_bsr_spmv<T><<<g, BLOCK, 0, h.stream>>>(h.n_nodes, h.brow_ptr.as<int32_t>(), h.pat_bcol.as<int32_t>(),
                                                     h.solver_vals.as<T>(), nullptr, 1.0, v, y, h.skip);
        };
        if (h.precision == PFGPU_MIXED)
            spmv.template operator()<float>();
        else
            spmv.template operator()<double>();
        add_regularization(h, v, y);
        PF_LAUNCH_CHECK();
        return;
    }
    if (h.op_valid && h.op_records)
        factored_hvp(h, v, y); // the records prepare_operator computed at x
    else if (h.path == PFGPU_PATH_FUSED)
        fused_hvp(h, x, v, y);
    else
        plain_hvp(h, x, v, y);
    if (inertia(h))
        k_mass_apply<<<grid_for(h.n_nodes, BLOCK / SPMV_LANES), BLOCK, 0, h.stream>>>(h.n_nodes, h.brow_ptr.as<int32_t>(),
                                                                   h.pat_bcol.as<int32_t>(), h.mass.as<double>(), v,
                                                                   h.obj_w, y);
    else if (h.obj_w != 1.0)
        k_combine_gradient<<<grid_for(h.n_dofs), BLOCK, 0, h.stream>>>(y, h.obj_w, nullptr, nullptr, h.n_dofs);
    add_regularization(h, v, y);
    PF_LAUNCH_CHECK();
}

void build_preconditioner(Handle &h, double scale, bool diag);

// The CG operator at x: with an assembled operator or the block-Jacobi
// preconditioner, the Hessian values at x (plain or fused assembly) into the
// solver's value buffer and, for the preconditioner, the inverse diagonal
// blocks of w H + M with the fixed dofs as identity.  Valid until x, the
// objective or the mask change (op_valid is cleared then).
void prepare_operator(Handle &h, const double *x, const uint8_t *fixed)
{
    h.op_assembled = assembled_operator(h);
    h.op_precond = h.solver_precond == PFGPU_PRECOND_BLOCK_JACOBI;
    h.op_records = false;
    h.op_reg = 0.0;
    // a matrix-free operator on the factored path needs only the diagonal
    // blocks for its preconditioner: those come from the factor records
    const bool diag_only = !h.op_assembled && h.op_precond && h.path == PFGPU_PATH_FACTORED &&
                           h.solver_hessian != PFGPU_SOLVER_HESSIAN_ASSEMBLED_PSD &&
                           (h.material == PFGPU_NEOHOOKEAN || h.material == PFGPU_LINEAR_ELASTICITY);
    const bool mixed_diag = !h.groups.empty() && h.op_precond;
    if (diag_only || mixed_diag) {
        if (!h.brow_ptr.get()) h.brow_ptr = upload(h.pat.brow_ptr, h.stream);
        h.node_diag.reserve(h.value_size() * 9 * size_t(h.n_nodes));
        if (mixed_diag) {
            mixed_node_diagonal(h, x, h.node_diag.get());
        } else {
            factored_node_diagonal(h, x, h.node_diag.get());
            h.op_records = true; // the matrix-free products of this operator use them
        }
    } else if (h.op_assembled || h.op_precond) {
        if (!h.brow_ptr.get()) h.brow_ptr = upload(h.pat.brow_ptr, h.stream);
        if (!h.pat_bcol.get()) h.pat_bcol = upload(h.pat.bcol, h.stream);
        h.solver_vals.reserve(h.value_size() * size_t(h.pat.nnz));
        if (h.solver_hessian == PFGPU_SOLVER_HESSIAN_ASSEMBLED_PSD)
            plain_hessian(h, x, h.solver_vals.get(), true);
        else if (h.path == PFGPU_PATH_FUSED)
            fused_energy_gradient_hessian(h, x, nullptr, nullptr, h.solver_vals.get());
        else
            plain_hessian(h, x, h.solver_vals.get());
    }
    const double *mass = inertia(h) ? h.mass.as<double>() : nullptr;
    const bool fold = h.op_assembled && (mass || h.obj_w != 1.0);
    if (fold) {
        auto f = [&]<typename T>() {
            k_fold_operator<T><<<grid_for(h.n_nodes, BLOCK / SPMV_LANES), BLOCK, 0, h.stream>>>(
                h.n_nodes, h.brow_ptr.as<int32_t>(), mass, h.obj_w, h.solver_vals.as<T>());
        };
        if (h.precision == PFGPU_MIXED)
            f.template operator()<float>();
        else
            f.template operator()<double>();
        PF_LAUNCH_CHECK();
    }
    if (h.op_precond) {
        if (!h.diag_pos.get()) {
            std::vector<int32_t> dp(size_t(h.n_nodes));
            for (int i = 0; i < h.n_nodes; ++i) {
                const int32_t *c0 = h.pat.bcol.data() + h.pat.brow_ptr[i], *c1 = h.pat.bcol.data() + h.pat.brow_ptr[i + 1];
                const int32_t *f = std::lower_bound(c0, c1, i);
                if (f == c1 || *f != i) fail(PFGPU_ERR_INVALID_STATE, "pattern without a diagonal block");
                dp[size_t(i)] = int32_t(f - c0);
            }
            h.diag_pos = upload(dp, h.stream);
        }
        h.op_ndiag = diag_only || mixed_diag;
        h.op_fold = fold;
        h.op_fixed = fixed;
        build_preconditioner(h, 1.0, false);
    }
    h.op_valid = true;
}

// the block-Jacobi preconditioner of the prepared operator scaled by scale
// (1 + delta for the regularized operator); with diag, also the unscaled
// diagonal blocks into reg_diag
void build_preconditioner(Handle &h, double scale, bool diag)
{
    const double *mass = inertia(h) ? h.mass.as<double>() : nullptr;
    h.pinv.reserve(sizeof(float) * PINV * size_t(h.n_nodes));
    if (diag) h.reg_diag.reserve(sizeof(float) * 9 * size_t(h.n_nodes));
    auto jac = [&]<typename T>() {
        k_block_jacobi<T><<<grid_for(h.n_nodes), BLOCK, 0, h.stream>>>(
            h.n_nodes, h.brow_ptr.as<int32_t>(), h.diag_pos.as<int32_t>(), h.solver_vals.as<T>(),
            h.op_ndiag ? h.node_diag.as<T>() : nullptr, h.op_fold ? nullptr : mass, h.op_fold ? 1.0 : h.obj_w,
            h.op_fixed, scale, h.pinv.as<float>(), diag ? h.reg_diag.as<float>() : nullptr);
    };
    if (h.precision == PFGPU_MIXED)
        jac.template operator()<float>();
    else
        jac.template operator()<double>();
    PF_LAUNCH_CHECK();
}

// the prepared operator becomes w H + M + delta D (delta = 0: unregularized);
// needs the block-Jacobi preconditioner, whose blocks D are
void set_regularization(Handle &h, double delta)
{
    if (!h.op_valid || !h.op_precond) fail(PFGPU_ERR_INVALID_STATE, "regularization needs the block-Jacobi preconditioner");
    if (delta == h.op_reg) return;
    build_preconditioner(h, 1.0 + delta, delta > 0.0);
    h.op_reg = delta;
}

// objective gradient into g and its value into scal[slot] (device)
void objective_gradient(Handle &h, const Work &w, const double *x, double *g, int slot);
// objective value into scal[slot] (device)
void objective_energy(Handle &h, const Work &w, const double *x, int slot);

const float *precond(const Handle &h) { return h.op_valid && h.op_precond ? h.pinv.as<float>() : nullptr; }

// PCG start from b (a fixed kernel sequence)
void cg_start(Handle &h, const Work &w, const uint8_t *fixed, const double *b)
{
    k_pcg_start<<<DOT_BLOCKS, BLOCK, 0, h.stream>>>(h.n_nodes, b, fixed, precond(h), w.y, w.r, w.z, w.p, w.part,
                                                    w.part2);
    k_finish_pcg<<<1, BLOCK, 0, h.stream>>>(w.part, w.part2, w.scal, 1);
    PF_LAUNCH_CHECK();
}

// one PCG iteration (a fixed kernel sequence)
void cg_iteration(Handle &h, const Work &w, const double *x, const uint8_t *fixed)
{
    hvp_any(h, x, w.p, w.q);
    dot(h, w, w.p, w.q, fixed, S_PQ, AFTER_ALPHA);
    k_pcg_update<<<DOT_BLOCKS, BLOCK, 0, h.stream>>>(h.n_nodes, w.y, w.r, w.z, w.p, w.q, fixed, precond(h), w.scal,
                                                     w.part, w.part2);
    k_finish_pcg<<<1, BLOCK, 0, h.stream>>>(w.part, w.part2, w.scal, 0);
    k_cg_direction<<<grid_for(h.n_dofs), BLOCK, 0, h.stream>>>(w.p, w.z, w.scal, h.n_dofs);
    PF_LAUNCH_CHECK();
}

void objective_energy(Handle &h, const Work &w, const double *x, int slot)
{
    // elastic part into scal[slot]; the host combines the parts
    plain_energy(h, x, w.scal + slot, nullptr);
    if (inertia(h)) {
        const unsigned g = grid_for(h.n_dofs);
        k_sub<<<g, BLOCK, 0, h.stream>>>(w.t, x, h.obj_xt, h.n_dofs);
        k_mass_apply<<<grid_for(h.n_nodes, BLOCK / SPMV_LANES), BLOCK, 0, h.stream>>>(h.n_nodes, h.brow_ptr.as<int32_t>(),
                                                                   h.pat_bcol.as<int32_t>(), h.mass.as<double>(), w.t,
                                                                   0.0, w.mt);
        dot(h, w, w.t, w.mt, nullptr, S_M, AFTER_NONE);
    }
    if (h.obj_f) dot(h, w, h.obj_f, x, nullptr, S_F, AFTER_NONE);
    PF_LAUNCH_CHECK();
}

void objective_gradient(Handle &h, const Work &w, const double *x, double *g, int slot)
{
    if (h.path == PFGPU_PATH_FUSED)
        fused_gradient(h, x, g, w.scal + slot);
    else
        plain_gradient(h, x, g, w.scal + slot);
    const double *mt = nullptr;
    if (inertia(h)) {
        const unsigned gr = grid_for(h.n_dofs);
        k_sub<<<gr, BLOCK, 0, h.stream>>>(w.t, x, h.obj_xt, h.n_dofs);
        k_mass_apply<<<grid_for(h.n_nodes, BLOCK / SPMV_LANES), BLOCK, 0, h.stream>>>(h.n_nodes, h.brow_ptr.as<int32_t>(),
                                                                   h.pat_bcol.as<int32_t>(), h.mass.as<double>(), w.t,
                                                                   0.0, w.mt);
        dot(h, w, w.t, w.mt, nullptr, S_M, AFTER_NONE);
        mt = w.mt;
    }
    if (h.obj_f) dot(h, w, h.obj_f, x, nullptr, S_F, AFTER_NONE);
    if (mt || h.obj_f || h.obj_w != 1.0)
        k_combine_gradient<<<grid_for(h.n_dofs), BLOCK, 0, h.stream>>>(g, h.obj_w, mt, h.obj_f, h.n_dofs);
    PF_LAUNCH_CHECK();
}

// Rounding level of an objective value e whose elastic part is e_el: the FP64
// sums contribute a few ulps of |e|; FP32 element energies (mixed precision)
// contribute a few FP32 ulps of the weighted elastic part.  Energy differences
// below this are not resolved by the line search.
double roundoff_noise(const Handle &h, double e, double e_el)
{
    const double fp32 = h.precision == PFGPU_MIXED ? 16.0 * 5.96e-8 * std::fabs(h.obj_w * e_el) : 0.0;
    return 64.0 * 2.2e-16 * std::fabs(e) + fp32;
}

// host value of the objective whose parts objective_energy/gradient left in scal
double objective_value(Handle &h, const Work &w, int slot)
{
    double e = h.obj_w * read_scalar(h, w, slot);
    if (inertia(h)) e += 0.5 * read_scalar(h, w, S_M);
    if (h.obj_f) e -= h.obj_w * read_scalar(h, w, S_F);
    return e;
}

} // namespace

void mass_apply(Handle &h, const double *v, double *y)
{
    if (!h.has_mass) fail(PFGPU_ERR_INVALID_STATE, "no mass matrix: call pfgpu_set_density first");
    k_mass_apply<<<grid_for(h.n_nodes, BLOCK / SPMV_LANES), BLOCK, 0, h.stream>>>(h.n_nodes, h.brow_ptr.as<int32_t>(), h.pat_bcol.as<int32_t>(),
                                                               h.mass.as<double>(), v, 0.0, y);
    PF_LAUNCH_CHECK();
}

namespace {
void release_graph(Handle &h)
{
    if (h.cg_graph) (void)hipGraphExecDestroy(static_cast<hipGraphExec_t>(h.cg_graph));
    h.cg_graph = nullptr;
    h.cg_graph_key = {};
}
} // namespace

void solver_release(Handle &h)
{
    h.op_valid = false;
    release_graph(h);
}

// CG on H(x)_ff y = b_f starting from y = 0; the solution is left in the
// workspace's y.  Returns the iterations taken; *rel = ||r|| / ||b_f||.
int cg_solve_ws(Handle &h, const double *x, const uint8_t *fixed, const double *b, int max_it, double tol,
                bool use_graph, double *rel, int *exit_reason)
{
    const Work w = workspace(h);
    if (!h.op_valid) prepare_operator(h, x, fixed);
    cg_start(h, w, fixed, b);
    const double b2 = read_scalar(h, w, S_RR);
    if (rel) *rel = 0.0;
    if (exit_reason) *exit_reason = CG_TOLERANCE;
    if (b2 == 0.0 || max_it <= 0) return 0;
    if (!std::isfinite(b2)) fail(PFGPU_ERR_INVALID_STATE, "non-finite right-hand side in the CG solve");
    // the device stops the iteration where r.r falls to tol^2 b.b
    k_set_scalar<<<1, 1, 0, h.stream>>>(w.scal, S_TOL, tol * tol * b2);
    PF_LAUNCH_CHECK();
    // the graph replays one iteration; it is rebuilt when its arguments change
    // (scratch buffers are grow-only: a reallocation by another call rebuilds the graph)
    intptr_t reg_bits = 0;
    static_assert(sizeof(reg_bits) == sizeof(h.op_reg), "the regularization weight is keyed by its bits");
    std::memcpy(&reg_bits, &h.op_reg, sizeof(reg_bits));
    const std::array<const void *, 9> key{x,
                                          fixed,
                                          w.r,
                                          reinterpret_cast<const void *>(intptr_t(h.obj_generation)),
                                          h.loc.get(),
                                          reinterpret_cast<const void *>(intptr_t(h.op_assembled ? 1 : 0) |
                                                                         intptr_t(precond(h) ? 2 : 0)),
                                          h.solver_vals.get(),
                                          h.pinv.get(),
                                          reinterpret_cast<const void *>(reg_bits)};
    if (use_graph && (h.cg_graph == nullptr || h.cg_graph_key != key)) {
        release_graph(h); // (the operator prepared above stays valid)
        cg_iteration(h, w, x, fixed); // sizes every scratch buffer before capture
        cg_start(h, w, fixed, b);
        hipGraph_t graph = nullptr;
        PF_HIP(hipStreamBeginCapture(h.stream, hipStreamCaptureModeThreadLocal));
        try {
            for (int k = 0; k < CHECK; ++k) cg_iteration(h, w, x, fixed);
        } catch (...) {
            (void)hipStreamEndCapture(h.stream, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        PF_HIP(hipStreamEndCapture(h.stream, &graph));
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        PF_HIP(e);
        h.cg_graph = exec;
        h.cg_graph_key = key;
        h.cg_graph_key[4] = h.loc.get(); // as sized by the eager pass
    }
    int it = 0, done = 0;
    double rr = b2, flag = CG_RUNNING;
    while (it < max_it) {
        // the graph holds CHECK iterations; a shorter remainder runs eagerly
        // (iterations after a stop change nothing)
        const int n = std::min(CHECK, max_it - it);
        if (use_graph && n == CHECK)
            PF_HIP(hipGraphLaunch(static_cast<hipGraphExec_t>(h.cg_graph), h.stream));
        else
            for (int k = 0; k < n; ++k) cg_iteration(h, w, x, fixed);
        it += n;
        double sc[S_COUNT];
        PF_HIP(hipMemcpyAsync(sc, w.scal, sizeof(sc), hipMemcpyDeviceToHost, h.stream));
        PF_HIP(hipStreamSynchronize(h.stream));
        rr = sc[S_RR];
        flag = sc[S_FLAG];
        done = int(sc[S_ITER]);
        if (!std::isfinite(rr)) fail(PFGPU_ERR_INVALID_STATE, "non-finite residual in the CG solve (inverted element?)");
        if (flag != CG_RUNNING) break;
    }
    if (rel) *rel = std::sqrt(rr / b2);
    if (exit_reason) *exit_reason = flag == CG_RUNNING ? PFGPU_CG_ITERATION_CAP : (flag == CG_CURVATURE ? PFGPU_CG_CURVATURE : PFGPU_CG_TOLERANCE);
    return done;
}

double *cg_solution(Handle &h) { return workspace(h).y; }

void newton_solve(Handle &h, double *x, const uint8_t *fixed, const pfgpu_newton_params &prm, pfgpu_newton_report &rep)
{
    const Work w = workspace(h);
    const unsigned gr = grid_for(h.n_dofs);
    rep = pfgpu_newton_report{};
    if (inertia(h) && !h.has_mass) fail(PFGPU_ERR_INVALID_STATE, "inertia term without a mass matrix");
    double gn0 = 0; // gradient norm at the initial guess
    // regularized Newton: delta of H + delta D, kept across iterations
    const bool regularize = prm.regularization_min > 0.0 && h.solver_precond == PFGPU_PRECOND_BLOCK_JACOBI;
    double reg = 0.0;
    auto energy_at = [&](const double *xx) {
        objective_energy(h, w, xx, S_E1);
        return objective_value(h, w, S_E1);
    };
    for (int it = 0; it < prm.max_iterations; ++it) {
        h.op_valid = false; // a new iterate: the CG operator is rebuilt at x
        // energy and gradient at x; the reduced gradient norm
        objective_gradient(h, w, x, w.g, S_E0);
        dot(h, w, w.g, w.g, fixed, S_B2, AFTER_NONE);
        const double gn = std::sqrt(read_scalar(h, w, S_B2));
        const double e0 = objective_value(h, w, S_E0);
        rep.energy = e0;
        rep.grad_norm = gn;
        if (!std::isfinite(gn) || !std::isfinite(e0)) fail(PFGPU_ERR_INVALID_STATE, "non-finite energy or gradient at the Newton iterate");
        if (it == 0) {
            gn0 = gn;
            rep.initial_grad_norm = gn;
        }
        if (gn <= prm.grad_norm || gn <= prm.relative_grad_norm * gn0) {
            rep.converged = 1;
            return;
        }
        // Newton direction: H_ff d = -g_f; on non-positive curvature, again
        // on H + delta D with a growing delta (regularized Newton)
        k_neg_free<<<gr, BLOCK, 0, h.stream>>>(w.d, w.g, fixed, h.n_dofs);
        PF_LAUNCH_CHECK();
        double rel = 0;
        int why = 0;
        while (true) {
            if (reg > 0.0) {
                if (!h.op_valid) prepare_operator(h, x, fixed);
                set_regularization(h, reg);
            }
            rep.cg_iterations +=
                cg_solve_ws(h, x, fixed, w.d, prm.cg_max_iterations, prm.cg_tolerance, prm.use_graph != 0, &rel, &why);
            if (why != PFGPU_CG_CURVATURE || !regularize || reg >= prm.regularization_max) break;
            reg = reg > 0.0 ? std::min(10.0 * reg, prm.regularization_max) : prm.regularization_min;
            rep.regularized_solves++;
        }
        const double reg_used = reg;
        if (why == PFGPU_CG_CURVATURE)
            rep.cg_stops_curvature++;
        else if (why == PFGPU_CG_ITERATION_CAP)
            rep.cg_stops_cap++;
        else
            rep.cg_stops_tolerance++;
        // direction in w.y; g.d for the Armijo condition
        dot(h, w, w.g, w.y, fixed, S_GD, AFTER_NONE);
        double gd = read_scalar(h, w, S_GD);
        const double *dir = w.y;
        if (!(gd < 0.0)) {
            // not a descent direction (indefinite Hessian): steepest descent
            k_neg_free<<<gr, BLOCK, 0, h.stream>>>(w.y, w.g, fixed, h.n_dofs);
            PF_LAUNCH_CHECK();
            gd = -gn * gn;
        }
        // backtracking from alpha0 along dir; true when a step was accepted
        double alpha = 1.0;
        auto search = [&](const double *dd, double slope, double alpha0) {
            alpha = alpha0;
            for (int ls = 0; ls <= prm.max_line_search; ++ls) {
                k_axpy_to<<<gr, BLOCK, 0, h.stream>>>(w.xt, x, alpha, dd, h.n_dofs);
                PF_LAUNCH_CHECK();
                if (prm.line_search == 0) return true;
                const double e1 = energy_at(w.xt);
                // a NaN energy is an inverted element: rejected like an invalid step
                if (std::isfinite(e1) && e1 - e0 <= prm.armijo_c * alpha * slope) return true;
                alpha *= 0.5;
                rep.line_search_halvings++;
            }
            return false;
        };
        // Near the minimum the predicted decrease g.d falls below the rounding
        // of the energy and the Armijo test cannot resolve it: the full Newton
        // step is taken (the quadratic model is exact to that precision) if it
        // keeps every element valid.
        const double noise = roundoff_noise(h, e0, read_scalar(h, w, S_E0));
        bool accepted = false;
        if (prm.line_search != 0 && -gd <= noise) {
            k_axpy_to<<<gr, BLOCK, 0, h.stream>>>(w.xt, x, 1.0, dir, h.n_dofs);
            PF_LAUNCH_CHECK();
            alpha = 1.0;
            accepted = std::isfinite(energy_at(w.xt));
        }
        if (!accepted) accepted = search(dir, gd, 1.0);
        if (!accepted) {
            // Newton direction rejected (near-singular or indefinite Hessian):
            // steepest descent from the Cauchy step g^T g / g^T H g of the
            // quadratic model, or a unit-length step when the curvature along
            // g is not positive
            k_neg_free<<<gr, BLOCK, 0, h.stream>>>(w.y, w.g, fixed, h.n_dofs);
            hvp_any(h, x, w.y, w.q);
            dot(h, w, w.y, w.q, fixed, S_GD, AFTER_NONE);
            const double c = read_scalar(h, w, S_GD);
            const double a0 = c > 0.0 ? gn * gn / c : 1.0 / gn;
            accepted = search(w.y, -gn * gn, a0);
            rep.gradient_steps++;
        }
        if (!accepted) fail(PFGPU_ERR_INVALID_STATE, "line search found no acceptable step");
        PF_HIP(hipMemcpyAsync(x, w.xt, sizeof(double) * size_t(h.n_dofs), hipMemcpyDeviceToDevice, h.stream));
        rep.iterations = it + 1;
        rep.last_alpha = alpha;
        rep.last_regularization = reg_used;
        if (reg > 0.0) reg = reg / 10.0 < prm.regularization_min ? 0.0 : reg / 10.0;
    }
    // final state
    objective_gradient(h, w, x, w.g, S_E0);
    dot(h, w, w.g, w.g, fixed, S_B2, AFTER_NONE);
    rep.grad_norm = std::sqrt(read_scalar(h, w, S_B2));
    rep.energy = objective_value(h, w, S_E0);
    rep.converged = rep.grad_norm <= prm.grad_norm || rep.grad_norm <= prm.relative_grad_norm * gn0;
}

// ---- implicit Euler with checkpointed adjoint --------------------------
//
// Forward (ImplicitEuler.cpp:13-31; InertiaForm.cpp:17-35): with
// xt_i = x_{i-1} + dt v_{i-1}, x_i minimises
//   1/2 (x - xt_i)^T M (x - xt_i) + dt^2 (E(x) - f^T x)
// over the free dofs (fixed dofs keep their x0 values), and
// v_i = (x_i - x_{i-1}) / dt.  (x_i, v_i) is kept for i = 0 and every
// `every` steps; the backward pass recomputes the states in between.
//
// Backward: the step residuals R_i = M (x_i - xt_i) + dt^2 (g(x_i) - f)
// vanish at every step; with J = J(x_T, v_T) and multipliers p_i the
// stationarity conditions give, on the free dofs,
//   (M + dt^2 H(x_i)) p_i = dJ/dx_i + M (2 p_{i+1} - p_{i+2}),
//   dJ/dx_T += (1/dt) dJ/dv_T,  dJ/dx_{T-1} = -(1/dt) dJ/dv_T,
// and the total derivatives
//   dJ/dtheta = sum_i dt^2 (-p_i^T dg(x_i)/dtheta)   (material VJP at x_i),
//   dJ/dv_0 = dt M p_1,   dJ/dx_0 = M (p_1 - p_2),
// zero on fixed dofs.  This is the BDF-1 case of the reference's backward
// recurrence (StateDiff.cpp:320-413 with dJ_material_transient_adjoint_term,
// AdjointTools.cpp:852-887, and dJ_initial_condition_adjoint_term,
// AdjointTools.cpp:992-1010), written in one multiplier per step.

namespace {

struct ObjectiveScope {
    Handle &h;
    double w;
    const double *xt, *f;
    ObjectiveScope(Handle &hh, double ww, const double *x_tilde, const double *ff)
        : h(hh), w(hh.obj_w), xt(hh.obj_xt), f(hh.obj_f)
    {
        h.op_valid = false;
        h.obj_w = ww;
        h.obj_xt = x_tilde;
        h.obj_f = ff;
        h.obj_generation++;
    }
    ~ObjectiveScope()
    {
        h.op_valid = false;
        h.obj_w = w;
        h.obj_xt = xt;
        h.obj_f = f;
        h.obj_generation++;
    }
};

// one forward step: x_out minimises the step objective from (x_prev, v_prev);
// v_out = (x_out - x_prev) / dt.  xt is a scratch vector.
void euler_step(Handle &h, const double *xp, const double *vp, double *xt, double *x_out, double *v_out)
{
    const int n = h.n_dofs;
    const unsigned g = grid_for(n);
    const uint8_t *fixed = h.tr_has_fixed ? h.tr_fixed.as<uint8_t>() : nullptr;
    k_lincomb<<<g, BLOCK, 0, h.stream>>>(xt, 1.0, xp, h.tr_dt, vp, n);
    PF_HIP(hipMemcpyAsync(x_out, xp, sizeof(double) * size_t(n), hipMemcpyDeviceToDevice, h.stream));
    {
        ObjectiveScope os(h, h.tr_dt * h.tr_dt, xt, h.tr_has_f ? h.tr_f.as<double>() : nullptr);
        pfgpu_newton_report rep{};
        newton_solve(h, x_out, fixed, h.tr_newton, rep);
        if (h.tr_recompute) {
            h.tr_stats.recompute_newton_iterations += rep.iterations;
            h.tr_stats.recompute_cg_iterations += rep.cg_iterations;
        } else {
            h.tr_stats.steps++;
            h.tr_stats.newton_iterations += rep.iterations;
            h.tr_stats.cg_iterations += rep.cg_iterations;
            h.tr_stats.line_search_halvings += rep.line_search_halvings;
        }
        if (!rep.converged) fail(PFGPU_ERR_INVALID_STATE, "a time step's Newton solve did not converge");
    }
    k_euler_update<<<g, BLOCK, 0, h.stream>>>(x_out, xp, h.tr_dt, v_out, n);
    PF_LAUNCH_CHECK();
}

} // namespace

void transient_forward(Handle &h, const double *x0, const double *v0, const uint8_t *fixed, const double *f, double dt,
                       int steps, int every, const pfgpu_newton_params &np, double *xT, double *vT)
{
    if (!h.has_mass) fail(PFGPU_ERR_INVALID_STATE, "no mass matrix: call pfgpu_set_density first");
    const size_t n = size_t(h.n_dofs), bytes = sizeof(double) * n;
    h.tr_stats = pfgpu_transient_stats{};
    h.tr_recompute = false;
    h.tr_dt = dt;
    h.tr_steps = steps;
    h.tr_every = every;
    h.tr_newton = np;
    h.tr_has_fixed = fixed != nullptr;
    if (fixed) {
        h.tr_fixed.reserve(n);
        PF_HIP(hipMemcpyAsync(h.tr_fixed.get(), fixed, n, hipMemcpyDeviceToDevice, h.stream));
    }
    h.tr_has_f = f != nullptr;
    if (f) {
        h.tr_f.reserve(bytes);
        PF_HIP(hipMemcpyAsync(h.tr_f.get(), f, bytes, hipMemcpyDeviceToDevice, h.stream));
    }
    const int n_ck = steps / every + 1;
    h.ckpt.clear();
    h.ckpt.resize(size_t(n_ck));
    for (auto &c : h.ckpt) c.reserve(2 * bytes);
    PF_HIP(hipMemcpyAsync(h.ckpt[0].as<double>(), x0, bytes, hipMemcpyDeviceToDevice, h.stream));
    PF_HIP(hipMemcpyAsync(h.ckpt[0].as<double>() + n, v0, bytes, hipMemcpyDeviceToDevice, h.stream));
    // two ping-pong states and the predictor
    h.tr_seg.reserve(5 * bytes);
    double *S = h.tr_seg.as<double>();
    double *xa = S, *va = S + n, *xb = S + 2 * n, *vb = S + 3 * n, *xt = S + 4 * n;
    PF_HIP(hipMemcpyAsync(xa, x0, bytes, hipMemcpyDeviceToDevice, h.stream));
    PF_HIP(hipMemcpyAsync(va, v0, bytes, hipMemcpyDeviceToDevice, h.stream));
    for (int i = 1; i <= steps; ++i) {
        euler_step(h, xa, va, xt, xb, vb);
        std::swap(xa, xb);
        std::swap(va, vb);
        if (i % every == 0) {
            PF_HIP(hipMemcpyAsync(h.ckpt[i / every].as<double>(), xa, bytes, hipMemcpyDeviceToDevice, h.stream));
            PF_HIP(hipMemcpyAsync(h.ckpt[i / every].as<double>() + n, va, bytes, hipMemcpyDeviceToDevice, h.stream));
        }
    }
    if (xT) PF_HIP(hipMemcpyAsync(xT, xa, bytes, hipMemcpyDeviceToDevice, h.stream));
    if (vT) PF_HIP(hipMemcpyAsync(vT, va, bytes, hipMemcpyDeviceToDevice, h.stream));
}

void transient_adjoint(Handle &h, const double *dJdx, const double *dJdv, int cg_max, double cg_tol, double *dtheta,
                       double *dx0, double *dv0)
{
    if (h.tr_steps <= 0 || h.ckpt.empty()) fail(PFGPU_ERR_INVALID_STATE, "no forward run to differentiate");
    const int T = h.tr_steps, K = h.tr_every, n = h.n_dofs, E = h.n_elements;
    const size_t bytes = sizeof(double) * size_t(n);
    const unsigned g = grid_for(n);
    const double dt = h.tr_dt;
    const uint8_t *fixed = h.tr_has_fixed ? h.tr_fixed.as<uint8_t>() : nullptr;
    // segment states x_{c..c+K}, v_{c..c+K}; multipliers p_{i+1}, p_{i+2}; rhs;
    // VJP of one step; running sum of the material terms
    DevBuf seg(size_t(2 * (K + 1)) * bytes), pv(size_t(4) * bytes), acc(sizeof(double) * 4 * size_t(E));
    double *SX = seg.as<double>(), *SV = SX + size_t(K + 1) * n;
    double *p1 = pv.as<double>(), *p2 = p1 + n, *rhs = p2 + n, *tmp = rhs + n;
    double *vjp = acc.as<double>(), *sum = vjp + 2 * size_t(E);
    k_lincomb<<<g, BLOCK, 0, h.stream>>>(p1, 0.0, nullptr, 0.0, nullptr, n);
    k_lincomb<<<g, BLOCK, 0, h.stream>>>(p2, 0.0, nullptr, 0.0, nullptr, n);
    k_lincomb<<<grid_for(2 * E), BLOCK, 0, h.stream>>>(sum, 0.0, nullptr, 0.0, nullptr, 2 * E);
    PF_LAUNCH_CHECK();
    DevBuf xt_buf(bytes);
    h.tr_stats.recompute_newton_iterations = h.tr_stats.recompute_cg_iterations = h.tr_stats.adjoint_cg_iterations = 0;
    struct Recompute {
        Handle &h;
        ~Recompute() { h.tr_recompute = false; }
    } recompute_scope{h};
    h.tr_recompute = true;
    for (int c = ((T - 1) / K) * K; c >= 0; c -= K) {
        // recompute the states of segment [c, min(c + K, T)] from checkpoint c
        const int e = std::min(c + K, T);
        const double *ck = h.ckpt[c / K].as<double>();
        PF_HIP(hipMemcpyAsync(SX, ck, bytes, hipMemcpyDeviceToDevice, h.stream));
        PF_HIP(hipMemcpyAsync(SV, ck + n, bytes, hipMemcpyDeviceToDevice, h.stream));
        for (int i = c + 1; i <= e; ++i)
            euler_step(h, SX + size_t(i - 1 - c) * n, SV + size_t(i - 1 - c) * n, xt_buf.as<double>(),
                       SX + size_t(i - c) * n, SV + size_t(i - c) * n);
        for (int i = e; i > c; --i) {
            // rhs = dJ/dx_i + M (2 p_{i+1} - p_{i+2})
            k_lincomb<<<g, BLOCK, 0, h.stream>>>(tmp, 2.0, p1, -1.0, p2, n);
            mass_apply(h, tmp, rhs);
            if (i == T) {
                if (dJdx) k_axpy<<<g, BLOCK, 0, h.stream>>>(rhs, 1.0, dJdx, n);
                if (dJdv) k_axpy<<<g, BLOCK, 0, h.stream>>>(rhs, 1.0 / dt, dJdv, n);
            }
            if (i == T - 1 && dJdv) k_axpy<<<g, BLOCK, 0, h.stream>>>(rhs, -1.0 / dt, dJdv, n);
            PF_LAUNCH_CHECK();
            const double *xi = SX + size_t(i - c) * n;
            {
                // (M + dt^2 H(x_i)) p_i = rhs on the free dofs
                ObjectiveScope os(h, dt * dt, xi, nullptr);
                double rel = 0;
                h.tr_stats.adjoint_cg_iterations += cg_solve_ws(h, xi, fixed, rhs, cg_max, cg_tol, true, &rel);
                if (!(rel <= std::max(cg_tol, 1e-300) * 10.0) && rel > 1e-6)
                    fail(PFGPU_ERR_INVALID_STATE, "adjoint CG solve did not converge");
            }
            // shift: p_{i+2} <- p_{i+1}, p_{i+1} <- p_i
            std::swap(p1, p2);
            PF_HIP(hipMemcpyAsync(p1, cg_solution(h), bytes, hipMemcpyDeviceToDevice, h.stream));
            // material term of step i
            element_material_vjp(h, xi, p1, vjp);
            k_axpy<<<grid_for(2 * E), BLOCK, 0, h.stream>>>(sum, dt * dt, vjp, 2 * E);
            PF_LAUNCH_CHECK();
        }
    }
    // p1 = p_1, p2 = p_2
    if (dtheta) PF_HIP(hipMemcpyAsync(dtheta, sum, sizeof(double) * 2 * size_t(E), hipMemcpyDeviceToDevice, h.stream));
    if (dv0) {
        mass_apply(h, p1, dv0);
        k_lincomb<<<g, BLOCK, 0, h.stream>>>(dv0, dt, dv0, 0.0, nullptr, n);
        if (fixed) k_mask<<<g, BLOCK, 0, h.stream>>>(dv0, fixed, n);
    }
    if (dx0) {
        k_lincomb<<<g, BLOCK, 0, h.stream>>>(tmp, 1.0, p1, -1.0, p2, n);
        mass_apply(h, tmp, dx0);
        if (fixed) k_mask<<<g, BLOCK, 0, h.stream>>>(dx0, fixed, n);
    }
    PF_LAUNCH_CHECK();
    PF_HIP(hipStreamSynchronize(h.stream));
}

// ---- batches: independent problems concatenated in one handle ---------
//
// A batch handle holds several meshes of the same order and material as one
// mesh with disjoint node ranges, so every assembly kernel is one launch for
// the whole batch and the Hessian is block diagonal.  Problem s owns elements
// [eo[s], eo[s+1]) and nodes [no[s], no[s+1]).  The Newton iteration runs
// on all problems at once, and each problem keeps the iteration it would
// have alone: the CG on the block-diagonal system carries one (alpha, beta,
// r.r) per problem and stops each problem at its own relative tolerance;
// line search, fallbacks, convergence test and step length are per problem
// too.  A converged problem is frozen (its dofs join the fixed mask).
//
// Per-problem sums run over fixed chunks of at most SEG_CHUNK dofs that never
// straddle two problems: one block per chunk writes a partial, then one
// thread per problem adds its chunks' partials in order, so a problem of a
// million dofs is summed by hundreds of blocks and every sum is reproducible.

namespace {

constexpr int SEG_CHUNK = 4095; // dofs: whole nodes

// per-problem PCG state (device), one entry per problem
struct SegState {
    double *rr, *rz, *b2, *alpha, *beta, *done, *iters, *active;
};

__device__ double block_sum(double v, double *sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (t < w) sh[t] += sh[t + w];
        __syncthreads();
    }
    return sh[0];
}

// part[c] = sum over chunk c's free dofs of a b; with done, 0 for the chunks
// of finished problems (cs: chunk -> problem), whose partials nobody reads
__global__ __launch_bounds__(BLOCK) void k_chunk_dot(const double *__restrict__ a, const double *__restrict__ b,
                                                     const uint8_t *__restrict__ mask, const int32_t *__restrict__ cb,
                                                     double *__restrict__ part, const int32_t *__restrict__ cs = nullptr,
                                                     const double *__restrict__ done = nullptr)
{
    __shared__ double sh[BLOCK];
    const int c = blockIdx.x;
    if (done && done[cs[c]] != 0.0) { // uniform over the workgroup
        if (threadIdx.x == 0) part[c] = 0.0;
        return;
    }
    double acc = 0;
    for (int i = cb[c] + threadIdx.x; i < cb[c + 1]; i += BLOCK)
        if (!mask || !mask[i]) acc = fma(a[i], b[i], acc);
    const double v = block_sum(acc, sh);
    if (threadIdx.x == 0) part[c] = v;
}

// The per-problem scalar kernels run as one workgroup of SEGW threads, one
// wavefront per problem at a time: lane l adds the chunk partials l, l+64, ...
// of the problem in order, then a fixed butterfly leaves the total in every lane.
constexpr int SEGW = 1024, WAVE = 64;

__device__ double seg_total(const double *__restrict__ part, const int32_t *__restrict__ scp, int s)
{
    double v = 0;
    for (int c = scp[s] + int(threadIdx.x % WAVE); c < scp[s + 1]; c += WAVE) v += part[c];
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

// sum over the workgroup's wavefronts of a per-wavefront value, in wave order
__device__ double wave_values_sum(double v, double *sh)
{
    if (threadIdx.x % WAVE == 0) sh[threadIdx.x / WAVE] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < SEGW / WAVE; ++k) t += sh[k];
    return t;
}

__global__ __launch_bounds__(SEGW) void k_seg_total(const double *__restrict__ part, const int32_t *__restrict__ scp,
                                                    int nseg, double *__restrict__ out)
{
    for (int s = int(threadIdx.x / WAVE); s < nseg; s += SEGW / WAVE) {
        const double v = seg_total(part, scp, s);
        if (threadIdx.x % WAVE == 0) out[s] = v;
    }
}

// PCG start over chunk c's nodes: y = 0, r = b on free dofs, z = P r, p = z;
// partials of r.z and r.r
__global__ __launch_bounds__(BLOCK) void k_chunk_pcg_start(const double *__restrict__ b, const uint8_t *__restrict__ mask,
                                                           const float *__restrict__ pinv, const int32_t *__restrict__ cb,
                                                           double *__restrict__ y, double *__restrict__ r,
                                                           double *__restrict__ z, double *__restrict__ p,
                                                           double *__restrict__ part_rz, double *__restrict__ part_rr)
{
    __shared__ double sh[BLOCK];
    const int c = blockIdx.x;
    double rz = 0, rr = 0;
    for (int i = cb[c] / 3 + threadIdx.x; i < cb[c + 1] / 3; i += BLOCK) {
        const int64_t o = 3 * int64_t(i);
        double ri[3], zi[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) ri[d] = mask[o + d] ? 0.0 : b[o + d];
        precondition(pinv, i, ri, zi);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            y[o + d] = 0.0;
            r[o + d] = ri[d];
            z[o + d] = zi[d];
            p[o + d] = zi[d];
            rz = fma(ri[d], zi[d], rz);
            rr = fma(ri[d], ri[d], rr);
        }
    }
    const double a = block_sum(rz, sh), e = block_sum(rr, sh);
    if (threadIdx.x == 0) {
        part_rz[c] = a;
        part_rr[c] = e;
    }
}

// PCG start from the r.z and r.r partials: b2 = r.r; a problem with b = 0 is done
__global__ __launch_bounds__(SEGW) void k_seg_cg_start(const double *__restrict__ part_rz,
                                                       const double *__restrict__ part_rr, const int32_t *__restrict__ scp,
                                                       int nseg, SegState st)
{
    __shared__ double sh[SEGW / WAVE];
    double act = 0;
    for (int s = int(threadIdx.x / WAVE); s < nseg; s += SEGW / WAVE) {
        const double v = seg_total(part_rr, scp, s), z = seg_total(part_rz, scp, s);
        if (threadIdx.x % WAVE == 0) {
            st.rr[s] = v;
            st.rz[s] = z;
            st.b2[s] = v;
            st.done[s] = v > 0.0 ? 0.0 : 1.0;
            st.iters[s] = 0.0;
        }
        act += v > 0.0 ? 1.0 : 0.0;
    }
    const double a = wave_values_sum(act, sh);
    if (threadIdx.x == 0) *st.active = a;
}

// Per-iteration scalar kernels: one wavefront per problem over a grid of
// seg_grid(nseg) workgroups (problem s on wavefront s mod 16 of workgroup
// s / 16); the count of problems still iterating is added into *active by
// each workgroup (exact: a sum of small integers in FP64), which k_seg_alpha
// clears first.
inline unsigned seg_grid(int nseg) { return unsigned(std::max(1, (nseg + SEGW / WAVE - 1) / (SEGW / WAVE))); }

// alpha = r.z / p.Hp per problem; a non-positive curvature ends that problem
__global__ __launch_bounds__(SEGW) void k_seg_alpha(const double *__restrict__ part, const int32_t *__restrict__ scp,
                                                    int nseg, SegState st)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *st.active = 0.0;
    for (int s = int(blockIdx.x * (SEGW / WAVE) + threadIdx.x / WAVE); s < nseg; s += int(gridDim.x) * (SEGW / WAVE)) {
        const double pq = seg_total(part, scp, s);
        if (threadIdx.x % WAVE != 0) continue;
        if (st.done[s] != 0.0) {
            st.alpha[s] = 0.0;
        } else if (!(pq > 0.0)) {
            st.alpha[s] = 0.0;
            st.done[s] = 1.0;
        } else {
            st.alpha[s] = st.rz[s] / pq;
        }
    }
}

// y += alpha p, r -= alpha q on free dofs, z = P r over chunk c's nodes;
// partials of the new r.z and r.r.  The chunks of a finished problem (done)
// keep their vectors and give zero partials, which nobody reads.
__global__ __launch_bounds__(BLOCK) void k_chunk_pcg_update(double *__restrict__ y, double *__restrict__ r,
                                                            double *__restrict__ z, const double *__restrict__ p,
                                                            const double *__restrict__ q, const uint8_t *__restrict__ mask,
                                                            const float *__restrict__ pinv, const int32_t *__restrict__ cb,
                                                            const int32_t *__restrict__ cs, const double *__restrict__ alpha,
                                                            const double *__restrict__ done, double *__restrict__ part_rz,
                                                            double *__restrict__ part_rr)
{
    __shared__ double sh[BLOCK];
    const int c = blockIdx.x;
    if (done[cs[c]] != 0.0) { // uniform over the workgroup
        if (threadIdx.x == 0) part_rz[c] = part_rr[c] = 0.0;
        return;
    }
    const double a = alpha[cs[c]];
    double rz = 0, rr = 0;
    for (int i = cb[c] / 3 + threadIdx.x; i < cb[c + 1] / 3; i += BLOCK) {
        const int64_t o = 3 * int64_t(i);
        // every load of the node first, then the masked update
        double ri[3], zi[3], pi[3], qi[3], yi[3];
        bool upd[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            ri[d] = r[o + d];
            pi[d] = p[o + d];
            qi[d] = q[o + d];
            yi[d] = y[o + d];
            upd[d] = a != 0.0 && !mask[o + d];
        }
#pragma unroll
        for (int d = 0; d < 3; ++d)
            if (upd[d]) {
                y[o + d] = fma(a, pi[d], yi[d]);
                ri[d] = fma(-a, qi[d], ri[d]);
                r[o + d] = ri[d];
            }
        precondition(pinv, i, ri, zi);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            z[o + d] = zi[d];
            rz = fma(ri[d], zi[d], rz);
            rr = fma(ri[d], ri[d], rr);
        }
    }
    const double e = block_sum(rz, sh), f = block_sum(rr, sh);
    if (threadIdx.x == 0) {
        part_rz[c] = e;
        part_rr[c] = f;
    }
}

// beta = r'.z' / r.z per problem; convergence at ||r|| <= tol ||b||
__global__ __launch_bounds__(SEGW) void k_seg_beta(const double *__restrict__ part_rz, const double *__restrict__ part_rr,
                                                   const int32_t *__restrict__ scp, int nseg, double tol2, SegState st)
{
    __shared__ double sh[SEGW / WAVE];
    double act = 0;
    for (int s = int(blockIdx.x * (SEGW / WAVE) + threadIdx.x / WAVE); s < nseg; s += int(gridDim.x) * (SEGW / WAVE)) {
        const double rn = seg_total(part_rr, scp, s), zn = seg_total(part_rz, scp, s);
        if (threadIdx.x % WAVE != 0) continue;
        if (!isfinite(rn) || !isfinite(zn)) act += __builtin_nan(""); // reported to the host as a failed solve
        if (st.done[s] == 0.0) {
            st.iters[s] += 1.0;
            st.beta[s] = st.rz[s] > 0.0 ? zn / st.rz[s] : 0.0;
            st.rr[s] = rn;
            st.rz[s] = zn;
            if (rn <= tol2 * st.b2[s]) st.done[s] = 1.0;
        } else {
            st.beta[s] = 0.0;
        }
        act += st.done[s] == 0.0 ? 1.0 : 0.0;
    }
    const double a = wave_values_sum(act, sh);
    if (threadIdx.x == 0) atomicAdd(st.active, a);
}

// p = z + beta p for the problems still iterating
__global__ __launch_bounds__(BLOCK) void k_chunk_cg_direction(double *__restrict__ p, const double *__restrict__ z,
                                                              const int32_t *__restrict__ cb, const int32_t *__restrict__ cs,
                                                              const double *__restrict__ beta, const double *__restrict__ done)
{
    const int c = blockIdx.x, s = cs[c];
    if (done[s] != 0.0) return;
    const double b = beta[s];
    for (int i = cb[c] + threadIdx.x; i < cb[c + 1]; i += BLOCK) p[i] = fma(b, p[i], z[i]);
}

// out[s] = sum of v[off[s] .. off[s+1]) (one block per segment, fixed order)
__global__ __launch_bounds__(BLOCK) void k_seg_sum(const double *__restrict__ v, const int32_t *__restrict__ off,
                                                   double *__restrict__ out)
{
    __shared__ double sh[BLOCK];
    const int sgm = blockIdx.x;
    double a = 0;
    for (int i = off[sgm] + threadIdx.x; i < off[sgm + 1]; i += BLOCK) a += v[i];
    const double tot = block_sum(a, sh);
    if (threadIdx.x == 0) out[sgm] = tot;
}

// out = x + alpha[seg(i)] d over all dofs; seg: node -> segment
__global__ __launch_bounds__(BLOCK) void k_axpy_seg(double *__restrict__ out, const double *__restrict__ x,
                                                    const double *__restrict__ alpha, const double *__restrict__ d,
                                                    const int32_t *__restrict__ seg, int n)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) out[i] = fma(alpha[seg[i / 3]], d[i], x[i]);
}

// working mask: fixed dofs plus every dof of a frozen segment
__global__ __launch_bounds__(BLOCK) void k_seg_mask(uint8_t *__restrict__ mask, const uint8_t *__restrict__ fixed,
                                                    const uint8_t *__restrict__ frozen, const int32_t *__restrict__ seg, int n)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) mask[i] = (fixed && fixed[i]) || frozen[seg[i / 3]] ? 1 : 0;
}

// d = -g on the free dofs of the flagged segments (others keep d)
__global__ __launch_bounds__(BLOCK) void k_seg_steepest(double *__restrict__ d, const double *__restrict__ g,
                                                        const uint8_t *__restrict__ mask, const uint8_t *__restrict__ flag,
                                                        const int32_t *__restrict__ seg, int n)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n && flag[seg[i / 3]]) d[i] = mask[i] ? 0.0 : -g[i];
}

// chunks of at most SEG_CHUNK dofs inside each problem's dof range
struct SegChunks {
    DevBuf cb, cs, scp, part, part2;
    int n = 0;
};

SegChunks seg_chunks(const int32_t *no, int nseg, hipStream_t st)
{
    std::vector<int32_t> cb{0}, cs, scp{0};
    for (int s = 0; s < nseg; ++s) {
        const int32_t b = 3 * no[s + 1];
        for (int32_t i = 3 * no[s]; i < b; i += SEG_CHUNK) {
            cb.push_back(std::min(b, i + SEG_CHUNK));
            cs.push_back(s);
        }
        scp.push_back(int32_t(cs.size()));
    }
    SegChunks c;
    c.n = int(cs.size());
    c.cb = upload(cb, st);
    c.cs = upload(cs, st);
    c.scp = upload(scp, st);
    c.part = DevBuf{sizeof(double) * size_t(c.n)};
    c.part2 = DevBuf{sizeof(double) * size_t(c.n)};
    return c;
}

// PCG start of every active problem (a fixed kernel sequence)
void seg_cg_start(Handle &h, const Work &w, const uint8_t *mask, const SegChunks &ch, int nseg, const SegState &st)
{
    k_chunk_pcg_start<<<ch.n, BLOCK, 0, h.stream>>>(w.d, mask, precond(h), ch.cb.as<int32_t>(), w.y, w.r, w.z, w.p,
                                                    ch.part.as<double>(), ch.part2.as<double>());
    k_seg_cg_start<<<1, SEGW, 0, h.stream>>>(ch.part.as<double>(), ch.part2.as<double>(), ch.scp.as<int32_t>(), nseg, st);
    PF_LAUNCH_CHECK();
}

// One PCG iteration on every active problem at once (a fixed kernel sequence).
void seg_cg_iteration(Handle &h, const Work &w, const double *x, const uint8_t *mask, const SegChunks &ch, int nseg,
                      double tol2, const SegState &st)
{
    hvp_any(h, x, w.p, w.q);
    k_chunk_dot<<<ch.n, BLOCK, 0, h.stream>>>(w.p, w.q, mask, ch.cb.as<int32_t>(), ch.part.as<double>(),
                                              ch.cs.as<int32_t>(), st.done);
    k_seg_alpha<<<seg_grid(nseg), SEGW, 0, h.stream>>>(ch.part.as<double>(), ch.scp.as<int32_t>(), nseg, st);
    k_chunk_pcg_update<<<ch.n, BLOCK, 0, h.stream>>>(w.y, w.r, w.z, w.p, w.q, mask, precond(h), ch.cb.as<int32_t>(),
                                                     ch.cs.as<int32_t>(), st.alpha, st.done, ch.part.as<double>(),
                                                     ch.part2.as<double>());
    k_seg_beta<<<seg_grid(nseg), SEGW, 0, h.stream>>>(ch.part.as<double>(), ch.part2.as<double>(), ch.scp.as<int32_t>(),
                                                      nseg, tol2,
                                          st);
    k_chunk_cg_direction<<<ch.n, BLOCK, 0, h.stream>>>(w.p, w.z, ch.cb.as<int32_t>(), ch.cs.as<int32_t>(), st.beta,
                                                       st.done);
    PF_LAUNCH_CHECK();
}

} // namespace

void newton_solve_batch(Handle &h, double *x, const uint8_t *fixed, int nseg, const int32_t *eo_h, const int32_t *no_h,
                        const pfgpu_newton_params &prm, pfgpu_newton_report *reps)
{
    if (h.obj_xt || h.obj_f || h.obj_w != 1.0) fail(PFGPU_ERR_UNSUPPORTED, "batched solves take the elastic energy only");
    const int n = h.n_dofs, E = h.n_elements;
    const unsigned gr = grid_for(n);
    const Work w = workspace(h);
    // per-batch device tables
    std::vector<int32_t> seg_of(size_t(h.n_nodes));
    for (int sg = 0; sg < nseg; ++sg)
        for (int i = no_h[sg]; i < no_h[sg + 1]; ++i) seg_of[size_t(i)] = sg;
    DevBuf d_eo = upload(std::vector<int32_t>(eo_h, eo_h + nseg + 1), h.stream);
    DevBuf d_seg = upload(seg_of, h.stream);
    std::vector<int32_t> eseg_of(static_cast<size_t>(E));
    for (int sg = 0; sg < nseg; ++sg)
        for (int e = eo_h[sg]; e < eo_h[sg + 1]; ++e) eseg_of[size_t(e)] = sg;
    DevBuf d_eseg = upload(eseg_of, h.stream);
    const SegChunks ch = seg_chunks(no_h, nseg, h.stream);
    const size_t ns = size_t(nseg);
    // doubles: energy, gradient norm, g.d, step (4 ns), PCG state (7 ns + 1); then two byte flags per problem
    DevBuf scal{sizeof(double) * (11 * ns + 1) + 2 * ns}, pe{sizeof(double) * size_t(E)}, mask{size_t(n)};
    double *s_e = scal.as<double>(), *s_g = s_e + ns, *s_gd = s_g + ns, *s_a = s_gd + ns;
    const SegState st{s_a + ns,     s_a + 2 * ns, s_a + 3 * ns, s_a + 4 * ns,
                      s_a + 5 * ns, s_a + 6 * ns, s_a + 7 * ns, s_a + 8 * ns};
    uint8_t *frozen = reinterpret_cast<uint8_t *>(s_a + 8 * ns + 1), *flag = frozen + ns;
    std::vector<double> e0(ns), e1(ns), gn(ns), gn0(ns), gd(ns), alpha(ns), it_cg(ns);
    std::vector<uint8_t> fr(ns, 0), fl(ns, 0);
    std::vector<pfgpu_newton_report> rep(ns);
    auto read = [&](const double *dev, std::vector<double> &host) {
        PF_HIP(hipMemcpyAsync(host.data(), dev, sizeof(double) * ns, hipMemcpyDeviceToHost, h.stream));
        PF_HIP(hipStreamSynchronize(h.stream));
    };
    auto seg_dot = [&](const double *a, const double *b, const uint8_t *m, double *out) {
        k_chunk_dot<<<ch.n, BLOCK, 0, h.stream>>>(a, b, m, ch.cb.as<int32_t>(), ch.part.as<double>());
        k_seg_total<<<1, SEGW, 0, h.stream>>>(ch.part.as<double>(), ch.scp.as<int32_t>(), nseg, out);
        PF_LAUNCH_CHECK();
    };
    auto seg_energy = [&](const double *xx, std::vector<double> &host) {
        plain_energy(h, xx, w.scal + S_E1, pe.as<double>());
        k_seg_sum<<<nseg, BLOCK, 0, h.stream>>>(pe.as<double>(), d_eo.as<int32_t>(), s_e);
        PF_LAUNCH_CHECK();
        read(s_e, host);
    };
    auto seg_grad = [&]() {
        if (h.path == PFGPU_PATH_FUSED)
            fused_gradient(h, x, w.g, nullptr);
        else
            plain_gradient(h, x, w.g, nullptr);
        seg_dot(w.g, w.g, fixed, s_g);
        read(s_g, gn);
        for (double &v : gn) v = std::sqrt(v);
    };
    auto upload_flags = [&](uint8_t *dev, const std::vector<uint8_t> &host) {
        PF_HIP(hipMemcpyAsync(dev, host.data(), ns, hipMemcpyHostToDevice, h.stream));
    };
    // the CG iteration, captured once per call: every pointer it reads is fixed
    // for the call (x, the working mask and the workspace change contents only)
    hipGraphExec_t graph = nullptr;
    // the scratch and operator buffers the captured iteration uses
    const void *graph_loc = nullptr, *graph_vals = nullptr, *graph_pinv = nullptr;
    struct GraphGuard {
        hipGraphExec_t &g;
        ~GraphGuard()
        {
            if (g) (void)hipGraphExecDestroy(g);
        }
    } guard_graph{graph};
    const double tol2 = prm.cg_tolerance * prm.cg_tolerance;
    // per-problem CG on H_ff y = d over the active problems; iterations per problem into it_cg
    auto seg_cg = [&]() {
        // problems whose CG has finished (or that are frozen) skip their element work
        struct SkipScope {
            Handle &h;
            ~SkipScope() { h.skip = SegSkip{}; }
        } skip_scope{h};
        h.skip = SegSkip{d_eseg.as<int32_t>(), d_seg.as<int32_t>(), st.done};
        seg_cg_start(h, w, mask.as<uint8_t>(), ch, nseg, st);
        if (graph && (graph_loc != h.loc.get() || graph_vals != h.solver_vals.get() || graph_pinv != h.pinv.get())) {
            (void)hipGraphExecDestroy(graph);
            graph = nullptr;
        }
        if (prm.use_graph && !graph) {
            // an eager iteration sizes every scratch buffer; then restart and capture
            seg_cg_iteration(h, w, x, mask.as<uint8_t>(), ch, nseg, tol2, st);
            seg_cg_start(h, w, mask.as<uint8_t>(), ch, nseg, st);
            hipGraph_t g = nullptr;
            PF_HIP(hipStreamBeginCapture(h.stream, hipStreamCaptureModeThreadLocal));
            try {
                for (int k = 0; k < CHECK; ++k) seg_cg_iteration(h, w, x, mask.as<uint8_t>(), ch, nseg, tol2, st);
            } catch (...) {
                (void)hipStreamEndCapture(h.stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            PF_HIP(hipStreamEndCapture(h.stream, &g));
            const hipError_t e = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            PF_HIP(e);
            graph_loc = h.loc.get();
            graph_vals = h.solver_vals.get();
            graph_pinv = h.pinv.get();
        }
        int it = 0;
        std::vector<double> act(1);
        while (it < prm.cg_max_iterations) {
            // the graph holds CHECK iterations; a shorter remainder runs eagerly
            const int k = std::min(CHECK, prm.cg_max_iterations - it);
            if (graph && k == CHECK)
                PF_HIP(hipGraphLaunch(graph, h.stream));
            else
                for (int j = 0; j < k; ++j) seg_cg_iteration(h, w, x, mask.as<uint8_t>(), ch, nseg, tol2, st);
            it += k;
            PF_HIP(hipMemcpyAsync(act.data(), st.active, sizeof(double), hipMemcpyDeviceToHost, h.stream));
            PF_HIP(hipStreamSynchronize(h.stream));
            if (!std::isfinite(act[0])) fail(PFGPU_ERR_INVALID_STATE, "non-finite residual in the batched CG solve");
            if (act[0] == 0.0) break;
        }
        read(st.iters, it_cg);
    };
    // per-problem backtracking over the problems flagged in `todo`; alpha holds
    // each flagged problem's first trial step, slope its g.d
    auto search = [&](std::vector<uint8_t> &todo, const std::vector<double> &slope) {
        for (int ls = 0; ls <= prm.max_line_search; ++ls) {
            PF_HIP(hipMemcpyAsync(s_a, alpha.data(), sizeof(double) * ns, hipMemcpyHostToDevice, h.stream));
            k_axpy_seg<<<gr, BLOCK, 0, h.stream>>>(w.xt, x, s_a, w.y, d_seg.as<int32_t>(), n);
            PF_LAUNCH_CHECK();
            if (prm.line_search == 0) {
                std::fill(todo.begin(), todo.end(), 0);
                return;
            }
            seg_energy(w.xt, e1);
            bool all = true;
            for (int sg = 0; sg < nseg; ++sg) {
                if (!todo[sg]) continue;
                // Armijo, or the roundoff regime: a predicted decrease below the
                // rounding of the FP64 energy sum is taken if every element stays valid
                const double noise = roundoff_noise(h, e0[sg], e0[sg]);
                if (std::isfinite(e1[sg]) &&
                    (e1[sg] - e0[sg] <= prm.armijo_c * alpha[sg] * slope[sg] || (-slope[sg] <= noise && alpha[sg] == 1.0))) {
                    todo[sg] = 0;
                } else {
                    alpha[sg] *= 0.5;
                    rep[sg].line_search_halvings++;
                    all = false;
                }
            }
            if (all) return;
        }
    };
    for (int it = 0; it < prm.max_iterations; ++it) {
        seg_energy(x, e0);
        seg_grad();
        bool any = false;
        for (int sg = 0; sg < nseg; ++sg) {
            if (!std::isfinite(e0[sg]) || !std::isfinite(gn[sg]))
                fail(PFGPU_ERR_INVALID_STATE, "non-finite energy or gradient at a batched Newton iterate");
            rep[sg].energy = e0[sg];
            rep[sg].grad_norm = gn[sg];
            if (it == 0) {
                gn0[sg] = gn[sg];
                rep[sg].initial_grad_norm = gn[sg];
            }
            if (!fr[sg] && (gn[sg] <= prm.grad_norm || gn[sg] <= prm.relative_grad_norm * gn0[sg])) {
                fr[sg] = 1;
                rep[sg].converged = 1;
            }
            any |= !fr[sg];
        }
        if (!any) break;
        upload_flags(frozen, fr);
        k_seg_mask<<<gr, BLOCK, 0, h.stream>>>(mask.as<uint8_t>(), fixed, frozen, d_seg.as<int32_t>(), n);
        PF_LAUNCH_CHECK();
        h.op_valid = false;
        prepare_operator(h, x, mask.as<uint8_t>());
        // Newton directions of all active problems: H_ff y = -g_f, per-problem PCG
        k_neg_free<<<gr, BLOCK, 0, h.stream>>>(w.d, w.g, mask.as<uint8_t>(), n);
        PF_LAUNCH_CHECK();
        seg_cg();
        seg_dot(w.g, w.y, mask.as<uint8_t>(), s_gd);
        read(s_gd, gd);
        // a problem whose direction is not a descent direction takes -g
        bool steep = false;
        for (int sg = 0; sg < nseg; ++sg) {
            fl[sg] = !fr[sg] && !(gd[sg] < 0.0);
            if (fl[sg]) {
                gd[sg] = -gn[sg] * gn[sg];
                steep = true;
            }
            if (!fr[sg]) rep[sg].cg_iterations += int(it_cg[sg]);
        }
        if (steep) {
            upload_flags(flag, fl);
            k_seg_steepest<<<gr, BLOCK, 0, h.stream>>>(w.y, w.g, mask.as<uint8_t>(), flag, d_seg.as<int32_t>(), n);
            PF_LAUNCH_CHECK();
        }
        std::vector<uint8_t> todo(ns);
        for (int sg = 0; sg < nseg; ++sg) {
            todo[sg] = !fr[sg];
            alpha[sg] = fr[sg] ? 0.0 : 1.0;
        }
        search(todo, gd);
        // Newton direction rejected (near-singular or indefinite Hessian):
        // steepest descent from the Cauchy step g^T g / g^T H g of each problem's
        // quadratic model, or a unit-length step when its curvature along g is not positive
        bool cauchy = false;
        for (int sg = 0; sg < nseg; ++sg) {
            fl[sg] = todo[sg];
            cauchy |= todo[sg] != 0;
        }
        if (cauchy) {
            upload_flags(flag, fl);
            k_seg_steepest<<<gr, BLOCK, 0, h.stream>>>(w.y, w.g, mask.as<uint8_t>(), flag, d_seg.as<int32_t>(), n);
            hvp_any(h, x, w.y, w.q);
            seg_dot(w.y, w.q, mask.as<uint8_t>(), s_gd);
            std::vector<double> c(ns), slope(ns);
            read(s_gd, c);
            for (int sg = 0; sg < nseg; ++sg) {
                slope[sg] = -gn[sg] * gn[sg];
                if (todo[sg]) {
                    alpha[sg] = c[sg] > 0.0 ? gn[sg] * gn[sg] / c[sg] : 1.0 / gn[sg];
                    rep[sg].gradient_steps++;
                }
            }
            search(todo, slope);
            for (int sg = 0; sg < nseg; ++sg)
                if (todo[sg]) fail(PFGPU_ERR_INVALID_STATE, "batched line search found no acceptable step");
        }
        PF_HIP(hipMemcpyAsync(x, w.xt, sizeof(double) * size_t(n), hipMemcpyDeviceToDevice, h.stream));
        for (int sg = 0; sg < nseg; ++sg)
            if (!fr[sg]) {
                rep[sg].iterations = it + 1;
                rep[sg].last_alpha = alpha[sg];
            }
    }
    seg_energy(x, e0);
    seg_grad();
    for (int sg = 0; sg < nseg; ++sg) {
        rep[sg].energy = e0[sg];
        rep[sg].grad_norm = gn[sg];
        rep[sg].converged = gn[sg] <= prm.grad_norm || gn[sg] <= prm.relative_grad_norm * gn0[sg];
        reps[sg] = rep[sg];
    }
}

} // namespace pfgpu
