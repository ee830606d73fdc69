// Plain assembly path: one thread per element writes its element-local
// result to HBM, then a gather kernel sums the contributions of every output
// entry in element order.  Deterministic, no atomics.  This is the
// straightforward GPU restatement of NLAssembler's element loop
// (Assembler.cpp:495-771) and the baseline the fused path is measured against.
#include <algorithm>
#include <thread>

#include "dispatch.hpp"
#include "plans.hpp"

namespace pfgpu {

namespace {

__device__ __forceinline__ double block_sum(double v)
{
    __shared__ double s[BLOCK];
    const int t = threadIdx.x;
    s[t] = v;
    __syncthreads();
#pragma unroll
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (t < w) s[t] += s[t + w];
        __syncthreads();
    }
    return s[0];
}

// reference tables in constant memory, one per (value type, order, points)
template <typename T, int ORDER, int NQ> __constant__ RefTab<T, ORDER, NQ> c_ref = RefTab<T, ORDER, NQ>();

template <typename T, int ORDER, int NQ> __device__ __forceinline__ const RefTab<T, ORDER, NQ> &ref_tab()
{
    return c_ref<T, ORDER, NQ>;
}

// ------------------------------------------------------------ element kernels

__device__ __forceinline__ float fma_t(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return fma(a, b, c); }

// Hessian factor record of the factored route (see k_hfactors): values per element
__host__ __device__ constexpr int fac_size(int nq) { return (11 * nq + 6 + 3) & ~3; }

// the point-independent tail of a factor record, and the record's store
template <typename T, int NQ> __device__ __forceinline__ void factors_tail(const T *jit, T mu, T w, T (&r)[fac_size(NQ)])
{
    constexpr int R = fac_size(NQ);
    T g6[6];
    metric(jit, g6);
    const T cm = w * mu;
#pragma unroll
    for (int i = 0; i < 6; ++i) r[11 * NQ + i] = cm * g6[i];
#pragma unroll
    for (int i = 11 * NQ + 6; i < R; ++i) r[i] = T(0);
}

template <typename T, int ORDER, int NQ, int MAT>
__global__ __launch_bounds__(BLOCK) void k_energy(DevMesh<T> m, const double *__restrict__ x,
                                                  double *__restrict__ per_elem, double *__restrict__ partial)
{
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    double val = 0;
    if (e < m.n_elements) {
        Elem<T, ORDER, NQ> el;
        load_geometry(m, e, el);
        gather(m, e, x, el.U);
        T acc = T(0);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const T w = el.det * T(quad_weight(NQ, q));
            Kin<T> k;
            field_grad<T, ORDER, NQ>(q, el.U, el.jit, k.H);
            kinematics(k);
            acc += energy_density<T, MAT>(k, el.lam, el.mu, el.p3) * w;
        }
        val = double(acc);
        if (per_elem) per_elem[e] = val;
    }
    const double s = block_sum(val);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// explicit derivative of the element energy w.r.t. its Lame pair: the energy
// density is linear in (lambda, mu), so out[e] = sum_q w psi(F; 1, 0) and
// out[E + e] = sum_q w psi(F; 0, 1)
template <typename T, int ORDER, int NQ, int MAT>
__global__ __launch_bounds__(BLOCK) void k_energy_lame_grad(DevMesh<T> m, const double *__restrict__ x,
                                                            double *__restrict__ out)
{
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= m.n_elements) return;
    Elem<T, ORDER, NQ> el;
    load_geometry(m, e, el);
    gather(m, e, x, el.U);
    T al = T(0), am = T(0);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const T w = el.det * T(quad_weight(NQ, q));
        Kin<T> k;
        field_grad<T, ORDER, NQ>(q, el.U, el.jit, k.H);
        kinematics(k);
        al += energy_density<T, MAT>(k, T(1), T(0), T(0)) * w;
        am += energy_density<T, MAT>(k, T(0), T(1), T(0)) * w;
    }
    out[e] = double(al);
    out[m.n_elements + e] = double(am);
}

// local gradient (+ energy partials); gloc[(e*NB + a)*3 + d]
template <typename T, int ORDER, int NQ, int MAT>
__global__ __launch_bounds__(BLOCK) void k_gradient(DevMesh<T> m, const double *__restrict__ x, T *__restrict__ gloc,
                                                    double *__restrict__ partial)
{
    constexpr int NB = NLocal<ORDER>::value;
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    double val = 0;
    if (e < m.n_elements) {
        Elem<T, ORDER, NQ> el;
        load_geometry(m, e, el);
        gather(m, e, x, el.U);
        T g[NB][3];
#pragma unroll
        for (int a = 0; a < NB; ++a)
#pragma unroll
            for (int d = 0; d < 3; ++d) g[a][d] = T(0);
        T acc = T(0);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const T w = el.det * T(quad_weight(NQ, q));
            Kin<T> k;
            field_grad<T, ORDER, NQ>(q, el.U, el.jit, k.H);
            kinematics(k);
            acc += energy_density<T, MAT>(k, el.lam, el.mu, el.p3) * w;
            T P[9];
            pk1<T, MAT>(k, el.lam, el.mu, el.p3, P);
            contract_add<T, ORDER, NQ>(q, P, el.jit, w, g);
        }
        val = double(acc);
        T *out = gloc + int64_t(e) * NB * 3;
#pragma unroll
        for (int a = 0; a < NB; ++a)
#pragma unroll
            for (int d = 0; d < 3; ++d) out[a * 3 + d] = g[a][d];
    }
    if (partial) {
        const double s = block_sum(val);
        if (threadIdx.x == 0) partial[blockIdx.x] = s;
    }
}

// k_gradient plus the element's Hessian factor record (k_hfactors' record)
// from the same kinematics, with one wavefront per workgroup: the 64
// elements' local gradients and records are staged in LDS and stored as
// contiguous runs (a thread's own 120- and 208-byte results would be stored
// at a per-lane stride); energy partials per 64 elements
template <typename T, int ORDER, int NQ, int MAT>
__global__ __launch_bounds__(64) void k_grad_factors(DevMesh<T> m, const double *__restrict__ x, T *__restrict__ gloc,
                                                     double *__restrict__ partial, T *__restrict__ rec)
{
    constexpr int NB = NLocal<ORDER>::value, R = fac_size(NQ), G = 3 * NB;
    // one buffer for both stagings (records, then gradients): the workgroup's
    // LDS sets how many wavefronts a CU holds
    constexpr int S = R > G ? R : G;
    __shared__ T srec[64 * (S + 1)];
    T *sgr = srec;
    const int lane = threadIdx.x, e0 = blockIdx.x * 64, e = e0 + lane;
    double val = 0;
    T g[NB][3];
    if (e < m.n_elements) {
        Elem<T, ORDER, NQ> el;
        load_geometry(m, e, el);
        gather(m, e, x, el.U);
        // a rule that weighs every point alike (the factored path takes no other)
        const T w = el.det * T(quad_weight(NQ, 0));
#pragma unroll
        for (int a = 0; a < NB; ++a)
#pragma unroll
            for (int d = 0; d < 3; ++d) g[a][d] = T(0);
        T acc = T(0);
        T r[R];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            Kin<T> k;
            field_grad<T, ORDER, NQ>(q, el.U, el.jit, k.H);
            kinematics(k);
            acc += energy_density<T, MAT>(k, el.lam, el.mu, el.p3) * w;
            T P[9];
            pk1<T, MAT>(k, el.lam, el.mu, el.p3, P);
            contract_add<T, ORDER, NQ>(q, P, el.jit, w, g);
            hess_qp<T, MAT>(k, el.jit, el.lam, el.mu, w, r + 9 * q, r[9 * NQ + q], r[10 * NQ + q]);
        }
        factors_tail<T, NQ>(el.jit, el.mu, w, r);
        val = double(acc);
#pragma unroll
        for (int i = 0; i < R; ++i) srec[lane * (R + 1) + i] = r[i];
    }
    __syncthreads();
    const int ne = min(64, m.n_elements - e0);
    for (int i = lane; i < ne * R; i += 64) {
        const int le = i / R;
        rec[int64_t(e0) * R + i] = srec[le * (R + 1) + (i - le * R)];
    }
    __syncthreads(); // the records are read before the gradients overwrite them
    if (e < m.n_elements) {
#pragma unroll
        for (int a = 0; a < NB; ++a)
#pragma unroll
            for (int d = 0; d < 3; ++d) sgr[lane * (G + 1) + a * 3 + d] = g[a][d];
    }
    __syncthreads();
    for (int i = lane; i < ne * G; i += 64) {
        const int le = i / G;
        gloc[int64_t(e0) * G + i] = sgr[le * (G + 1) + (i - le * G)];
    }
    if (partial) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) val += __shfl_xor(val, o, 64);
        if (lane == 0) partial[blockIdx.x] = val;
    }
}

// Upper-triangle local Hessian blocks (a <= b, s(a,b) = a NB - a(a-1)/2 + b - a)
// in tiles of 64 elements: block s of element e is the 9 values at
// hloc[(((e / 64) NBLK + s) 64 + e % 64) 9], row-major.  A wavefront computes
// the blocks of its 64 elements and writes each block index s as one
// contiguous 2304-byte run, staged through LDS.
template <typename T, int ORDER, int NQ, int MAT, int HB>
__global__ __launch_bounds__(HB) void k_hessian(DevMesh<T> m, const double *__restrict__ x, T *__restrict__ hloc)
{
    constexpr int NB = NLocal<ORDER>::value, NBLK = NB * (NB + 1) / 2;
    __shared__ T stage[HB / 64][64 * 9];
    const int e = blockIdx.x * HB + threadIdx.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // every lane of a wavefront takes part in the staged stores; lanes past
    // the last element compute on element 0 and their blocks are never read
    const bool valid = e < m.n_elements;
    const int ee = valid ? e : 0;
    const int64_t tile = (int64_t(blockIdx.x) * HB + wv * 64) / 64;
    if (tile * 64 >= m.n_elements) return; // whole wavefront past the end (uniform)
    Elem<T, ORDER, NQ> el;
    load_geometry(m, ee, el);
    if (MAT != MAT_LE) gather(m, ee, x, el.U);
    // the weight of the rules that weigh every point alike (isotropic term)
    const T w0 = el.det * T(quad_weight(NQ, 0));
    // per point: M = F^{-T} J^{-T} (k_a = M RG(q,a)) and the two curvature
    // weights; the k vectors are formed per block row and column to keep the
    // register footprint small.  SaintVenant adds the S-weighted metric and
    // mu F F^T per point.
    // FixedCorotational adds the rotation R and K^{-1} per point.
    constexpr int NX = MAT == MAT_SV || MAT == MAT_FC ? NQ : 1;
    // MooneyRivlin forms its blocks from the directional stress derivative
    // (below) and keeps nothing per point
    constexpr int NM = MAT == MAT_MR ? 1 : NQ;
    T Mq[NM][9], cl[NM], cc[NM], sg6[NX][6], ff6[NX][9];
#pragma unroll
    for (int q = 0; q < NM; ++q) {
        const T w = el.det * T(quad_weight(NQ, q));
        if constexpr (MAT == MAT_MR) {
#pragma unroll
            for (int i = 0; i < 9; ++i) Mq[q][i] = T(0);
            cl[q] = cc[q] = T(0);
            continue;
        }
        Kin<T> k;
        if (MAT != MAT_LE) field_grad<T, ORDER, NQ>(q, el.U, el.jit, k.H);
        if (MAT == MAT_NH || MAT == MAT_FC) kinematics(k);
        hess_qp<T, MAT>(k, el.jit, el.lam, el.mu, w, Mq[q], cl[q], cc[q]);
        if constexpr (MAT == MAT_SV) sv_extra(k.H, el.jit, el.lam, el.mu, w, sg6[q], ff6[q]);
        if constexpr (MAT == MAT_FC) {
            // sg6: K^{-1}; ff6: the rows of R
            FcKin<T> f;
            fc_kinematics(k.H, f);
#pragma unroll
            for (int i = 0; i < 6; ++i) sg6[q][i] = f.Ki6[i];
#pragma unroll
            for (int i = 0; i < 9; ++i) ff6[q][i] = f.Q[i] + (i % 4 == 0 ? T(1) : T(0));
        }
    }
    T g6[6];
    metric(el.jit, g6);
    const T cm = (MAT == MAT_FC ? T(2) : T(1)) * w0 * el.mu;
    T *st = stage[wv];
    T *out = hloc + tile * NBLK * 576;
    const RefTab<T, ORDER, NQ> &tb = ref_tab<T, ORDER, NQ>();
    // the isotropic term sum_q w_q cm' RG(q,a)^T G RG(q,b): from the summed
    // table for the rules that weigh every point alike, point by point otherwise
    auto iso = [&](int a, int b, const T *c) {
        if constexpr (quad_uniform(NQ)) {
            return cm * (g6[0] * c[0] + g6[1] * c[1] + g6[2] * c[2] + g6[3] * c[3] + g6[4] * c[4] + g6[5] * c[5]);
        } else {
            T s = T(0);
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const T *ra = tb.rg + (q * NB + a) * 3, *rb = tb.rg + (q * NB + b) * 3;
                s += (MAT == MAT_FC ? T(2) : T(1)) * el.det * T(quad_weight(NQ, q)) * el.mu * sym_form(ra, g6, rb);
            }
            return s;
        }
    };
    // a and b are uniform over the wavefront: the table reads are scalar
#pragma unroll 1
    for (int a = 0; a < NB; ++a) {
        T Ka[NM][3];
#pragma unroll
        for (int q = 0; q < NM; ++q)
#pragma unroll
            for (int mm = 0; mm < 3; ++mm) {
                const T *r = tb.rg + (q * NB + a) * 3;
                Ka[q][mm] = Mq[q][mm * 3] * r[0] + Mq[q][mm * 3 + 1] * r[1] + Mq[q][mm * 3 + 2] * r[2];
            }
        // not unrolled: the k_b of all columns at once would not fit in registers
#pragma unroll 1
        for (int b = a; b < NB; ++b) {
            T Kb[NM][3];
#pragma unroll
            for (int q = 0; q < NM; ++q)
#pragma unroll
                for (int mm = 0; mm < 3; ++mm) {
                    const T *r = tb.rg + (q * NB + b) * 3;
                    Kb[q][mm] = Mq[q][mm * 3] * r[0] + Mq[q][mm * 3 + 1] * r[1] + Mq[q][mm * 3 + 2] * r[2];
                }
            const int s = a * NB - a * (a - 1) / 2 + (b - a);
            const T *c = tb.rgg + (a * NB + b) * 6;
            T gd = T(0), xf[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
            T xr[9] = {T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
            if constexpr (MAT == MAT_FC) {
                // -2 mu da sum_q (G_a x r_m)^T K^{-1} (G_b x r_n), r_m the rows of R
                gd = iso(a, b, c);
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    const T w = el.det * T(quad_weight(NQ, q));
                    const T *ra = tb.rg + (q * NB + a) * 3, *rb = tb.rg + (q * NB + b) * 3;
                    T Ga[3], Gb[3];
#pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        Ga[i] = ra[0] * el.jit[0 * 3 + i] + ra[1] * el.jit[1 * 3 + i] + ra[2] * el.jit[2 * 3 + i];
                        Gb[i] = rb[0] * el.jit[0 * 3 + i] + rb[1] * el.jit[1 * 3 + i] + rb[2] * el.jit[2 * 3 + i];
                    }
                    const T *R = ff6[q], *Ki = sg6[q];
                    T ca[3][3], kb[3][3];
#pragma unroll
                    for (int mm = 0; mm < 3; ++mm) {
                        const T *r = R + mm * 3;
                        ca[mm][0] = Ga[1] * r[2] - Ga[2] * r[1];
                        ca[mm][1] = Ga[2] * r[0] - Ga[0] * r[2];
                        ca[mm][2] = Ga[0] * r[1] - Ga[1] * r[0];
                        const T cb0 = Gb[1] * r[2] - Gb[2] * r[1], cb1 = Gb[2] * r[0] - Gb[0] * r[2],
                                cb2 = Gb[0] * r[1] - Gb[1] * r[0];
                        kb[mm][0] = Ki[0] * cb0 + Ki[3] * cb1 + Ki[4] * cb2;
                        kb[mm][1] = Ki[3] * cb0 + Ki[1] * cb1 + Ki[5] * cb2;
                        kb[mm][2] = Ki[4] * cb0 + Ki[5] * cb1 + Ki[2] * cb2;
                    }
                    const T cr = -T(2) * el.mu * w;
#pragma unroll
                    for (int mm = 0; mm < 3; ++mm)
#pragma unroll
                        for (int nn = 0; nn < 3; ++nn)
                            xr[mm * 3 + nn] += cr * (ca[mm][0] * kb[nn][0] + ca[mm][1] * kb[nn][1] + ca[mm][2] * kb[nn][2]);
                }
            } else if constexpr (MAT == MAT_MR) {
                // column n of the block: sum_q da dP[e_n G_b^T] G_a (the
                // kinematics of each point are formed again per block)
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    const T w = el.det * T(quad_weight(NQ, q));
                    Kin<T> k;
                    field_grad<T, ORDER, NQ>(q, el.U, el.jit, k.H);
                    kinematics(k);
                    const T *ra = tb.rg + (q * NB + a) * 3, *rb = tb.rg + (q * NB + b) * 3;
                    T Ga[3], Gb[3];
#pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        Ga[i] = ra[0] * el.jit[0 * 3 + i] + ra[1] * el.jit[1 * 3 + i] + ra[2] * el.jit[2 * 3 + i];
                        Gb[i] = rb[0] * el.jit[0 * 3 + i] + rb[1] * el.jit[1 * 3 + i] + rb[2] * el.jit[2 * 3 + i];
                    }
#pragma unroll
                    for (int nn = 0; nn < 3; ++nn) {
                        T W[9], dP[9];
#pragma unroll
                        for (int i = 0; i < 9; ++i) W[i] = i / 3 == nn ? Gb[i % 3] : T(0);
                        mr_dstress(k, el.lam, el.mu, el.p3, W, dP);
#pragma unroll
                        for (int mm = 0; mm < 3; ++mm)
                            xr[mm * 3 + nn] += w * (dP[mm * 3] * Ga[0] + dP[mm * 3 + 1] * Ga[1] + dP[mm * 3 + 2] * Ga[2]);
                    }
                }
            } else if constexpr (MAT == MAT_SV) {
                // sum_q [ (G_a^T S G_b) I + mu (G_a . G_b) F F^T ] da
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    const T *ra = tb.rg + (q * NB + a) * 3, *rb = tb.rg + (q * NB + b) * 3;
                    gd += sym_form(ra, sg6[q], rb);
                    const T gq = sym_form(ra, g6, rb);
#pragma unroll
                    for (int i = 0; i < 6; ++i) xf[i] += gq * ff6[q][i];
                }
            } else {
                gd = iso(a, b, c);
            }
            const int sym[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
#pragma unroll
            for (int mm = 0; mm < 3; ++mm)
#pragma unroll
                for (int nn = 0; nn < 3; ++nn) {
                    T v = mm == nn ? gd : T(0);
                    if constexpr (MAT == MAT_SV) v += xf[sym[mm * 3 + nn]];
                    if constexpr (MAT == MAT_FC || MAT == MAT_MR) v += xr[mm * 3 + nn];
#pragma unroll
                    for (int q = 0; q < NM; ++q) v += cl[q] * Ka[q][mm] * Kb[q][nn] + cc[q] * Kb[q][mm] * Ka[q][nn];
                    st[lane * 9 + mm * 3 + nn] = v;
                }
            // the wavefront's 64 blocks s are 576 consecutive values
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int j = 0; j < 9; ++j) out[s * 576 + j * 64 + lane] = st[j * 64 + lane];
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// local H(x) v; yloc[(e*NB + a)*3 + d].  One wavefront per workgroup: the
// 64 elements' results are staged in LDS and stored as one contiguous run
// (elements of finished problems compute and store nothing).  FP32 with the
// NeoHookean, linear and SaintVenant tangents fits 168 VGPRs without spills:
// three wavefronts per SIMD instead of two.
template <typename T, int MAT> constexpr int hvp_waves()
{
    return sizeof(T) == 4 && (MAT == MAT_NH || MAT == MAT_LE || MAT == MAT_SV) ? 3 : 1;
}
template <typename T, int ORDER, int NQ, int MAT>
__global__ __launch_bounds__(64, (hvp_waves<T, MAT>())) void k_hvp(DevMesh<T> m, const double *__restrict__ x,
                                                                   const double *__restrict__ v, T *__restrict__ yloc,
                                                                   SegSkip skip)
{
    constexpr int NB = NLocal<ORDER>::value, G = 3 * NB;
    __shared__ T sy[64 * (G + 1)];
    const int lane = threadIdx.x, e0 = blockIdx.x * 64, e = e0 + lane;
    if (e < m.n_elements && !skip.element(e)) {
        Elem<T, ORDER, NQ> el;
        load_geometry(m, e, el);
        gather(m, e, x, el.U);
        T V[NB][3];
        gather(m, e, v, V);
        T y[NB][3];
#pragma unroll
        for (int a = 0; a < NB; ++a)
#pragma unroll
            for (int d = 0; d < 3; ++d) y[a][d] = T(0);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const T w = el.det * T(quad_weight(NQ, q));
            Kin<T> k;
            field_grad<T, ORDER, NQ>(q, el.U, el.jit, k.H);
            if (MAT == MAT_NH || MAT == MAT_FC || MAT == MAT_MR) kinematics(k);
            T Gv[9], A[9];
            field_grad<T, ORDER, NQ>(q, V, el.jit, Gv);
            dstress<T, MAT>(k, el.lam, el.mu, el.p3, Gv, A);
            contract_add<T, ORDER, NQ>(q, A, el.jit, w, y);
        }
#pragma unroll
        for (int a = 0; a < NB; ++a)
#pragma unroll
            for (int d = 0; d < 3; ++d) sy[lane * (G + 1) + a * 3 + d] = y[a][d];
    }
    __syncthreads();
    const int ne = min(64, m.n_elements - e0);
    for (int i = lane; i < ne * G; i += 64) {
        const int le = i / G;
        if (!skip.element(e0 + le)) yloc[int64_t(e0) * G + i] = sy[le * (G + 1) + (i - le * G)];
    }
}

// local H(x) v from the element's Hessian factor record (k_hfactors) instead
// of the kinematics of x: with D_q = sum_b v_b RG(q,b)^T (the reference
// gradient of v) the block form of k_gather_factored gives
//   y_a = sum_q Z_q RG(q,a),  Z_q = cc_q M_q D_q^T M_q + cl_q (M_q : D_q) M_q + D_q G,
// G the record's metric tail.  Valid while the records belong to x (inside a
// CG solve, where x is fixed); about half the arithmetic of k_hvp and no
// gather of x.  Layout and staging as k_hvp.
// PARTS = 2 (rules of an even number of points): two wavefronts per 64
// elements, each taking half of the points and reading only their part of
// the record; the halves are added in LDS (first + second), so twice the
// wavefronts share the kernel's loads and latency.
template <typename T, int ORDER, int NQ, int PARTS>
__global__ __launch_bounds__(64 * PARTS) void k_hvp_factored(DevMesh<T> m, const T *__restrict__ rec,
                                                             const double *__restrict__ v, T *__restrict__ yloc,
                                                             SegSkip skip)
{
    constexpr int NB = NLocal<ORDER>::value, G = 3 * NB, R = fac_size(NQ), QP = NQ / PARTS;
    static_assert(NQ % PARTS == 0, "the parts take equal numbers of points");
    __shared__ T sy[64 * (G + 1)];
    const int lane = threadIdx.x % 64, w = threadIdx.x / 64, e0 = blockIdx.x * 64, e = e0 + lane;
    const bool on = e < m.n_elements && !skip.element(e);
    T y[NB][3];
#pragma unroll
    for (int a = 0; a < NB; ++a) y[a][0] = y[a][1] = y[a][2] = T(0);
    if (on) {
        T V[NB][3];
        gather(m, e, v, V);
        const T *rp = static_cast<const T *>(__builtin_assume_aligned(rec, 16)) + int64_t(e) * R;
        const T *G6 = rp + 11 * NQ; // 00 11 22 01 02 12
        const T Gm[9] = {G6[0], G6[3], G6[4], G6[3], G6[1], G6[5], G6[4], G6[5], G6[2]};
        // the points q = W QP .. (W + 1) QP - 1 (W uniform over the wavefront)
        auto points = [&]<int W>() {
#pragma unroll
            for (int qq = 0; qq < QP; ++qq) {
                const int q = W * QP + qq;
                T M[9];
#pragma unroll
                for (int i = 0; i < 9; ++i) M[i] = rp[9 * q + i];
                T D[9];
#pragma unroll
                for (int i = 0; i < 9; ++i) D[i] = T(0);
#pragma unroll
                for (int b = 0; b < NB; ++b)
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const double rg = RG<ORDER, NQ>(q, b, k);
                        if (rg != 0.0)
#pragma unroll
                            for (int n = 0; n < 3; ++n) D[n * 3 + k] = fma_t(V[b][n], T(rg), D[n * 3 + k]);
                    }
                T s1 = T(0);
#pragma unroll
                for (int i = 0; i < 9; ++i) s1 = fma_t(M[i], D[i], s1);
                const T cl = rp[9 * NQ + q], cc = rp[10 * NQ + q], ls = cl * s1;
                // W = cc M D^T
                T Wm[9];
#pragma unroll
                for (int mm = 0; mm < 3; ++mm)
#pragma unroll
                    for (int nn = 0; nn < 3; ++nn)
                        Wm[mm * 3 + nn] =
                            cc * (M[mm * 3] * D[nn * 3] + M[mm * 3 + 1] * D[nn * 3 + 1] + M[mm * 3 + 2] * D[nn * 3 + 2]);
                T Z[9];
#pragma unroll
                for (int mm = 0; mm < 3; ++mm)
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        T z = ls * M[mm * 3 + k];
#pragma unroll
                        for (int j = 0; j < 3; ++j)
                            z = fma_t(Wm[mm * 3 + j], M[j * 3 + k], fma_t(D[mm * 3 + j], Gm[j * 3 + k], z));
                        Z[mm * 3 + k] = z;
                    }
#pragma unroll
                for (int a = 0; a < NB; ++a)
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const double rg = RG<ORDER, NQ>(q, a, k);
                        if (rg != 0.0)
#pragma unroll
                            for (int mm = 0; mm < 3; ++mm) y[a][mm] = fma_t(Z[mm * 3 + k], T(rg), y[a][mm]);
                    }
            }
        };
        // the arithmetic at a raised wave priority, the loads and stores of
        // other wavefronts around it (36.5 against 37.3 us at 40^3 P2)
        __builtin_amdgcn_s_setprio(1);
        if constexpr (PARTS == 1) {
            points.template operator()<0>();
        } else {
            if (w == 0)
                points.template operator()<0>();
            else
                points.template operator()<1>();
        }
        __builtin_amdgcn_s_setprio(0);
    }
    if constexpr (PARTS == 2) {
        if (w == 1 && on)
#pragma unroll
            for (int a = 0; a < NB; ++a)
#pragma unroll
                for (int d = 0; d < 3; ++d) sy[lane * (G + 1) + a * 3 + d] = y[a][d];
        __syncthreads();
        if (w == 0 && on)
#pragma unroll
            for (int a = 0; a < NB; ++a)
#pragma unroll
                for (int d = 0; d < 3; ++d) y[a][d] += sy[lane * (G + 1) + a * 3 + d];
        __syncthreads(); // the second half is read before the sums overwrite it
    }
    if (w == 0 && on)
#pragma unroll
        for (int a = 0; a < NB; ++a)
#pragma unroll
            for (int d = 0; d < 3; ++d) sy[lane * (G + 1) + a * 3 + d] = y[a][d];
    __syncthreads();
    const int ne = min(64, m.n_elements - e0);
    for (int i = int(threadIdx.x); i < ne * G; i += 64 * PARTS) {
        const int le = i / G;
        if (!skip.element(e0 + le)) yloc[int64_t(e0) * G + i] = sy[le * (G + 1) + (i - le * G)];
    }
}

// (d grad/d lambda) dl + (d grad/d mu) dm, element-local
template <typename T, int ORDER, int NQ, int MAT>
__global__ __launch_bounds__(BLOCK) void k_material_jvp(DevMesh<T> m, const double *__restrict__ x,
                                                        const double *__restrict__ dl, const double *__restrict__ dm,
                                                        T *__restrict__ yloc)
{
    constexpr int NB = NLocal<ORDER>::value;
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= m.n_elements) return;
    Elem<T, ORDER, NQ> el;
    load_geometry(m, e, el);
    gather(m, e, x, el.U);
    const T sl = T(dl[e]), sm = T(dm[e]);
    T y[NB][3];
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) y[a][d] = T(0);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const T w = el.det * T(quad_weight(NQ, q));
        Kin<T> k;
        field_grad<T, ORDER, NQ>(q, el.U, el.jit, k.H);
        T A[9];
        if (MAT == MAT_NH) {
            kinematics(k);
            // dP/dmu = F - F^{-T}, dP/dlambda = ln J F^{-T} (NeoHookeanElasticity.cpp:747-762)
#pragma unroll
            for (int i = 0; i < 9; ++i) A[i] = sm * (k.H[i] + k.D[i]) + sl * k.lnJ * k.FmT[i];
        } else if (MAT == MAT_FC) {
            // dP/dmu = 2 (F - R), dP/dlambda = (J - 1) cof F (FixedCorotational.cpp:493-546)
            kinematics(k);
            FcKin<T> f;
            fc_kinematics(k.H, f);
#pragma unroll
            for (int i = 0; i < 9; ++i) A[i] = sm * T(2) * (k.H[i] - f.Q[i]) + sl * k.delta * k.J * k.FmT[i];
        } else if (MAT == MAT_SV) {
            // dP/dmu = 2 F E, dP/dlambda = tr(E) F
            SvKin<T> s;
            sv_kinematics(k.H, s);
            T FE[9];
            mul3(s.F, s.E, FE);
#pragma unroll
            for (int i = 0; i < 9; ++i) A[i] = sm * T(2) * FE[i] + sl * s.trE * s.F[i];
        } else {
            // d sigma/d mu = grad u + grad u^T, d sigma/d lambda = tr(grad u) I (LinearElasticity.cpp:290-299)
            const T tr = k.H[0] + k.H[4] + k.H[8];
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) A[r * 3 + c] = sm * (k.H[r * 3 + c] + k.H[c * 3 + r]) + (r == c ? sl * tr : T(0));
        }
        contract_add<T, ORDER, NQ>(q, A, el.jit, w, y);
    }
    T *out = yloc + int64_t(e) * NB * 3;
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) out[a * 3 + d] = y[a][d];
}

// force_material_derivative term (ElasticForceDerivative.cpp:97-129): out[e], out[E+e]
template <typename T, int ORDER, int NQ, int MAT>
__global__ __launch_bounds__(BLOCK) void k_material_vjp(DevMesh<T> m, const double *__restrict__ x,
                                                        const double *__restrict__ p, double *__restrict__ out)
{
    constexpr int NB = NLocal<ORDER>::value;
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= m.n_elements) return;
    Elem<T, ORDER, NQ> el;
    load_geometry(m, e, el);
    gather(m, e, x, el.U);
    T Pn[NB][3];
    gather(m, e, p, Pn);
    T tl = T(0), tm = T(0);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const T w = el.det * T(quad_weight(NQ, q));
        Kin<T> k;
        field_grad<T, ORDER, NQ>(q, el.U, el.jit, k.H);
        T Gp[9];
        field_grad<T, ORDER, NQ>(q, Pn, el.jit, Gp);
        T il = T(0), im = T(0);
        if (MAT == MAT_NH) {
            kinematics(k);
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                im += (k.H[i] + k.D[i]) * Gp[i];
                il += k.FmT[i] * Gp[i];
            }
            il *= k.lnJ;
        } else if (MAT == MAT_FC) {
            kinematics(k);
            FcKin<T> f;
            fc_kinematics(k.H, f);
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                im += T(2) * (k.H[i] - f.Q[i]) * Gp[i];
                il += k.FmT[i] * Gp[i];
            }
            il *= k.delta * k.J;
        } else if (MAT == MAT_SV) {
            SvKin<T> s;
            sv_kinematics(k.H, s);
            T FE[9];
            mul3(s.F, s.E, FE);
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                im += T(2) * FE[i] * Gp[i];
                il += s.F[i] * Gp[i];
            }
            il *= s.trE;
        } else {
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) im += (k.H[r * 3 + c] + k.H[c * 3 + r]) * Gp[r * 3 + c];
            il = (k.H[0] + k.H[4] + k.H[8]) * (Gp[0] + Gp[4] + Gp[8]);
        }
        tl -= il * w;
        tm -= im * w;
    }
    out[e] = double(tl);
    out[m.n_elements + e] = double(tm);
}

// force_shape_derivative term per geometric vertex (ElasticForceDerivative.cpp:208-259):
// -da W g_v with W = grad u^T (C : grad p) + grad p^T P - <P, grad p> I;
// sloc[(e*4 + v)*3 + d]
template <typename T, int ORDER, int NQ, int MAT>
__global__ __launch_bounds__(BLOCK) void k_shape_vjp(DevMesh<T> m, const double *__restrict__ x,
                                                     const double *__restrict__ p, T *__restrict__ sloc)
{
    constexpr int NB = NLocal<ORDER>::value;
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= m.n_elements) return;
    Elem<T, ORDER, NQ> el;
    load_geometry(m, e, el);
    gather(m, e, x, el.U);
    T Pn[NB][3];
    gather(m, e, p, Pn);
    T s[4][3];
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int d = 0; d < 3; ++d) s[v][d] = T(0);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const T w = el.det * T(quad_weight(NQ, q));
        Kin<T> k;
        field_grad<T, ORDER, NQ>(q, el.U, el.jit, k.H);
        if (MAT == MAT_NH || MAT == MAT_FC || MAT == MAT_MR) kinematics(k);
        T Gp[9], P[9], B[9];
        field_grad<T, ORDER, NQ>(q, Pn, el.jit, Gp);
        pk1<T, MAT>(k, el.lam, el.mu, el.p3, P);
        dstress<T, MAT>(k, el.lam, el.mu, el.p3, Gp, B);
        T sp = T(0);
#pragma unroll
        for (int i = 0; i < 9; ++i) sp += P[i] * Gp[i];
        T W[9];
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                T acc = d == c ? -sp : T(0);
#pragma unroll
                for (int r = 0; r < 3; ++r) acc += k.H[r * 3 + d] * B[r * 3 + c] + Gp[r * 3 + d] * P[r * 3 + c];
                W[d * 3 + c] = acc;
            }
        // the P1 geometric basis at the same point
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            T gv[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                T g = T(0);
#pragma unroll
                for (int kk = 0; kk < 3; ++kk) {
                    const double r = RG<1, NQ>(q, v, kk);
                    if (r != 0.0) g += T(r) * el.jit[kk * 3 + c];
                }
                gv[c] = g;
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) s[v][d] -= w * (W[d * 3 + 0] * gv[0] + W[d * 3 + 1] * gv[1] + W[d * 3 + 2] * gv[2]);
        }
    }
    T *out = sloc + int64_t(e) * 12;
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int d = 0; d < 3; ++d) out[v * 3 + d] = s[v][d];
}

// -------------------------------------------------------------- gathers

// out[3i+d] = sum_{k in ptr[i]..ptr[i+1]} loc[inc[k]*3 + d], accumulated in double
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_gather_rows(int n, const int32_t *__restrict__ ptr, const int32_t *__restrict__ inc,
                                                       const T *__restrict__ loc, double *__restrict__ out)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    double s0 = 0, s1 = 0, s2 = 0;
    for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
        const T *l = loc + int64_t(inc[k]) * 3;
        s0 += double(l[0]);
        s1 += double(l[1]);
        s2 += double(l[2]);
    }
    out[3 * int64_t(i)] = s0;
    out[3 * int64_t(i) + 1] = s1;
    out[3 * int64_t(i) + 2] = s2;
}

// Segmented gathers.  The output items (node rows, or 3x3 blocks of the
// pattern) are cut into chunks of consecutive items holding at most SEG
// contributions (chunk[c] = first item of chunk c).  A workgroup takes one
// chunk: thread t fetches contribution t of the chunk into LDS, so all of the
// chunk's loads are in flight at once; then one thread per item sums its
// contributions in their stored (element) order and writes the item.
constexpr int SEG = BLOCK;

// Workgroups are dealt round-robin to the 8 XCDs (each with its own L2).
// Chunk c of n is given to workgroup (c mod per) * 8 + c / per, per =
// ceil(n / 8), so that each XCD walks one contiguous eighth of the chunks and
// its L2 sees the element blocks that neighbouring rows share.  The grid is
// 8 per workgroups; those past the last chunk exit.
constexpr int N_XCD = 8;
__device__ __forceinline__ int xcd_chunk(int n_chunks)
{
    const int per = (n_chunks + N_XCD - 1) / N_XCD;
    return int(blockIdx.x % N_XCD) * per + int(blockIdx.x / N_XCD);
}
inline unsigned xcd_grid(int n_chunks) { return unsigned((n_chunks + N_XCD - 1) / N_XCD * N_XCD); }

// Node-row gather of element-local 3-vectors over a row plan: chunk c covers
// the rows chunk[c] .. chunk[c+1]-1, whose at most SEG contributions (local
// vector indices) sit in the slots c SEG .. (c+1) SEG - 1 (-1 when unused);
// rmeta[i] = q0 | count << 10 gives row i's contributions within its chunk.
// Both dependent load chains (slot, then the local vector; chunk start,
// then the row metadata) are two deep and issued before the barrier.
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_gather_rows_seg(int n_chunks, const int32_t *__restrict__ chunk,
                                                           const int32_t *__restrict__ slot,
                                                           const int32_t *__restrict__ rmeta, const T *__restrict__ loc,
                                                           double *__restrict__ out, SegSkip skip)
{
    __shared__ double sv[SEG * 3];
    const int c = xcd_chunk(n_chunks);
    if (c >= n_chunks) return;
    const int t = threadIdx.x;
    const int32_t w = slot[int64_t(c) * SEG + t];
    const int i0 = chunk[c], i1 = chunk[c + 1];
    // rows of skipped problems are written as zero; a chunk lies in at most a
    // few problems, so a fully skipped chunk reads nothing
    if (skip.node(i0) && skip.node(i1 - 1)) {
        for (int i = i0 + t; i < i1; i += BLOCK)
            if (skip.node(i)) out[3 * int64_t(i)] = out[3 * int64_t(i) + 1] = out[3 * int64_t(i) + 2] = 0.0;
        bool all = true;
        for (int i = i0; i < i1 && all; ++i) all = skip.node(i);
        if (all) return;
    }
    T l0 = T(0), l1 = T(0), l2 = T(0);
    if (w >= 0) {
        const T *l = loc + int64_t(w) * 3;
        l0 = l[0];
        l1 = l[1];
        l2 = l[2];
    }
    const int i = i0 + t;
    const int32_t mt = i < i1 ? rmeta[i] : 0;
    if (w >= 0) {
        sv[t * 3] = double(l0);
        sv[t * 3 + 1] = double(l1);
        sv[t * 3 + 2] = double(l2);
    }
    __syncthreads();
    auto row = [&](int ii, int32_t m) {
        if (skip.node(ii)) {
            out[3 * int64_t(ii)] = out[3 * int64_t(ii) + 1] = out[3 * int64_t(ii) + 2] = 0.0;
            return;
        }
        const int q0 = m & 1023, q1 = q0 + (m >> 10);
        double s0 = 0, s1 = 0, s2 = 0;
        for (int k = q0; k < q1; ++k) {
            s0 += sv[k * 3];
            s1 += sv[k * 3 + 1];
            s2 += sv[k * 3 + 2];
        }
        out[3 * int64_t(ii)] = s0;
        out[3 * int64_t(ii) + 1] = s1;
        out[3 * int64_t(ii) + 2] = s2;
    };
    if (i < i1) row(i, mt);
    // rows without contributions (nodes no element uses)
    for (int ii = i + BLOCK; ii < i1; ii += BLOCK) row(ii, rmeta[ii]);
}

// Gather plan of a workgroup size GS (built on the host, see build_plan):
// chunk c covers the pattern blocks chunk[c] .. chunk[c+1]-1 and holds at
// most GS contributions, stored in the slots c GS .. (c+1) GS - 1 (-1 when
// unused); out[b] = 9 r0 + 3 (b - r0) is block b's first value (r0 the first
// block of its row) and meta[b] = q0 | cnt << 10 | deg << 20 gives its
// contributions' first slot in the chunk, their count and the blocks of its
// row.  Every load the workgroup needs is at most two levels deep (the
// slot's contribution, then its data; the chunk's first block, then the
// block's out and meta), and both chains are issued before the barrier.

// block b: the sum of its staged contributions, into the CSR values
template <typename T>
__device__ __forceinline__ void sum_block(const T *sv, int32_t off, int32_t mt, T *__restrict__ values)
{
    const int q0 = mt & 1023, cnt = (mt >> 10) & 1023, deg = int(uint32_t(mt) >> 20);
    T acc[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) acc[i] = T(0);
    for (int k = q0; k < q0 + cnt; ++k)
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[i] += sv[k * 9 + i];
    T *out = values + off;
#pragma unroll
    for (int mm = 0; mm < 3; ++mm)
#pragma unroll
        for (int nn = 0; nn < 3; ++nn) out[mm * 3 * deg + nn] = acc[mm * 3 + nn];
}

// slot word: index of the contribution's block in hloc, bit 31 when the
// block is stored transposed (the index is below 2^31 - 1, so -1 is free)
template <typename T, int GS>
__global__ __launch_bounds__(GS) void k_gather_blocks_seg(int n_chunks, const int32_t *__restrict__ chunk,
                                                             const int32_t *__restrict__ slot,
                                                             const int32_t *__restrict__ bout,
                                                             const int32_t *__restrict__ meta,
                                                             const T *__restrict__ hloc, T *__restrict__ values)
{
    __shared__ T sv[GS * 9];
    const int c = xcd_chunk(n_chunks);
    if (c >= n_chunks) return;
    const int t = threadIdx.x;
    const int32_t cs = slot[int64_t(c) * GS + t];
    const int b0 = chunk[c], b1 = chunk[c + 1];
    T h[9];
    if (cs != -1) {
        const T *p = hloc + int64_t(cs & 0x7fffffff) * 9;
#pragma unroll
        for (int i = 0; i < 9; ++i) h[i] = p[i];
    }
    const int bq = b0 + t;
    int32_t off = 0, mt = 0;
    if (bq < b1) {
        off = bout[bq];
        mt = meta[bq];
    }
    if (cs >= 0) {
#pragma unroll
        for (int i = 0; i < 9; ++i) sv[t * 9 + i] = h[i];
    } else if (cs != -1) {
#pragma unroll
        for (int mm = 0; mm < 3; ++mm)
#pragma unroll
            for (int nn = 0; nn < 3; ++nn) sv[t * 9 + mm * 3 + nn] = h[nn * 3 + mm];
    }
    __syncthreads();
    if (bq < b1) sum_block(sv, off, mt, values);
    // blocks without contributions (none in a pattern built from elements)
    for (int bb = bq + GS; bb < b1; bb += GS) sum_block(sv, bout[bb], meta[bb], values);
}

// ------------------------------------------------------------ factored Hessian
//
// With k_a = M_q RG(q, a), M_q = F^{-T} J^{-T} (J^{-T} for linear
// elasticity), the NeoHookean / linear-elasticity block is
//   H_e(a, b) = (cm sum_q G_a . G_b) I + sum_q [cl_q k_a k_b^T + cc_q k_b k_a^T],
// the form k_hessian evaluates.  k_hfactors writes per element the record
//   [M_q (9 per point) | cl_q | cc_q | cm g6 (the reference metric, 6)],
// padded to a multiple of 4 values, and k_gather_factored forms each
// contribution's block from its element's record where k_gather_blocks_seg
// reads a stored block: 11 NQ + 6 values per element travel through HBM
// instead of the 9 n(n+1)/2 of the element blocks, and the lanes of a
// wavefront that share an element read its record through the same cache
// lines.  (Records holding every node's k vectors instead take less
// arithmetic in the gather but cost more in their strided writes.)
template <typename T, int ORDER, int NQ, int MAT>
__global__ __launch_bounds__(64) void k_hfactors(DevMesh<T> m, const double *__restrict__ x, T *__restrict__ rec)
{
    // one wavefront per workgroup: the 64 records are staged in LDS and
    // stored as one contiguous run
    constexpr int R = fac_size(NQ);
    __shared__ T srec[64 * (R + 1)];
    const int lane = threadIdx.x, e0 = blockIdx.x * 64, e = e0 + lane;
    if (e < m.n_elements) {
        Elem<T, ORDER, NQ> el;
        load_geometry(m, e, el);
        if (MAT != MAT_LE) gather(m, e, x, el.U);
        // a rule that weighs every point alike (the factored path takes no other)
        const T w = el.det * T(quad_weight(NQ, 0));
        T r[R];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            Kin<T> k;
            if (MAT != MAT_LE) field_grad<T, ORDER, NQ>(q, el.U, el.jit, k.H);
            if (MAT == MAT_NH) kinematics(k);
            hess_qp<T, MAT>(k, el.jit, el.lam, el.mu, w, r + 9 * q, r[9 * NQ + q], r[10 * NQ + q]);
        }
        factors_tail<T, NQ>(el.jit, el.mu, w, r);
#pragma unroll
        for (int i = 0; i < R; ++i) srec[lane * (R + 1) + i] = r[i];
    }
    __syncthreads();
    const int ne = min(64, m.n_elements - e0);
    for (int i = lane; i < ne * R; i += 64) {
        const int le = i / R;
        rec[int64_t(e0) * R + i] = srec[le * (R + 1) + (i - le * R)];
    }
}

// gather plan as for k_gather_blocks_seg, except that a chunk's slots hold
// its contributions ordered by element (the lanes of a wavefront then share
// records, and so cache lines, in runs) and dst[slot] is the contribution's
// position in block order, where its block is staged; slot word
// e NB^2 + a NB + b.
//
// Persistent: the workgroups of XCD x (blockIdx mod 8) walk that XCD's
// contiguous eighth of the chunks with stride gridDim / 8, and the slot
// words and block metadata of a workgroup's next chunk are loaded while it
// processes the current one, so only the record loads wait in each chunk
// (requesting the next chunk's records before the current chunk's sums was
// measured slower).  The grid is the device's resident workgroups; gridDim
// must be a multiple of 8.
template <typename T, int ORDER, int NQ, int GS>
__global__ __launch_bounds__(GS) void k_gather_factored(int n_chunks, const int32_t *__restrict__ chunk,
                                                        const int32_t *__restrict__ slot,
                                                        const uint16_t *__restrict__ dst,
                                                        const int32_t *__restrict__ bout,
                                                        const int32_t *__restrict__ meta, const T *__restrict__ rec,
                                                        T *__restrict__ values)
{
    constexpr int NB = NLocal<ORDER>::value, R = fac_size(NQ);
    __shared__ T sv[GS * 9];
    __shared__ T srg[NQ * NB * 3], srgg[NB * NB * 6];
    const int per = (n_chunks + N_XCD - 1) / N_XCD, xcd = int(blockIdx.x % N_XCD);
    const int stride = int(gridDim.x / N_XCD), c_end = min(n_chunks, (xcd + 1) * per);
    int c = xcd * per + int(blockIdx.x / N_XCD);
    if (c >= c_end) return; // uniform over the workgroup
    const int t = threadIdx.x;
    const RefTab<T, ORDER, NQ> &tb = ref_tab<T, ORDER, NQ>();
    for (int i = t; i < NQ * NB * 3; i += GS) srg[i] = tb.rg[i];
    for (int i = t; i < NB * NB * 6; i += GS) srgg[i] = tb.rgg[i];
    // indices of chunk cc: this thread's slot word and stage position, the
    // chunk's blocks, this thread's output block's offset and metadata
    auto fetch = [&](int cc, int32_t &cd, int &ds, int &b0, int &b1, int32_t &off, int32_t &mt) {
        cd = slot[int64_t(cc) * GS + t];
        ds = dst[int64_t(cc) * GS + t];
        b0 = chunk[cc];
        b1 = chunk[cc + 1];
        off = 0;
        mt = 0;
        if (b0 + t < b1) {
            off = bout[b0 + t];
            mt = meta[b0 + t];
        }
    };
    int32_t cd, off, mt;
    int ds, b0, b1;
    fetch(c, cd, ds, b0, b1, off, mt);
    while (c < c_end) { // c, c_end and the chunk bounds are uniform over the workgroup
        const int cn = c + stride;
        int32_t cdn = -1, offn = 0, mtn = 0;
        int dsn = 0, b0n = 0, b1n = 0;
        if (cn < c_end) fetch(cn, cdn, dsn, b0n, b1n, offn, mtn);
        // this thread's contribution: its element's record
        int a = 0, b = 0;
        T r[R];
        if (cd >= 0) {
            const int e = cd / (NB * NB), ab = cd - e * (NB * NB);
            a = ab / NB;
            b = ab - a * NB;
            const T *rp = static_cast<const T *>(__builtin_assume_aligned(rec, 16)) + int64_t(e) * R;
#pragma unroll
            for (int i = 0; i < R; ++i) r[i] = rp[i];
        }
        // the tables are staged, and the previous chunk's sums have read sv
        __syncthreads();
        if (cd >= 0) {
            const T *cr = srgg + (a * NB + b) * 6;
            const T gd = r[11 * NQ] * cr[0] + r[11 * NQ + 1] * cr[1] + r[11 * NQ + 2] * cr[2] + r[11 * NQ + 3] * cr[3] +
                         r[11 * NQ + 4] * cr[4] + r[11 * NQ + 5] * cr[5];
            T v[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) v[i] = i % 4 == 0 ? gd : T(0);
            // written as fused multiply-adds: 42 per point
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const T *M = r + 9 * q, *ra = srg + (q * NB + a) * 3, *rb = srg + (q * NB + b) * 3;
                T ka[3], kb[3];
#pragma unroll
                for (int mm = 0; mm < 3; ++mm) {
                    ka[mm] = fma_t(M[mm * 3 + 2], ra[2], fma_t(M[mm * 3 + 1], ra[1], M[mm * 3] * ra[0]));
                    kb[mm] = fma_t(M[mm * 3 + 2], rb[2], fma_t(M[mm * 3 + 1], rb[1], M[mm * 3] * rb[0]));
                }
                const T cl = r[9 * NQ + q], cc = r[10 * NQ + q];
                const T la[3] = {cl * ka[0], cl * ka[1], cl * ka[2]}, lb[3] = {cc * kb[0], cc * kb[1], cc * kb[2]};
#pragma unroll
                for (int mm = 0; mm < 3; ++mm)
#pragma unroll
                    for (int nn = 0; nn < 3; ++nn)
                        v[mm * 3 + nn] = fma_t(lb[mm], ka[nn], fma_t(la[mm], kb[nn], v[mm * 3 + nn]));
            }
#pragma unroll
            for (int i = 0; i < 9; ++i) sv[ds * 9 + i] = v[i];
        }
        __syncthreads();
        if (b0 + t < b1) sum_block(sv, off, mt, values);
        // blocks without contributions (none in a pattern built from elements)
        for (int bb = b0 + t + GS; bb < b1; bb += GS) sum_block(sv, bout[bb], meta[bb], values);
        c = cn;
        cd = cdn;
        ds = dsn;
        b0 = b0n;
        b1 = b1n;
        off = offn;
        mt = mtn;
    }
}

// Row-unit gather of the factored route.  A unit is one node incidence
// (row node i = local node a of element e); the units of a node row are its
// incident elements in element order (node_inc).  A chunk is a run of whole
// rows holding at most GS units and GS - 1 rows; thread t of the workgroup
// takes unit t of the chunk: it reads the element's factor record once,
// forms the n_local blocks H_e(a, b) of its row, b = 0 .. n_local - 1 (b is
// unrolled, so the reference gradients RG(q, b) are compile-time constants
// and their zeros cost nothing), and stages each block at its position udst
// (the blocks of the chunk in pattern order, the contributions of a block in
// unit order).  The block sums are then split into pieces of at most
// UNIT_PIECE contributions (plans.hpp); thread t sums its run of pieces,
// lanes[c (GS + 1) + t] .. lanes[c (GS + 1) + t + 1] (balanced by the plan),
// storing a block that is one piece straight into the CSR values (3 runs of
// 3 values per block, consecutive blocks of a row contiguous) and the sum of
// a piece of a split block over the piece's first staged contribution; after
// a barrier the thread that summed the first piece of a split block adds the
// piece sums in piece order and stores the block.  One thread per (row,
// component) sums the local gradient rows of the row's units in FP64 when
// GRAD.  Summation orders are fixed by the tables and do not depend on GS:
// results are bitwise reproducible.
//
//   H_e(a, b) = (sum_q RG(q,a)^T G RG(q,b)) I + sum_q [cl_q k_a k_b^T + cc_q k_b k_a^T],
//   k_a = M_q RG(q, a),  G = w mu J^{-1} J^{-T} (the record's metric tail)
//
// The staged blocks (GS n_local 9 values) bound the workgroups per CU to a
// few, so the kernel is persistent and software-pipelined: the workgroups of
// XCD x (blockIdx mod 8) walk that XCD's contiguous eighth of the chunks with
// stride gridDim / 8; while chunk k is formed and summed, the records, stage
// positions, local gradients, lane entries and row / block tables of chunk
// k + 1 are in flight, and the unit codes of chunk k + 2 (on which the record
// addresses depend).  Every load of the pipeline is issued unconditionally
// (indices clamped into their arrays, results of lanes past the chunk
// unused), so the compiler can wait for exactly the loads of the chunk at
// hand.  ctab[c] = (first row, first unit, first pattern block) of chunk c,
// ctab[n_chunks] the totals; n_words = 32-bit words of bq0.  gridDim must be
// a multiple of 8.
template <typename T, int ORDER, int NQ, int GS, bool GRAD, int SPLIT, bool PREFETCH>
__global__ __launch_bounds__(GS) void k_gather_units(int n_chunks, int n_words, const int4 *__restrict__ ctab,
                                                     const int32_t *__restrict__ node_ptr,
                                                     const int32_t *__restrict__ node_inc,
                                                     const int32_t *__restrict__ brow_ptr,
                                                     const uint16_t *__restrict__ udst, const uint16_t *__restrict__ bq0,
                                                     const uint32_t *__restrict__ lanes, const T *__restrict__ rec,
                                                     const T *__restrict__ gloc, T *__restrict__ values,
                                                     double *__restrict__ grad)
{
    // UPC units per chunk; with SPLIT = 2 each unit has two threads in
    // different wavefronts, each forming half of its blocks
    constexpr int NB = NLocal<ORDER>::value, R = fac_size(NQ), DW = NB / 2, L = UNIT_PIECE, UPC = GS / SPLIT;
    static_assert(SPLIT == 1 || (SPLIT == 2 && UPC % 64 == 0), "the halves of a split unit lie in different wavefronts");
    // 32-bit words of block starts per thread: a chunk has at most UPC NB blocks
    constexpr int NWMAX = UPC * NB / 2 + 2, QW = (NWMAX + GS - 1) / GS;
    static_assert(NB % 2 == 0, "stage positions are read as 32-bit pairs");
    __shared__ T srg[NQ * NB * 3];
    __shared__ int32_t srb[UPC + 1], sru[UPC + 1];
    __shared__ uint32_t sq[NWMAX];
    extern __shared__ __align__(16) unsigned char smem[];
    T *sv = reinterpret_cast<T *>(smem); // UPC NB staged blocks of 9 values
    T *sg = sv + UPC * NB * 9;           // UPC local gradient rows
    const int per = (n_chunks + N_XCD - 1) / N_XCD, xcd = int(blockIdx.x % N_XCD);
    const int stride = int(gridDim.x / N_XCD), c_end = min(n_chunks, (xcd + 1) * per);
    int c = xcd * per + int(blockIdx.x / N_XCD);
    if (c >= c_end) return; // uniform over the workgroup
    const int t = threadIdx.x, last_unit = ctab[n_chunks].y - 1;
    const int ut = t % UPC, hp = t / UPC; // the thread's unit and half (uniform per wavefront)
    const RefTab<T, ORDER, NQ> &tb = ref_tab<T, ORDER, NQ>();
    for (int k = t; k < NQ * NB * 3; k += GS) srg[k] = tb.rg[k];
    __syncthreads();

    struct Meta {
        int i0, nr, u0, nu, B0, nb;
    };
    // chunk cc's metadata; past the XCD's last chunk, that chunk's again
    auto meta = [&](int cc) {
        cc = min(cc, c_end - 1);
        const int4 a = ctab[cc], b = ctab[cc + 1];
        return Meta{a.x, b.x - a.x, a.y, b.y - a.y, a.z, b.z - a.z};
    };
    // the unit code (e n_local + a) of thread t, some other unit's past the chunk
    auto code_of = [&](const Meta &m) { return node_inc[min(m.u0 + ut, last_unit)]; };
    // everything else a chunk needs, in registers
    struct Unit {
        T r[R];
        uint32_t dw[DW];
        T l0, l1, l2;
        int32_t rb0, ru;
        uint32_t qw[QW];
        uint32_t lane0, lane1;
    };
    auto load_unit = [&](int cc, const Meta &m, int code, Unit &u) {
        const int e = code / NB;
        const T *rp = static_cast<const T *>(__builtin_assume_aligned(rec, 16)) + int64_t(e) * R;
#pragma unroll
        for (int i = 0; i < R; ++i) u.r[i] = rp[i];
        const uint32_t *dp = reinterpret_cast<const uint32_t *>(udst + int64_t(min(m.u0 + ut, last_unit)) * NB);
#pragma unroll
        for (int j = 0; j < DW; ++j) u.dw[j] = dp[j];
        if constexpr (GRAD) {
            const T *gp = gloc + int64_t(code) * 3;
            u.l0 = gp[0];
            u.l1 = gp[1];
            u.l2 = gp[2];
        }
        const int rowi = m.i0 + min(t, m.nr);
        u.rb0 = brow_ptr[rowi];
        u.ru = node_ptr[rowi];
        // block starts: bq0[B0 .. B0 + nb - 1] as 32-bit words from an even index
        const uint32_t *qp = reinterpret_cast<const uint32_t *>(bq0);
        const int w0 = m.B0 >> 1;
#pragma unroll
        for (int j = 0; j < QW; ++j) u.qw[j] = qp[min(w0 + t + j * GS, n_words - 1)];
        const uint32_t *lp = lanes + int64_t(min(cc, c_end - 1)) * (GS + 1) + t;
        u.lane0 = lp[0];
        u.lane1 = lp[1];
    };

    Meta m0 = meta(c);
    int code0 = code_of(m0);
    Unit u0;
    if constexpr (PREFETCH) load_unit(c, m0, code0, u0);
    int c1 = c + stride;
    Meta m1 = meta(c1);
    int code1 = code_of(m1);
    while (true) { // c, c1 and the chunk metadata are uniform over the workgroup
        // in flight while chunk c is processed: chunk c1's data (PREFETCH;
        // otherwise chunk c's data is loaded here, the registers of a second
        // chunk left to more resident wavefronts), chunk c2's codes
        Unit u1;
        if constexpr (PREFETCH)
            load_unit(c1, m1, code1, u1);
        else
            load_unit(c, m0, code0, u0);
        const int c2 = c1 + stride;
        const Meta m2 = meta(c2);
        const int code2 = code_of(m2);
        // ---- chunk c: blocks of this thread's unit into the stage.  The
        // formation (VALU-bound) issues at a raised wave priority and the rest
        // of the chunk (LDS-bound sums, tables, gradient) at the normal one:
        // the workgroups of a CU drift into different phases, so one's
        // formation overlaps another's sums (the gather ran about 4 % faster
        // with either phase raised than with none)
        __builtin_amdgcn_s_setprio(1);
        if (ut < m0.nu) {
            const int a = code0 - (code0 / NB) * NB;
            const T *r = u0.r;
            if constexpr (GRAD) {
                if (hp == 0) {
                    sg[ut * 3] = u0.l0;
                    sg[ut * 3 + 1] = u0.l1;
                    sg[ut * 3 + 2] = u0.l2;
                }
            }
            const T *G6 = r + 11 * NQ; // 00 11 22 01 02 12
            // per quadrature point: cl k_a, cc k_a and G RG(q, a)
            T la[NQ][3], lc[NQ][3], ha[NQ][3];
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const T *M = r + 9 * q, *ra = srg + (q * NB + a) * 3;
                const T ra0 = ra[0], ra1 = ra[1], ra2 = ra[2];
#pragma unroll
                for (int mm = 0; mm < 3; ++mm) {
                    const T ka = fma_t(M[mm * 3 + 2], ra2, fma_t(M[mm * 3 + 1], ra1, M[mm * 3] * ra0));
                    la[q][mm] = r[9 * NQ + q] * ka;
                    lc[q][mm] = r[10 * NQ + q] * ka;
                }
                ha[q][0] = fma_t(G6[4], ra2, fma_t(G6[3], ra1, G6[0] * ra0));
                ha[q][1] = fma_t(G6[5], ra2, fma_t(G6[1], ra1, G6[3] * ra0));
                ha[q][2] = fma_t(G6[2], ra2, fma_t(G6[5], ra1, G6[4] * ra0));
            }
            // the blocks b = B0 .. B1 - 1 of this unit's row (b compile-time)
            auto form = [&]<int B0, int B1>() {
#pragma unroll
            for (int b = B0; b < B1; ++b) {
                T gd = T(0);
#pragma unroll
                for (int q = 0; q < NQ; ++q)
#pragma unroll
                    for (int l = 0; l < 3; ++l) {
                        const double rb = RG<ORDER, NQ>(q, b, l);
                        if (rb != 0.0) gd = fma_t(ha[q][l], T(rb), gd);
                    }
                T v[9];
#pragma unroll
                for (int i = 0; i < 9; ++i) v[i] = i % 4 == 0 ? gd : T(0);
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    const T *M = r + 9 * q;
                    T kb[3];
#pragma unroll
                    for (int mm = 0; mm < 3; ++mm) {
                        T s = T(0);
                        bool first = true;
#pragma unroll
                        for (int l = 0; l < 3; ++l) {
                            const double rb = RG<ORDER, NQ>(q, b, l);
                            if (rb != 0.0) {
                                s = first ? M[mm * 3 + l] * T(rb) : fma_t(M[mm * 3 + l], T(rb), s);
                                first = false;
                            }
                        }
                        kb[mm] = s;
                    }
                    // cl k_a k_b^T + cc k_b k_a^T
#pragma unroll
                    for (int mm = 0; mm < 3; ++mm)
#pragma unroll
                        for (int nn = 0; nn < 3; ++nn)
                            v[mm * 3 + nn] = fma_t(kb[mm], lc[q][nn], fma_t(la[q][mm], kb[nn], v[mm * 3 + nn]));
                }
                const int d = int((u0.dw[b >> 1] >> (16 * (b & 1))) & 0xffffu);
                T *o = sv + d * 9;
#pragma unroll
                for (int i = 0; i < 9; ++i) o[i] = v[i];
            }
            };
            if constexpr (SPLIT == 1) {
                form.template operator()<0, NB>();
            } else {
                if (hp == 0)
                    form.template operator()<0, NB / 2>();
                else
                    form.template operator()<NB / 2, NB>();
            }
        }
        __builtin_amdgcn_s_setprio(0);
        // ---- chunk c: row and block tables
        if (t <= m0.nr) {
            srb[t] = u0.rb0 - m0.B0;
            sru[t] = u0.ru - m0.u0;
        }
        {
            const int nw = (m0.nb + (m0.B0 & 1) + 1) / 2;
#pragma unroll
            for (int j = 0; j < QW; ++j)
                if (t + j * GS < nw) sq[t + j * GS] = u0.qw[j];
        }
        __syncthreads();
        // ---- chunk c: this thread's run of pieces
        const uint16_t *sq16 = reinterpret_cast<const uint16_t *>(sq) + (m0.B0 & 1);
        const int total = m0.nu * NB;
        auto block_end = [&](int k) { return k + 1 < m0.nb ? int(sq16[k + 1]) : total; };
        // block k of the row whose blocks are rs .. re - 1 (chunk-local)
        // (32-bit value indices: the plan is built for fewer than 2^31 values)
        auto write_block = [&](int k, int rs, int re, const T(&acc)[9]) {
            const int deg = re - rs, o = 9 * (m0.B0 + rs) + 3 * (k - rs);
#pragma unroll
            for (int mm = 0; mm < 3; ++mm)
#pragma unroll
                for (int nn = 0; nn < 3; ++nn) values[o + mm * 3 * deg + nn] = acc[mm * 3 + nn];
        };
        // first pieces of split blocks: block, its row's block range
        int nh = 0, hk0 = 0, hs0 = 0, he0 = 0, hk1 = 0, hs1 = 0, he1 = 0;
        {
            int k = int(u0.lane0 & 0xfffu), j = int((u0.lane0 >> 12) & 0xffu), rr = int(u0.lane0 >> 20);
            const int ke = int(u0.lane1 & 0xfffu), je = int((u0.lane1 >> 12) & 0xffu);
            if ((k != ke || j != je) && k < m0.nb) {
                int rs = srb[rr], re = srb[rr + 1];
                int q0 = sq16[k], q1 = block_end(k), qn = block_end(k + 1);
                while (true) {
                    const int s = q0 + j * L, pe = min(q1, s + L);
                    // the piece's first contribution (a piece is never empty:
                    // s < pe), then two in flight, in unit order
                    T acc[9];
#pragma unroll
                    for (int i = 0; i < 9; ++i) acc[i] = sv[s * 9 + i];
                    int x = s + 1;
                    for (; x + 2 <= pe; x += 2) {
                        T c0[9], c1[9];
#pragma unroll
                        for (int i = 0; i < 9; ++i) {
                            c0[i] = sv[x * 9 + i];
                            c1[i] = sv[x * 9 + 9 + i];
                        }
#pragma unroll
                        for (int i = 0; i < 9; ++i) acc[i] = (acc[i] + c0[i]) + c1[i];
                    }
                    if (x < pe)
#pragma unroll
                        for (int i = 0; i < 9; ++i) acc[i] += sv[x * 9 + i];
                    const bool whole = pe == q1;
                    if (j == 0 && whole) {
                        write_block(k, rs, re, acc);
                    } else {
                        // a piece of a split block: its sum over its first contribution
#pragma unroll
                        for (int i = 0; i < 9; ++i) sv[s * 9 + i] = acc[i];
                        if (j == 0) {
                            if (nh == 0) {
                                hk0 = k;
                                hs0 = rs;
                                he0 = re;
                            } else {
                                hk1 = k;
                                hs1 = rs;
                                he1 = re;
                            }
                            ++nh;
                        }
                    }
                    if (whole) {
                        ++k;
                        j = 0;
                    } else {
                        ++j;
                    }
                    if ((k == ke && j == je) || k >= m0.nb) break;
                    if (j == 0) {
                        while (k >= re) {
                            rs = re;
                            re = srb[++rr + 1];
                        }
                        q0 = q1;
                        q1 = qn;
                        qn = block_end(k + 1);
                    }
                }
            }
        }
        // ---- chunk c: one thread per (row, component): the gradient, in FP64
        // (unit order; four staged rows in flight)
        if constexpr (GRAD) {
            for (int it = t; it < 3 * m0.nr; it += GS) {
                const int row = it / 3, comp = it - 3 * row, ue = sru[row + 1];
                int u = sru[row];
                double s = 0;
                for (; u + 4 <= ue; u += 4) {
                    const T g0 = sg[u * 3 + comp], g1 = sg[u * 3 + 3 + comp], g2 = sg[u * 3 + 6 + comp],
                            g3 = sg[u * 3 + 9 + comp];
                    s += double(g0);
                    s += double(g1);
                    s += double(g2);
                    s += double(g3);
                }
                for (; u < ue; ++u) s += double(sg[u * 3 + comp]);
                grad[3 * int64_t(m0.i0) + it] = s;
            }
        }
        __syncthreads();
        // ---- chunk c: split blocks, their piece sums in piece order
        for (int hh = 0; hh < nh; ++hh) {
            const int k = hh ? hk1 : hk0, rs = hh ? hs1 : hs0, re = hh ? he1 : he0;
            const int q0 = sq16[k], q1 = block_end(k);
            T acc[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) acc[i] = sv[q0 * 9 + i];
            for (int s = q0 + L; s < q1; s += L)
#pragma unroll
                for (int i = 0; i < 9; ++i) acc[i] += sv[s * 9 + i];
            write_block(k, rs, re, acc);
        }
        if (c1 >= c_end) break;
        __syncthreads(); // the stage and tables are read before chunk c1 overwrites them
        c = c1;
        m0 = m1;
        if constexpr (PREFETCH) u0 = u1;
        code0 = code1;
        c1 = c2;
        m1 = m2;
        code1 = code2;
    }
}

// Diagonal blocks of the Hessian per node from the factor records, over the
// row plan of k_gather_rows_seg (slot word e NB + a): H_e(a, a) =
// (cm sum_q G_a . G_a) I + sum_q (cl_q + cc_q) k_a k_a^T, summed over the
// node's (element, local node) pairs in incidence order into diag[9 i ..]
// (row-major 3x3).  The block-Jacobi preconditioner of a matrix-free
// operator needs nothing else of the Hessian.
template <typename T, int ORDER, int NQ>
__global__ __launch_bounds__(BLOCK) void k_node_diag(int n_chunks, const int32_t *__restrict__ chunk,
                                                     const int32_t *__restrict__ slot, const int32_t *__restrict__ rmeta,
                                                     const T *__restrict__ rec, T *__restrict__ diag)
{
    constexpr int NB = NLocal<ORDER>::value, R = fac_size(NQ);
    __shared__ T sv[SEG * 6];
    __shared__ T srg[NQ * NB * 3], srgg[NB * NB * 6];
    const int c = xcd_chunk(n_chunks);
    if (c >= n_chunks) return;
    const int t = threadIdx.x;
    const int32_t w = slot[int64_t(c) * SEG + t];
    const int i0 = chunk[c], i1 = chunk[c + 1];
    int a = 0;
    T r[R];
    if (w >= 0) {
        const int e = w / NB;
        a = w - e * NB;
        const T *rp = static_cast<const T *>(__builtin_assume_aligned(rec, 16)) + int64_t(e) * R;
#pragma unroll
        for (int k = 0; k < R; ++k) r[k] = rp[k];
    }
    const int i = i0 + t;
    const int32_t mt = i < i1 ? rmeta[i] : 0;
    const RefTab<T, ORDER, NQ> &tb = ref_tab<T, ORDER, NQ>();
    for (int k = t; k < NQ * NB * 3; k += BLOCK) srg[k] = tb.rg[k];
    for (int k = t; k < NB * NB * 6; k += BLOCK) srgg[k] = tb.rgg[k];
    __syncthreads();
    if (w >= 0) {
        const T *cr = srgg + (a * NB + a) * 6;
        const T gd = r[11 * NQ] * cr[0] + r[11 * NQ + 1] * cr[1] + r[11 * NQ + 2] * cr[2] + r[11 * NQ + 3] * cr[3] +
                     r[11 * NQ + 4] * cr[4] + r[11 * NQ + 5] * cr[5];
        T v6[6] = {gd, gd, gd, T(0), T(0), T(0)}; // 00 11 22 01 02 12
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const T *M = r + 9 * q, *ra = srg + (q * NB + a) * 3;
            T ka[3];
#pragma unroll
            for (int mm = 0; mm < 3; ++mm) ka[mm] = fma_t(M[mm * 3 + 2], ra[2], fma_t(M[mm * 3 + 1], ra[1], M[mm * 3] * ra[0]));
            const T s = r[9 * NQ + q] + r[10 * NQ + q];
            v6[0] = fma_t(s * ka[0], ka[0], v6[0]);
            v6[1] = fma_t(s * ka[1], ka[1], v6[1]);
            v6[2] = fma_t(s * ka[2], ka[2], v6[2]);
            v6[3] = fma_t(s * ka[0], ka[1], v6[3]);
            v6[4] = fma_t(s * ka[0], ka[2], v6[4]);
            v6[5] = fma_t(s * ka[1], ka[2], v6[5]);
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) sv[t * 6 + k] = v6[k];
    }
    __syncthreads();
    auto row = [&](int ii, int32_t m) {
        const int q0 = m & 1023, q1 = q0 + (m >> 10);
        T acc[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
        for (int k = q0; k < q1; ++k)
#pragma unroll
            for (int j = 0; j < 6; ++j) acc[j] += sv[k * 6 + j];
        T *o = diag + 9 * int64_t(ii);
        o[0] = acc[0], o[1] = acc[3], o[2] = acc[4];
        o[3] = acc[3], o[4] = acc[1], o[5] = acc[5];
        o[6] = acc[4], o[7] = acc[5], o[8] = acc[2];
    };
    if (i < i1) row(i, mt);
    for (int ii = i + BLOCK; ii < i1; ii += BLOCK) row(ii, rmeta[ii]);
}

} // namespace

// ------------------------------------------------------------------ launchers

// node-row gather of element-local 3-vectors (defined below)
void gather_rows(Handle &h, const void *loc, double *out);

void energy_into(Handle &h, const double *x, double *energy, double *partial, double *part)
{
    const unsigned g = grid_for(h.n_elements);
    dispatch(h, [&]<typename T, int O, int Q, int M>() {
        k_energy<T, O, Q, M><<<g, BLOCK, 0, h.stream>>>(h.mesh<T>(), x, nullptr, partial);
    });
    PF_LAUNCH_CHECK();
    sum_doubles_into(h, partial, int(g), energy, part);
}

void plain_energy(Handle &h, const double *x, double *energy, double *per_element)
{
    if (!h.groups.empty()) {
        mixed_energy(h, x, energy, per_element);
        return;
    }
    const unsigned g = grid_for(h.n_elements);
    h.partial.reserve(sizeof(double) * g);
    dispatch(h, [&]<typename T, int O, int Q, int M>() {
        k_energy<T, O, Q, M><<<g, BLOCK, 0, h.stream>>>(h.mesh<T>(), x, per_element, h.partial.as<double>());
    });
    PF_LAUNCH_CHECK();
    if (energy) sum_doubles(h, h.partial.as<double>(), int(g), energy);
}

void plain_gradient(Handle &h, const double *x, double *g, double *energy)
{
    if (!h.groups.empty()) {
        mixed_gradient(h, x, g, energy);
        return;
    }
    const unsigned grid = grid_for(h.n_elements);
    h.loc.reserve(h.value_size() * size_t(h.n_elements) * h.n_local * 3);
    if (energy) h.partial.reserve(sizeof(double) * grid);
    dispatch(h, [&]<typename T, int O, int Q, int M>() {
        k_gradient<T, O, Q, M><<<grid, BLOCK, 0, h.stream>>>(h.mesh<T>(), x, h.loc.as<T>(),
                                                            energy ? h.partial.as<double>() : nullptr);
        PF_LAUNCH_CHECK();
        gather_rows(h, h.loc.get(), g);
    });
    PF_LAUNCH_CHECK();
    if (energy) sum_doubles(h, h.partial.as<double>(), int(grid), energy);
}

// chunk starts over items 0..n-1 with contribution offsets ptr: consecutive
// items, at most SEG contributions per chunk
// the row plan of k_gather_rows_seg, from the node incidence (make_row_plan)
void ensure_row_chunks(Handle &h)
{
    if (h.row_chunk.get()) return;
    static_assert(ROW_SEG == SEG, "the row plan's chunk capacity is the row gather's workgroup");
    RowPlanHost r;
    if (!make_row_plan(h.pat, r))
        fail(PFGPU_ERR_UNSUPPORTED, "an output gathers more contributions than a gather workgroup holds");
    h.row_slot = upload(r.slot, h.stream);
    h.row_meta = upload(r.rmeta, h.stream);
    h.row_chunk = upload(r.chunk, h.stream);
    h.n_row_chunks = int(r.chunk.size()) - 1;
}

void gather_rows(Handle &h, const void *loc, double *out)
{
    ensure_row_chunks(h);
    if (h.precision == PFGPU_MIXED)
        k_gather_rows_seg<float><<<xcd_grid(h.n_row_chunks), BLOCK, 0, h.stream>>>(
            h.n_row_chunks, h.row_chunk.as<int32_t>(), h.row_slot.as<int32_t>(), h.row_meta.as<int32_t>(),
            static_cast<const float *>(loc), out, h.skip);
    else
        k_gather_rows_seg<double><<<xcd_grid(h.n_row_chunks), BLOCK, 0, h.stream>>>(
            h.n_row_chunks, h.row_chunk.as<int32_t>(), h.row_slot.as<int32_t>(), h.row_meta.as<int32_t>(),
            static_cast<const double *>(loc), out, h.skip);
    PF_LAUNCH_CHECK();
}

// Contributions of every pattern block in element order (shared by the
// block and factored gathers): offsets, the element-local codes
// e NB^2 + a NB + b, and every block's first CSR value.
void ensure_plain_hessian_maps(Handle &h)
{
    if (h.plain_hess_ready) return;
    if (9 * h.pat.n_blocks >= (int64_t(1) << 31))
        fail(PFGPU_ERR_UNSUPPORTED, "too many Hessian values for 32-bit value offsets");
    std::vector<int64_t> ptr64;
    std::vector<int32_t> block_row;
    build_block_contributions(h.pat, h.n_local, h.h_elem_nodes.data(), ptr64, h.h_contrib_code, block_row);
    h.h_contrib_ptr.assign(ptr64.begin(), ptr64.end());
    std::vector<int32_t> bout(size_t(h.pat.n_blocks));
    for (int i = 0; i < h.n_nodes; ++i) {
        const int32_t r0 = h.pat.brow_ptr[i];
        for (int32_t b = r0; b < h.pat.brow_ptr[i + 1]; ++b) bout[size_t(b)] = 9 * r0 + 3 * (b - r0);
    }
    h.blk_out = upload(bout, h.stream);
    if (!h.brow_ptr.get()) h.brow_ptr = upload(h.pat.brow_ptr, h.stream);
    h.blk_plan = GatherPlan{};
    h.fac_plan = GatherPlan{};
    h.plain_hess_ready = true;
}

// the gather plan of workgroup size gs for slot words word(k) of the
// contributions k (layout described at k_gather_blocks_seg, built by
// make_contrib_plan); by_element: each chunk's slots reordered by key(word),
// with their block-order positions in dst (the factored gather's layout)
template <typename W, typename K> void build_plan(Handle &h, GatherPlan &p, int gs, W &&word, bool by_element, K &&key)
{
    if (p.threads == gs) return;
    const std::vector<int32_t> &ptr = h.h_contrib_ptr;
    const size_t nk = size_t(ptr[size_t(h.pat.n_blocks)]);
    std::vector<int32_t> words(nk), keys(by_element ? nk : 0);
    for (size_t k = 0; k < nk; ++k) {
        words[k] = word(int32_t(k));
        if (by_element) keys[k] = key(words[k]);
    }
    ContribPlanHost c;
    if (!make_contrib_plan(h.pat, ptr, words, by_element ? &keys : nullptr, gs, c))
        fail(PFGPU_ERR_UNSUPPORTED, "an output gathers more contributions than a gather workgroup holds, "
                                    "or a node has more than 4095 neighbours");
    if (by_element) p.dst = upload(c.dst, h.stream);
    p.chunk = upload(c.chunk, h.stream);
    p.slot = upload(c.slot, h.stream);
    p.meta = upload(c.meta, h.stream);
    p.n_chunks = int(c.chunk.size()) - 1;
    p.threads = gs;
    p.grid = 0;
}

// persistent grid of a kernel: the workgroups of gs threads resident at
// once on the device (from the kernel's occupancy), a multiple of the 8
// XCDs, and no more than the chunks
template <typename K> unsigned persistent_grid(Handle &h, K kernel, int n_chunks, int gs)
{
    if (h.n_cu == 0) PF_HIP(hipDeviceGetAttribute(&h.n_cu, hipDeviceAttributeMultiprocessorCount, h.device));
    int per_cu = 0;
    PF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, gs, 0));
    const int g = std::max(1, h.n_cu) * std::max(1, per_cu);
    return std::min(xcd_grid(n_chunks), unsigned((g + N_XCD - 1) / N_XCD * N_XCD));
}

// the factored route's plan (element-ordered slots) for the tuned gather size
GatherPlan &factored_plan(Handle &h)
{
    if (h.material != MAT_NH && h.material != MAT_LE)
        fail(PFGPU_ERR_UNSUPPORTED, "the factored Hessian covers NeoHookean and linear elasticity");
    ensure_plain_hessian_maps(h);
    const int nb2 = h.n_local * h.n_local;
    build_plan(
        h, h.fac_plan, h.tune_gather_threads, [&](int32_t k) { return h.h_contrib_code[size_t(k)]; }, true,
        [&](int32_t w) { return w / nb2; });
    return h.fac_plan;
}

// the CSR values from the factor records in h.fac_rec
template <typename T, int O, int Q> void launch_factored_gather(Handle &h, GatherPlan &p, void *values)
{
    auto gath = [&]<int GS>() {
        if (p.grid == 0) p.grid = persistent_grid(h, k_gather_factored<T, O, Q, GS>, p.n_chunks, GS);
        k_gather_factored<T, O, Q, GS><<<p.grid, GS, 0, h.stream>>>(p.n_chunks, p.chunk.as<int32_t>(), p.slot.as<int32_t>(),
                                                                    p.dst.as<uint16_t>(), h.blk_out.as<int32_t>(),
                                                                    p.meta.as<int32_t>(), h.fac_rec.as<T>(),
                                                                    static_cast<T *>(values));
    };
    if (p.threads == 64)
        gath.template operator()<64>();
    else if (p.threads == 128)
        gath.template operator()<128>();
    else if (p.threads == 512)
        gath.template operator()<512>();
    else
        gath.template operator()<256>();
    PF_LAUNCH_CHECK();
}

// the row-unit plan of workgroups of gs threads, split threads per unit
// (layout at k_gather_units, built by make_unit_plan); false when a node has
// more incident elements than one chunk holds (the per-contribution gather
// then serves the handle)
bool build_unit_plan(Handle &h, UnitPlan &p, int gs, int split)
{
    if (p.threads == gs && p.split == split) return true;
    UnitPlanHost u;
    if (!make_unit_plan(h.pat, h.n_local, h.h_elem_nodes.data(), gs / split, gs, u)) return false;
    // per chunk boundary: first row, first unit, first pattern block
    const int nc = int(u.chunk_row.size()) - 1;
    std::vector<int32_t> ctab(size_t(nc + 1) * 4, 0);
    for (int c = 0; c <= nc; ++c) {
        const int32_t i = u.chunk_row[size_t(c)];
        ctab[size_t(c) * 4] = i;
        ctab[size_t(c) * 4 + 1] = h.pat.node_ptr[size_t(i)];
        ctab[size_t(c) * 4 + 2] = h.pat.brow_ptr[size_t(i)];
    }
    p.chunk_row = upload(ctab, h.stream);
    p.udst = upload(u.udst, h.stream);
    // two spare entries: the kernel reads the block starts as 32-bit words
    // from an even index, up to one entry past a chunk's last block
    u.bq0.push_back(0);
    u.bq0.push_back(0);
    if (u.bq0.size() % 2) u.bq0.push_back(0);
    p.bq0 = upload(u.bq0, h.stream);
    p.n_words = int(u.bq0.size() / 2);
    p.lanes = upload(u.lanes, h.stream);
    p.n_chunks = nc;
    p.threads = gs;
    p.split = split;
    p.grid = p.grid_grad = 0;
    return true;
}

// the CSR values (and, with gloc, the gradient) from the factor records in
// h.fac_rec over the row-unit plan
template <typename T, int O, int Q> void launch_unit_gather(Handle &h, UnitPlan &p, const T *gloc, double *g, void *values)
{
    constexpr int NB = NLocal<O>::value;
    auto gath = [&]<int GS, bool GRAD, int SPLIT>() {
        // two threads per unit: no prefetch of the next chunk (fewer
        // registers, more resident wavefronts)
        auto kern = k_gather_units<T, O, Q, GS, GRAD, SPLIT, SPLIT == 1>;
        const int lds = int(sizeof(T)) * (GS / SPLIT) * (NB * 9 + 3);
        unsigned &grid = GRAD ? p.grid_grad : p.grid;
        if (grid == 0) {
            PF_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern), hipFuncAttributeMaxDynamicSharedMemorySize, lds));
            // persistent: the workgroups resident at once, a multiple of the 8 XCDs
            if (h.n_cu == 0) PF_HIP(hipDeviceGetAttribute(&h.n_cu, hipDeviceAttributeMultiprocessorCount, h.device));
            int per_cu = 0;
            PF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, GS, size_t(lds)));
            const int g = std::max(1, h.n_cu) * std::max(1, per_cu);
            grid = std::min(xcd_grid(p.n_chunks), unsigned((g + N_XCD - 1) / N_XCD * N_XCD));
        }
        kern<<<grid, GS, lds, h.stream>>>(p.n_chunks, p.n_words, p.chunk_row.as<int4>(), h.node_ptr.as<int32_t>(),
                                          h.node_inc.as<int32_t>(), h.brow_ptr.as<int32_t>(), p.udst.as<uint16_t>(),
                                          p.bq0.as<uint16_t>(), p.lanes.as<uint32_t>(), h.fac_rec.as<T>(), gloc,
                                          static_cast<T *>(values), g);
    };
    auto sized = [&]<bool GRAD>() {
        if (p.threads == 64)
            gath.template operator()<64, GRAD, 1>();
        else if (p.threads == 256)
            gath.template operator()<256, GRAD, 1>();
        else
            gath.template operator()<128, GRAD, 2>();
    };
    if (gloc)
        sized.template operator()<true>();
    else
        sized.template operator()<false>();
    PF_LAUNCH_CHECK();
}

// the factored route's gather for the tuned workgroup size: the row-unit
// plan, or the per-contribution plan when a node has too many incident
// elements for it (or the tuned size is 512)
bool unit_plan_ready(Handle &h)
{
    if (h.material != MAT_NH && h.material != MAT_LE)
        fail(PFGPU_ERR_UNSUPPORTED, "the factored Hessian covers NeoHookean and linear elasticity");
    ensure_plain_hessian_maps(h);
    // 64: a chunk of 64 units, one thread each; 128: 64 units, two threads
    // each (in different wavefronts, each forming half of the blocks); 256:
    // 256 units, one thread each
    int gs = h.tune_gather_threads;
    if (gs > 256 || 9 * int64_t(h.pat.n_blocks) >= (int64_t(1) << 31)) return false;
    // dynamic stage + the static tables (row tables, block starts, reference gradients)
    auto lds = [&](int g) {
        const int u = g == 128 ? 64 : g;
        return int64_t(h.value_size()) * u * (h.n_local * 9 + 3) + 8 * (u + 1) + 4 * (u * h.n_local / 2 + 2) + 1024;
    };
    // the staged blocks must fit the 160 KB of LDS (FP64 P2 at 256 threads
    // does not): the largest fitting size at or below the tuned one
    while (gs > 64 && lds(gs) > 160 * 1024) gs /= 2;
    if (lds(gs) > 160 * 1024) return false;
    return build_unit_plan(h, h.unit_plan, gs, gs == 128 ? 2 : 1);
}

void factored_hessian(Handle &h, const double *x, void *values)
{
    const bool units = unit_plan_ready(h);
    GatherPlan *p = units ? nullptr : &factored_plan(h);
    dispatch<true>(h, [&]<typename T, int O, int Q, int M>() {
        if constexpr (M == MAT_NH || M == MAT_LE) {
            h.fac_rec.reserve(sizeof(T) * size_t(h.n_elements) * fac_size(Q));
            k_hfactors<T, O, Q, M><<<grid_for(h.n_elements, 64), 64, 0, h.stream>>>(h.mesh<T>(), x, h.fac_rec.as<T>());
            PF_LAUNCH_CHECK();
            if (units)
                launch_unit_gather<T, O, Q>(h, h.unit_plan, nullptr, nullptr, values);
            else
                launch_factored_gather<T, O, Q>(h, *p, values);
        }
    });
}

// E + G + H on the factored route: one element pass writes the local
// gradient, the energy partials and the Hessian factor records
void factored_energy_gradient_hessian(Handle &h, const double *x, double *energy, double *g, void *values)
{
    const bool units = unit_plan_ready(h);
    GatherPlan *p = units ? nullptr : &factored_plan(h);
    const unsigned grid = grid_for(h.n_elements, 64);
    h.loc.reserve(h.value_size() * size_t(h.n_elements) * h.n_local * 3);
    if (energy) h.partial.reserve(sizeof(double) * grid);
    dispatch<true>(h, [&]<typename T, int O, int Q, int M>() {
        if constexpr (M == MAT_NH || M == MAT_LE) {
            h.fac_rec.reserve(sizeof(T) * size_t(h.n_elements) * fac_size(Q));
            k_grad_factors<T, O, Q, M><<<grid, 64, 0, h.stream>>>(
                h.mesh<T>(), x, h.loc.as<T>(), energy ? h.partial.as<double>() : nullptr, h.fac_rec.as<T>());
            PF_LAUNCH_CHECK();
            if (units) {
                // the gradient rows are summed by the same pass
                launch_unit_gather<T, O, Q>(h, h.unit_plan, h.loc.as<T>(), g, values);
            } else {
                gather_rows(h, h.loc.get(), g);
                launch_factored_gather<T, O, Q>(h, *p, values);
            }
            if (energy) sum_doubles(h, h.partial.as<double>(), int(grid), energy);
        }
    });
}

// the Hessian's diagonal block of every node (n_nodes x 9, value type of the
// handle) from the factor records; NeoHookean and linear elasticity
void factored_node_diagonal(Handle &h, const double *x, void *diag)
{
    if (h.material != MAT_NH && h.material != MAT_LE)
        fail(PFGPU_ERR_UNSUPPORTED, "the factored Hessian covers NeoHookean and linear elasticity");
    ensure_row_chunks(h);
    dispatch<true>(h, [&]<typename T, int O, int Q, int M>() {
        if constexpr (M == MAT_NH || M == MAT_LE) {
            h.fac_rec.reserve(sizeof(T) * size_t(h.n_elements) * fac_size(Q));
            k_hfactors<T, O, Q, M><<<grid_for(h.n_elements, 64), 64, 0, h.stream>>>(h.mesh<T>(), x, h.fac_rec.as<T>());
            PF_LAUNCH_CHECK();
            k_node_diag<T, O, Q><<<xcd_grid(h.n_row_chunks), BLOCK, 0, h.stream>>>(
                h.n_row_chunks, h.row_chunk.as<int32_t>(), h.row_slot.as<int32_t>(), h.row_meta.as<int32_t>(),
                h.fac_rec.as<T>(), static_cast<T *>(diag));
            PF_LAUNCH_CHECK();
        }
    });
}

void plain_hessian(Handle &h, const double *x, void *values, bool project_to_psd)
{
    // the element-block route carries the PSD projection
    if (h.path == PFGPU_PATH_FACTORED && !project_to_psd) {
        factored_hessian(h, x, values);
        return;
    }
    ensure_plain_hessian_maps(h);
    // contribution (e, a, b) -> index of block s(min(a,b), max(a,b)) of element
    // e in the tiled hloc layout of k_hessian; bit 31 when a > b (transposed)
    const int NB = h.n_local, NBLK = h.n_blocks_sym();
    if (int64_t((h.n_elements + 63) / 64) * NBLK * 64 >= (int64_t(1) << 31) - 1)
        fail(PFGPU_ERR_UNSUPPORTED, "too many local Hessian blocks for 31-bit indices");
    build_plan(h, h.blk_plan, h.tune_gather_threads, [&](int32_t k) {
        const int32_t c = h.h_contrib_code[size_t(k)];
        const int e = c / (NB * NB), a = (c / NB) % NB, b = c % NB;
        const int lo = std::min(a, b), hi = std::max(a, b);
        const int s = lo * NB - lo * (lo - 1) / 2 + (hi - lo);
        const int32_t idx = int32_t((int64_t(e / 64) * NBLK + s) * 64 + e % 64);
        return a > b ? int32_t(uint32_t(idx) | 0x80000000u) : idx;
    }, false, [](int32_t w) { return w; });
    const GatherPlan &p = h.blk_plan;
    h.hloc.reserve(h.value_size() * size_t((h.n_elements + 63) / 64) * 64 * h.n_blocks_sym() * 9);
    dispatch(h, [&]<typename T, int O, int Q, int M>() {
        auto hess = [&]<int HB>() {
            k_hessian<T, O, Q, M, HB><<<grid_for(h.n_elements, HB), HB, 0, h.stream>>>(h.mesh<T>(), x, h.hloc.as<T>());
        };
        if (h.tune_hessian_threads == 64)
            hess.template operator()<64>();
        else if (h.tune_hessian_threads == 128)
            hess.template operator()<128>();
        else
            hess.template operator()<256>();
        PF_LAUNCH_CHECK();
        if (project_to_psd) psd_project_blocks(h, h.hloc.get());
        auto gath = [&]<int GS>() {
            k_gather_blocks_seg<T, GS><<<xcd_grid(p.n_chunks), GS, 0, h.stream>>>(
                p.n_chunks, p.chunk.as<int32_t>(), p.slot.as<int32_t>(), h.blk_out.as<int32_t>(), p.meta.as<int32_t>(),
                h.hloc.as<T>(), static_cast<T *>(values));
        };
        if (p.threads == 64)
            gath.template operator()<64>();
        else if (p.threads == 128)
            gath.template operator()<128>();
        else if (p.threads == 512)
            gath.template operator()<512>();
        else
            gath.template operator()<256>();
    });
    PF_LAUNCH_CHECK();
}

void plain_hvp(Handle &h, const double *x, const double *v, double *y)
{
    if (!h.groups.empty()) {
        mixed_hvp(h, x, v, y);
        return;
    }
    const unsigned grid = grid_for(h.n_elements, 64);
    h.loc.reserve(h.value_size() * size_t(h.n_elements) * h.n_local * 3);
    dispatch(h, [&]<typename T, int O, int Q, int M>() {
        k_hvp<T, O, Q, M><<<grid, 64, 0, h.stream>>>(h.mesh<T>(), x, v, h.loc.as<T>(), h.skip);
        PF_LAUNCH_CHECK();
        gather_rows(h, h.loc.get(), y);
    });
    PF_LAUNCH_CHECK();
}

void factored_hvp(Handle &h, const double *v, double *y)
{
    const unsigned grid = grid_for(h.n_elements, 64);
    dispatch<true>(h, [&]<typename T, int O, int Q, int M>() {
        if constexpr (M == MAT_NH || M == MAT_LE) {
            constexpr int PARTS = Q % 2 == 0 ? 2 : 1;
            k_hvp_factored<T, O, Q, PARTS>
                <<<grid, 64 * PARTS, 0, h.stream>>>(h.mesh<T>(), h.fac_rec.as<T>(), v, h.loc.as<T>(), h.skip);
            PF_LAUNCH_CHECK();
            gather_rows(h, h.loc.get(), y);
        } else {
            fail(PFGPU_ERR_INVALID_STATE, "factor records exist for NeoHookean and linear elasticity only");
        }
    });
    PF_LAUNCH_CHECK();
}

void plain_material_jvp(Handle &h, const double *x, const double *dl, const double *dm, double *y)
{
    if (h.material == MAT_MR)
        fail(PFGPU_ERR_UNSUPPORTED, "MooneyRivlin has no Lame-parameter derivatives (as in the reference)");
    const unsigned grid = grid_for(h.n_elements);
    h.loc.reserve(h.value_size() * size_t(h.n_elements) * h.n_local * 3);
    dispatch(h, [&]<typename T, int O, int Q, int M>() {
        k_material_jvp<T, O, Q, M><<<grid, BLOCK, 0, h.stream>>>(h.mesh<T>(), x, dl, dm, h.loc.as<T>());
        PF_LAUNCH_CHECK();
        gather_rows(h, h.loc.get(), y);
    });
    PF_LAUNCH_CHECK();
}

void element_material_vjp(Handle &h, const double *x, const double *p, double *out)
{
    if (h.material == MAT_MR)
        fail(PFGPU_ERR_UNSUPPORTED, "MooneyRivlin has no Lame-parameter derivatives (as in the reference)");
    dispatch(h, [&]<typename T, int O, int Q, int M>() {
        k_material_vjp<T, O, Q, M><<<grid_for(h.n_elements), BLOCK, 0, h.stream>>>(h.mesh<T>(), x, p, out);
    });
    PF_LAUNCH_CHECK();
}

void energy_lame_gradient(Handle &h, const double *x, double *out)
{
    if (h.material == MAT_MR)
        fail(PFGPU_ERR_UNSUPPORTED, "MooneyRivlin has no Lame-parameter derivatives (as in the reference)");
    dispatch(h, [&]<typename T, int O, int Q, int M>() {
        k_energy_lame_grad<T, O, Q, M><<<grid_for(h.n_elements), BLOCK, 0, h.stream>>>(h.mesh<T>(), x, out);
    });
    PF_LAUNCH_CHECK();
}

void plain_shape_vjp(Handle &h, const double *x, const double *p, double *out)
{
    const unsigned grid = grid_for(h.n_elements);
    h.loc.reserve(h.value_size() * size_t(h.n_elements) * 12);
    dispatch(h, [&]<typename T, int O, int Q, int M>() {
        k_shape_vjp<T, O, Q, M><<<grid, BLOCK, 0, h.stream>>>(h.mesh<T>(), x, p, h.loc.as<T>());
        PF_LAUNCH_CHECK();
        k_gather_rows<T><<<grid_for(h.n_vertices), BLOCK, 0, h.stream>>>(h.n_vertices, h.vert_ptr.as<int32_t>(),
                                                                           h.vert_inc.as<int32_t>(), h.loc.as<T>(), out);
    });
    PF_LAUNCH_CHECK();
}

} // namespace pfgpu
