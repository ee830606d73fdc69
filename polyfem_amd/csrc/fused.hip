// Fused energy + gradient + Hessian assembly: one persistent workgroup per CU
// walks a list of node patches (see fused_tables.cpp).
//
// The 16 wavefronts of the workgroup are split into
//   producers (the first WA waves): for the next chunk of incident elements,
//            one thread per (element, quadrature point) computes the
//            kinematics, the energy density, the stress and the Hessian
//            factors into an LDS staging buffer, and copies the chunk's
//            (row, element) items into LDS;
//   consumers (the other waves): lane groups of 3 n_local lanes, each group
//            owning whole rows of the chunk; for an item (row node = local
//            node a of element e) lane (b, m) computes row m of the 3x3
//            block H_e(a, b) and adds it to the row's LDS accumulator at the
//            column position of node_e(b); lanes (0, m) add the gradient.
// Staging and items are double-buffered, so the loads of chunk u+1 overlap
// the accumulation of chunk u.  When a patch is complete its rows are
// streamed to HBM, each as one contiguous segment.  Nothing is atomic: a row
// has exactly one writer and its contributions arrive in element order, so
// results are bitwise reproducible.
#include <type_traits>

#include "dispatch.hpp"
#include "fused.hpp"

namespace pfgpu {

namespace {

template <int ORDER, int NQ> struct RGTab {
    static constexpr int NB = NLocal<ORDER>::value;
    double rg[NQ * NB * 3];
    double rgg[NB * NB * 6];
    constexpr RGTab() : rg(), rgg()
    {
        for (int q = 0; q < NQ; ++q)
            for (int a = 0; a < NB; ++a)
                for (int k = 0; k < 3; ++k) rg[(q * NB + a) * 3 + k] = RG<ORDER, NQ>(q, a, k);
        const int kk[6] = {0, 1, 2, 0, 0, 1}, ll[6] = {0, 1, 2, 1, 2, 2};
        for (int a = 0; a < NB; ++a)
            for (int b = 0; b < NB; ++b)
                for (int i = 0; i < 6; ++i)
                    rgg[(a * NB + b) * 6 + i] = i < 3 ? RGG<ORDER, NQ>(a, b, kk[i], ll[i])
                                                      : RGG<ORDER, NQ>(a, b, kk[i], ll[i]) + RGG<ORDER, NQ>(a, b, ll[i], kk[i]);
    }
};

__constant__ RGTab<1, 1> c_tab11 = RGTab<1, 1>();
__constant__ RGTab<1, 4> c_tab14 = RGTab<1, 4>();
__constant__ RGTab<2, 1> c_tab21 = RGTab<2, 1>();
__constant__ RGTab<2, 4> c_tab24 = RGTab<2, 4>();

template <int ORDER, int NQ> __device__ __forceinline__ const RGTab<ORDER, NQ> &tab()
{
    if constexpr (ORDER == 1 && NQ == 1) return c_tab11;
    else if constexpr (ORDER == 1) return c_tab14;
    else if constexpr (NQ == 1) return c_tab21;
    else return c_tab24;
}

__device__ __forceinline__ float fmat(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double fmat(double a, double b, double c) { return fma(a, b, c); }

// byte offsets of the LDS regions
struct Lds {
    int acc, gacc, stg0, stg1, hdr0, hdr1, pos0, pos1, cg0, cg1, rg, rgg, racc, rdeg, red, total;
};

template <typename T, int ORDER, int NQ> Lds lds_layout(int max_acc, int max_rows)
{
    constexpr int NB = NLocal<ORDER>::value, PS = pos_stride(NB);
    constexpr int CH = chunk_elems<T, NQ>(), IC = item_cap<T, ORDER, NQ>(), NG = n_groups<T, ORDER, NQ>();
    auto al = [](int v) { return (v + 15) & ~15; };
    Lds L{};
    int off = 0;
    auto take = [&](int bytes) {
        const int o = off;
        off = al(off + bytes);
        return o;
    };
    L.acc = take(max_acc * int(sizeof(T)));
    L.gacc = take(3 * max_rows * int(sizeof(T)));
    L.stg0 = take(CH * Stage<T, ORDER, NQ>::SIZE * int(sizeof(T)));
    L.stg1 = take(CH * Stage<T, ORDER, NQ>::SIZE * int(sizeof(T)));
    L.hdr0 = take(IC * 4);
    L.hdr1 = take(IC * 4);
    L.pos0 = take(IC * PS);
    L.pos1 = take(IC * PS);
    L.cg0 = take((NG + 1) * 4);
    L.cg1 = take((NG + 1) * 4);
    L.rg = take(NQ * NB * 4 * int(sizeof(T)));
    L.rgg = take(NB * NB * 8 * int(sizeof(T)));
    L.racc = take(max_rows * 4);
    L.rdeg = take(max_rows * 4);
    L.red = take(2 * NWAVE * 8);
    L.total = off;
    return L;
}

// producer work for one (element, quadrature point Q): fills the point's
// parts of the staging record s of the element (layout Stage<T, ORDER, NQ>)
// and returns the point's contribution to the element gradient in gq
template <typename T, int ORDER, int NQ, int MAT, int Q>
__device__ __forceinline__ void setup_point(const DevMesh<T> &m, int e, const int32_t (&nodes)[NLocal<ORDER>::value],
                                            const double *__restrict__ x, T *__restrict__ s, bool own, double &eacc,
                                            T (&gq)[NLocal<ORDER>::value][3])
{
    using S = Stage<T, ORDER, NQ>;
    constexpr int NB = NLocal<ORDER>::value;
    Elem<T, ORDER, NQ> el;
    load_geometry(m, e, el);
    // relative to local node 0, in FP64 (as gather() in elem.hpp)
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) el.U[a][d] = T(x[3 * int64_t(nodes[a]) + d] - x[3 * int64_t(nodes[0]) + d]);
    const T w = el.det * T(quad_weight(NQ, 0));
    Kin<T> k;
    field_grad<T, ORDER, NQ>(Q, el.U, el.jit, k.H);
    if (MAT == MAT_NH) kinematics(k);
    if (own) eacc += double(energy_density<T, MAT>(k, el.lam, el.mu, el.p3) * w);
    T P[9];
    pk1<T, MAT>(k, el.lam, el.mu, el.p3, P);
    // gradient row of node a at this point: sum_k RG(Q,a,k) Qm[d][k], Qm = w P jit^T
    {
        T Qm[9];
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int kk = 0; kk < 3; ++kk)
                Qm[d * 3 + kk] = w * (P[d * 3 + 0] * el.jit[kk * 3 + 0] + P[d * 3 + 1] * el.jit[kk * 3 + 1] +
                                      P[d * 3 + 2] * el.jit[kk * 3 + 2]);
#pragma unroll
        for (int a = 0; a < NB; ++a)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                T g = T(0);
#pragma unroll
                for (int kk = 0; kk < 3; ++kk) {
                    const double r = RG<ORDER, NQ>(Q, a, kk);
                    if (r != 0.0) g = fmat(Qm[d * 3 + kk], T(r), g);
                }
                gq[a][d] = g;
            }
    }
    T M[9], cl, cc;
    hess_qp<T, MAT>(k, el.jit, el.lam, el.mu, w, M, cl, cc);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        T kb[4];
        khat<T, ORDER, NQ>(Q, b, M, kb);
        kb[3] = T(0);
#pragma unroll
        for (int i = 0; i < 4; ++i) s[S::K + (Q * NB + b) * 4 + i] = kb[i];
    }
    s[S::C + 2 * Q] = cl;
    s[S::C + 2 * Q + 1] = cc;
    if (Q == 0) {
        T g6[6];
        metric(el.jit, g6);
        const T cm = w * el.mu;
#pragma unroll
        for (int i = 0; i < 6; ++i) s[S::G + i] = cm * g6[i];
        s[S::G + 6] = T(0);
        s[S::G + 7] = T(0);
    }
}

template <typename T> struct Vec4 {
    T x, y, z, w;
};
__device__ __forceinline__ Vec4<float> ld4(const float *p)
{
    const float4 v = *reinterpret_cast<const float4 *>(p);
    return {v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ Vec4<double> ld4(const double *p)
{
    const double2 a = *reinterpret_cast<const double2 *>(p), b = *reinterpret_cast<const double2 *>(p + 2);
    return {a.x, a.y, b.x, b.y};
}
template <typename T> struct Vec2 {
    T x, y;
};
__device__ __forceinline__ Vec2<float> ld2(const float *p)
{
    const float2 v = *reinterpret_cast<const float2 *>(p);
    return {v.x, v.y};
}
__device__ __forceinline__ Vec2<double> ld2(const double *p)
{
    const double2 v = *reinterpret_cast<const double2 *>(p);
    return {v.x, v.y};
}

struct Unit {
    int p, c; // patch, chunk; p < 0: none
};

template <typename T, int ORDER, int NQ, int MAT>
__global__ __launch_bounds__(NT) void k_egh(DevMesh<T> m, const double *__restrict__ x, FusedDev fd, Lds L, int n_patches,
                                            T *__restrict__ values, double *__restrict__ grad, double *__restrict__ epart)
{
    constexpr int NB = NLocal<ORDER>::value;
    constexpr int GL = group_lanes<ORDER>();
    constexpr int GPW = groups_per_wave<ORDER>();
    constexpr int WA = producer_waves<T, NQ>();
    constexpr int NG = n_groups<T, ORDER, NQ>();
    constexpr int IC = item_cap<T, ORDER, NQ>();
    constexpr int PS = pos_stride(NB);
    constexpr int NP = WA * 64; // producer threads
    using S = Stage<T, ORDER, NQ>;
    constexpr int SE = S::SIZE;
    constexpr int CH = chunk_elems<T, NQ>();
    constexpr int NPART = grad_parts<T, NQ>();

    extern __shared__ __align__(16) unsigned char smem[];
    T *acc = reinterpret_cast<T *>(smem + L.acc);
    T *gacc = reinterpret_cast<T *>(smem + L.gacc);
    T *rg = reinterpret_cast<T *>(smem + L.rg);
    T *rgg = reinterpret_cast<T *>(smem + L.rgg);
    int *racc = reinterpret_cast<int *>(smem + L.racc);
    int *rdeg = reinterpret_cast<int *>(smem + L.rdeg);
    double *red = reinterpret_cast<double *>(smem + L.red);

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool producer = wave < WA;
    // the producers are the critical path of every chunk and share their
    // SIMDs with consumer waves: they issue first
    if (producer) __builtin_amdgcn_s_setprio(2);

    const RGTab<ORDER, NQ> &tb = tab<ORDER, NQ>();
    // reference-gradient tables padded to 4 and 8 entries for vector loads
    for (int i = tid; i < NQ * NB * 4; i += NT) rg[i] = (i & 3) == 3 ? T(0) : T(tb.rg[(i >> 2) * 3 + (i & 3)]);
    for (int i = tid; i < NB * NB * 8; i += NT) rgg[i] = (i & 7) >= 6 ? T(0) : T(tb.rgg[(i >> 3) * 6 + (i & 7)]);

    // per-patch bookkeeping (uniform)
    auto desc = [&](int p, int f) { return fd.desc[int64_t(p) * PD_STRIDE + f]; };
    auto begin_patch = [&](int p) {
        const int row0 = desc(p, PD_ROW0), nrows = desc(p, PD_NROWS), naccv = desc(p, PD_ACC);
        for (int i = tid; i < naccv; i += NT) acc[i] = T(0);
        for (int i = tid; i < 3 * nrows; i += NT) gacc[i] = T(0);
        for (int r = tid; r < nrows; r += NT) {
            const int node = fd.rows[row0 + r];
            racc[r] = fd.row_acc[row0 + r];
            rdeg[r] = fd.brow_ptr[node + 1] - fd.brow_ptr[node];
        }
    };
    auto next_unit = [&](Unit u) -> Unit {
        if (u.c + 1 < desc(u.p, PD_NCH)) return Unit{u.p, u.c + 1};
        const int np = u.p + int(gridDim.x);
        return np < n_patches ? Unit{np, 0} : Unit{-1, 0};
    };

    // producers.  prefetch(u) issues the loads of unit u's item tables and
    // element node lists into registers; they stay in flight across the
    // barrier and produce(u) stores the tables to LDS and stages the elements.
    using PU = std::conditional_t<PS == 16, uint4, uint32_t>; // column positions of one item
    constexpr int NH = (IC + NP - 1) / NP;
    constexpr int NC = (NG + 1 + NP - 1) / NP;
    struct Pref {
        int nit, cn;
        int32_t ev;
        int32_t nodes[NB];
        int hv[NH], cv[NC];
        PU pv[NH];
    };
    int pq, pkk;
    producer_task<T, NQ>(tid, pq, pkk);
    auto prefetch = [&](Unit u, Pref &f) {
        f.nit = 0;
        f.cn = 0;
        if (u.p < 0) return;
        const int el0 = desc(u.p, PD_EL0), cs0 = desc(u.p, PD_CS0), cgb = desc(u.p, PD_CG0);
        const int cbeg = fd.chunk_start[cs0 + u.c];
        f.cn = fd.chunk_start[cs0 + u.c + 1] - cbeg;
        const int it0 = fd.cg[cgb + u.c * NG];
        f.nit = fd.cg[cgb + (u.c + 1) * NG] - it0;
        const PU *gpos = reinterpret_cast<const PU *>(fd.pos) + it0;
#pragma unroll
        for (int j = 0; j < NH; ++j)
            if (tid + j * NP < f.nit) {
                f.hv[j] = fd.hdr[it0 + tid + j * NP];
                f.pv[j] = gpos[tid + j * NP];
            }
#pragma unroll
        for (int j = 0; j < NC; ++j) f.cv[j] = tid + j * NP <= NG ? fd.cg[cgb + u.c * NG + tid + j * NP] - it0 : 0;
        if (pkk < f.cn) {
            const int64_t v = int64_t(el0) + cbeg + pkk;
            f.ev = fd.elems[v];
#pragma unroll
            for (int a = 0; a < NB; ++a) f.nodes[a] = fd.vnodes[v * NB + a];
        }
    };
    auto produce = [&](int slot, const Pref &f) {
        T *stg = reinterpret_cast<T *>(smem + (slot ? L.stg1 : L.stg0));
        int *lhdr = reinterpret_cast<int *>(smem + (slot ? L.hdr1 : L.hdr0));
        PU *lpu = reinterpret_cast<PU *>(smem + (slot ? L.pos1 : L.pos0));
        int *lcg = reinterpret_cast<int *>(smem + (slot ? L.cg1 : L.cg0));
#pragma unroll
        for (int j = 0; j < NH; ++j)
            if (tid + j * NP < f.nit) {
                lhdr[tid + j * NP] = f.hv[j];
                lpu[tid + j * NP] = f.pv[j];
            }
#pragma unroll
        for (int j = 0; j < NC; ++j)
            if (tid + j * NP <= NG) lcg[tid + j * NP] = f.cv[j];
        double eacc = 0;
        T gq[NB][3];
#pragma unroll
        for (int a = 0; a < NB; ++a) gq[a][0] = gq[a][1] = gq[a][2] = T(0);
        T *s = stg + pkk * SE;
        if (pkk < f.cn) {
            const int e = f.ev & 0x7fffffff;
            const bool own = f.ev < 0;
            if constexpr (NQ == 1) {
                setup_point<T, ORDER, NQ, MAT, 0>(m, e, f.nodes, x, s, own, eacc, gq);
            } else {
                switch (pq) {
                case 0: setup_point<T, ORDER, NQ, MAT, 0>(m, e, f.nodes, x, s, own, eacc, gq); break;
                case 1: setup_point<T, ORDER, NQ, MAT, 1>(m, e, f.nodes, x, s, own, eacc, gq); break;
                case 2: setup_point<T, ORDER, NQ, MAT, 2>(m, e, f.nodes, x, s, own, eacc, gq); break;
                default: setup_point<T, ORDER, NQ, MAT, 3>(m, e, f.nodes, x, s, own, eacc, gq); break;
                }
            }
        }
        // element gradient: sum over the points held by this wavefront (lanes
        // CH apart), written by the lane of the wavefront's first point
#pragma unroll
        for (int off = CH; off < 64; off <<= 1)
#pragma unroll
            for (int a = 0; a < NB; ++a)
#pragma unroll
                for (int d = 0; d < 3; ++d) gq[a][d] += __shfl_xor(gq[a][d], off);
        if (pkk < f.cn && lane < CH) {
            const int h = CH >= 64 ? pq : wave;
#pragma unroll
            for (int a = 0; a < NB; ++a) {
                T *r = s + S::R + (h * NB + a) * 4;
                r[0] = gq[a][0];
                r[1] = gq[a][1];
                r[2] = gq[a][2];
                r[3] = T(0);
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) eacc += __shfl_xor(eacc, off);
        if (lane == 0) red[slot * NWAVE + wave] = eacc;
    };

    // consumers: accumulate unit u from buffer slot.  Group lane (b, mr)
    // forms row mr of the 3x3 block (a, b) of each item; lanes with b = 0 add
    // gradient component mr of row a.
    const int cw = wave - WA;
    const int gi = lane / GL, gl = lane - gi * GL;
    const int b = gl / 3, mr = gl - 3 * (gl / 3);
    const bool in_group = !producer && gi < GPW;
    const int group = cw * GPW + gi;
    auto consume = [&](int slot) {
        const T *stg = reinterpret_cast<const T *>(smem + (slot ? L.stg1 : L.stg0));
        const int *lhdr = reinterpret_cast<const int *>(smem + (slot ? L.hdr1 : L.hdr0));
        const uint8_t *lpos = reinterpret_cast<const uint8_t *>(smem + (slot ? L.pos1 : L.pos0));
        const int *lcg = reinterpret_cast<const int *>(smem + (slot ? L.cg1 : L.cg0));
        if (!in_group) return;
        const int i0 = lcg[group], i1 = lcg[group + 1];
        for (int it = i0; it < i1; ++it) {
            const int hdr = lhdr[it];
            const int kk = hdr & 1023, a = (hdr >> 10) & 15, r = hdr >> 14;
            const T *s = stg + kk * SE;
            const int pos = lpos[it * PS + b];
            if (b == 0) {
                // gradient entry (row a, component mr)
                T gs = s[S::R + a * 4 + mr];
#pragma unroll
                for (int h = 1; h < NPART; ++h) gs += s[S::R + (h * NB + a) * 4 + mr];
                gacc[3 * r + mr] += gs;
            }
            T o0 = T(0), o1 = T(0), o2 = T(0);
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const Vec4<T> ka = ld4(s + S::K + (q * NB + a) * 4), kb = ld4(s + S::K + (q * NB + b) * 4);
                const Vec2<T> c = ld2(s + S::C + 2 * q);
                const T u = c.x * s[S::K + (q * NB + a) * 4 + mr], v = c.y * s[S::K + (q * NB + b) * 4 + mr];
                o0 = fmat(u, kb.x, fmat(v, ka.x, o0));
                o1 = fmat(u, kb.y, fmat(v, ka.y, o1));
                o2 = fmat(u, kb.z, fmat(v, ka.z, o2));
            }
            {
                const Vec4<T> g0 = ld4(s + S::G), c0 = ld4(rgg + (a * NB + b) * 8);
                const Vec2<T> g4 = ld2(s + S::G + 4), c4 = ld2(rgg + (a * NB + b) * 8 + 4);
                const T sab = fmat(g0.x, c0.x, fmat(g0.y, c0.y, fmat(g0.z, c0.z, fmat(g0.w, c0.w, fmat(g4.x, c4.x, g4.y * c4.y)))));
                o0 += mr == 0 ? sab : T(0);
                o1 += mr == 1 ? sab : T(0);
                o2 += mr == 2 ? sab : T(0);
            }
            T *ar = acc + racc[r] + mr * 3 * rdeg[r] + 3 * pos;
            ar[0] += o0;
            ar[1] += o1;
            ar[2] += o2;
        }
    };

    Unit cur{int(blockIdx.x), 0};
    if (cur.p >= n_patches) return; // uniform
    begin_patch(cur.p);
    Pref pf;
    if (producer) prefetch(cur, pf);
    __syncthreads(); // rg tables
    if (producer) {
        produce(0, pf);
        prefetch(next_unit(cur), pf);
    }
    __syncthreads();
    if (tid == 0) {
        double t = 0;
        for (int w = 0; w < WA; ++w) t += red[w];
        epart[desc(cur.p, PD_CHUNK_BASE) + cur.c] = t;
    }
    int slot = 0;
    while (true) {
        const Unit nxt = next_unit(cur);
        if (producer) {
            if (nxt.p >= 0) {
                produce(slot ^ 1, pf);
                prefetch(next_unit(nxt), pf);
            }
        } else {
            consume(slot);
        }
        __syncthreads();
        if (nxt.p >= 0 && tid == 0) {
            double t = 0;
            for (int w = 0; w < WA; ++w) t += red[(slot ^ 1) * NWAVE + w];
            epart[desc(nxt.p, PD_CHUNK_BASE) + nxt.c] = t;
        }
        if (nxt.p != cur.p) {
            // patch complete: stream its rows out, one wavefront per row
            const int row0 = desc(cur.p, PD_ROW0), nrows = desc(cur.p, PD_NROWS);
            for (int r = wave; r < nrows; r += NWAVE) {
                const int node = fd.rows[row0 + r];
                const int64_t goff = 9 * int64_t(fd.brow_ptr[node]);
                const int len = 9 * rdeg[r], a0 = racc[r];
                for (int j = lane; j < len; j += 64) values[goff + j] = acc[a0 + j];
            }
            for (int i = tid; i < 3 * nrows; i += NT) grad[3 * int64_t(fd.rows[row0 + i / 3]) + i % 3] = double(gacc[i]);
            if (nxt.p < 0) break;
            __syncthreads();
            begin_patch(nxt.p);
            __syncthreads();
        }
        cur = nxt;
        slot ^= 1;
    }
}

FusedDev device_view(const Handle &h)
{
    return FusedDev{h.patch_desc.as<int32_t>(), h.patch_rows.as<int32_t>(), h.patch_row_acc.as<int32_t>(),
                    h.patch_elems.as<int32_t>(), h.patch_vnodes.as<int32_t>(), h.patch_cs.as<int32_t>(), h.patch_cg.as<int32_t>(),
                    h.patch_hdr.as<int32_t>(), h.patch_pos.as<uint8_t>(), h.brow_ptr.as<int32_t>()};
}

template <typename T, int ORDER, int NQ> FusedShape shape_for()
{
    FusedShape s{};
    s.groups = n_groups<T, ORDER, NQ>();
    s.chunk = chunk_elems<T, NQ>();
    s.item_cap = item_cap<T, ORDER, NQ>();
    s.max_rows = 256;
    const Lds L = lds_layout<T, ORDER, NQ>(0, s.max_rows);
    s.acc_capacity = (LDS_LIMIT - L.total) / int(sizeof(T));
    return s;
}

template <typename T, int ORDER, int NQ, int MAT>
void launch_egh(Handle &h, const double *x, void *values, double *g, double *energy)
{
    const Lds L = lds_layout<T, ORDER, NQ>(h.fused_max_acc, h.fused_max_rows);
    if (L.total > LDS_LIMIT) fail(PFGPU_ERR_UNSUPPORTED, "fused kernel LDS layout exceeds 160 KB");
    auto kern = k_egh<T, ORDER, NQ, MAT>;
    PF_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern), hipFuncAttributeMaxDynamicSharedMemorySize, L.total));
    int dev = 0, n_cu = 0;
    PF_HIP(hipGetDevice(&dev));
    PF_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    const int grid = std::min(h.n_patches, std::max(1, n_cu));
    h.partial.reserve(sizeof(double) * size_t(h.fused_chunks));
    kern<<<grid, NT, L.total, h.stream>>>(h.mesh<T>(), x, device_view(h), L, h.n_patches, static_cast<T *>(values), g,
                                          h.partial.as<double>());
    PF_LAUNCH_CHECK();
    if (energy) sum_doubles(h, h.partial.as<double>(), int(h.fused_chunks), energy);
}

} // namespace

void ensure_fused_tables(Handle &h)
{
    if (h.fused_ready) return;
    FusedShape s{};
    dispatch<true>(h, [&]<typename T, int O, int Q, int M>() { s = shape_for<T, O, Q>(); });
    FusedTables t = build_fused_tables(h.pat, h.n_local, h.h_elem_nodes.data(), h.n_elements, h.h_vertices.data(),
                                       h.h_elem_vertices.data(), s);
    h.patch_desc = upload(t.desc, h.stream);
    h.patch_rows = upload(t.rows, h.stream);
    h.patch_row_acc = upload(t.row_acc, h.stream);
    h.patch_elems = upload(t.elems, h.stream);
    h.patch_vnodes = upload(t.vnodes, h.stream);
    h.patch_cs = upload(t.chunk_start, h.stream);
    h.patch_cg = upload(t.cg, h.stream);
    h.patch_hdr = upload(t.hdr, h.stream);
    h.patch_pos = upload(t.pos, h.stream);
    if (!h.brow_ptr.get()) h.brow_ptr = upload(h.pat.brow_ptr, h.stream);
    h.n_patches = t.n_patches;
    h.fused_max_acc = t.max_acc;
    h.fused_max_rows = t.max_rows;
    h.fused_chunks = t.n_chunks;
    h.fused_visits = t.visits;
    // one energy partial per chunk: sized with the tables, not by a later call
    h.partial.reserve(sizeof(double) * size_t(h.fused_chunks));
    h.fused_ready = true;
}

void fused_gradient(Handle &h, const double *x, double *g, double *energy) { plain_gradient(h, x, g, energy); }

void fused_energy_gradient_hessian(Handle &h, const double *x, double *energy, double *g, void *values)
{
    if (!values) {
        if (g)
            plain_gradient(h, x, g, energy);
        else if (energy)
            plain_energy(h, x, energy, nullptr);
        return;
    }
    if (h.material != MAT_NH && h.material != MAT_LE)
        fail(PFGPU_ERR_UNSUPPORTED, "the fused path covers NeoHookean and linear elasticity");
    ensure_fused_tables(h);
    double *gv = g;
    if (!gv) {
        h.loc.reserve(sizeof(double) * size_t(h.n_dofs));
        gv = h.loc.as<double>();
    }
    dispatch<true>(h, [&]<typename T, int O, int Q, int M>() {
        if constexpr (M == MAT_NH || M == MAT_LE) launch_egh<T, O, Q, M>(h, x, values, gv, energy);
    });
}

void fused_hvp(Handle &h, const double *x, const double *v, double *y) { plain_hvp(h, x, v, y); }

} // namespace pfgpu
