// Fused assembly path: one workgroup per node patch (see fused_tables.cpp).
//
// Per chunk of incident elements:
//   phase A  one thread per (element, quadrature point), the point uniform
//            across a wavefront: kinematics, energy density, stress and the
//            Hessian factors go to LDS;
//   phase B  lane groups of n_local lanes, each group owning whole rows: for
//            an item (row node = local node a of element e) lane b computes
//            the 3x3 block H_e(a, b) and adds it to the row's LDS
//            accumulator at the column position of node_e(b); lanes 0..2 add
//            the gradient entries.
// The owned rows are then streamed to HBM, each as one contiguous segment.
// Nothing is atomic: a row has exactly one writer and its contributions
// arrive in element order.
#include "dispatch.hpp"
#include "fused.hpp"

namespace pfgpu {

namespace {

enum { MODE_EGH = 0, MODE_EG = 1, MODE_HV = 2 };

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

// byte offsets of the LDS regions
struct FusedLds {
    int acc, gacc, stg, rg, rgg, racc, rdeg, red, total;
};

constexpr int SQ = 20; // staging per (element, q): M[9], cl, cc, Q[9]

template <int NQ> constexpr int stage_stride() { return SQ * NQ + 8; } // + metric[6] + pad

template <typename T, int NQ> constexpr int chunk_for(int threads)
{
    return (sizeof(T) == 4 ? threads : threads / 2) / NQ;
}

// phase A for one (element, quadrature point Q)
template <typename T, int ORDER, int NQ, int MAT, int MODE, int Q>
__device__ __forceinline__ void setup_point(const DevMesh<T> &m, int e, const double *__restrict__ x,
                                            const double *__restrict__ v, T *__restrict__ s, bool own, double &eacc)
{
    Elem<T, ORDER, NQ> el;
    load_geometry(m, e, el);
    gather(m, e, x, el.U);
    const T w = el.det * T(quad_weight(NQ));
    Kin<T> k;
    field_grad<T, ORDER, NQ>(Q, el.U, el.jit, k.H);
    if (MAT == MAT_NH) kinematics(k);
    T S[9];
    if (MODE == MODE_HV) {
        T V[NLocal<ORDER>::value][3];
        gather(m, e, v, V);
        T Gv[9];
        field_grad<T, ORDER, NQ>(Q, V, el.jit, Gv);
        dstress<T, MAT>(k, el.lam, el.mu, Gv, S);
    } else {
        if (own) eacc += double(energy_density<T, MAT>(k, el.lam, el.mu) * w);
        pk1<T, MAT>(k, el.lam, el.mu, S);
    }
    T *sq = s + Q * SQ;
    // Q = w S jit^T: the gradient row of node a is sum_q sum_k RG(q,a,k) Q[d][k]
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int kk = 0; kk < 3; ++kk)
            sq[11 + d * 3 + kk] = w * (S[d * 3 + 0] * el.jit[kk * 3 + 0] + S[d * 3 + 1] * el.jit[kk * 3 + 1] + S[d * 3 + 2] * el.jit[kk * 3 + 2]);
    if (MODE == MODE_EGH) {
        T M[9], cl, cc;
        hess_qp<T, MAT>(k, el.jit, el.lam, el.mu, w, M, cl, cc);
#pragma unroll
        for (int i = 0; i < 9; ++i) sq[i] = M[i];
        sq[9] = cl;
        sq[10] = cc;
        if (Q == 0) {
            T g6[6];
            metric(el.jit, g6);
            const T cm = w * el.mu;
#pragma unroll
            for (int i = 0; i < 6; ++i) s[NQ * SQ + i] = cm * g6[i];
        }
    }
}

template <typename T, int ORDER, int NQ, int MAT, int MODE, int NT>
__global__ __launch_bounds__(NT) void k_fused(DevMesh<T> m, const double *__restrict__ x, const double *__restrict__ v,
                                              FusedDev fd, FusedLds L, T *__restrict__ values, double *__restrict__ out_vec,
                                              double *__restrict__ epart)
{
    constexpr int NB = NLocal<ORDER>::value;
    constexpr int GPW = 64 / NB;
    constexpr int NG = (NT / 64) * GPW;
    constexpr int CH = chunk_for<T, NQ>(NT);
    constexpr int SE = stage_stride<NQ>();
    constexpr bool HESS = MODE == MODE_EGH;

    extern __shared__ __align__(16) unsigned char smem[];
    T *acc = reinterpret_cast<T *>(smem + L.acc);
    T *gacc = reinterpret_cast<T *>(smem + L.gacc);
    T *stg = reinterpret_cast<T *>(smem + L.stg);
    T *rg = reinterpret_cast<T *>(smem + L.rg);
    T *rgg = reinterpret_cast<T *>(smem + L.rgg);
    int *racc = reinterpret_cast<int *>(smem + L.racc);
    int *rdeg = reinterpret_cast<int *>(smem + L.rdeg);
    double *red = reinterpret_cast<double *>(smem + L.red);

    const int tid = threadIdx.x;
    const int32_t *D = fd.desc + 8 * blockIdx.x;
    const int row0 = D[0], nrows = D[1], el0 = D[2], nel = D[3], cg0 = D[4], nchunks = D[5], naccv = D[6];

    if (HESS)
        for (int i = tid; i < naccv; i += NT) acc[i] = T(0);
    for (int i = tid; i < 3 * nrows; i += NT) gacc[i] = T(0);
    for (int r = tid; r < nrows; r += NT) {
        const int node = fd.rows[row0 + r];
        racc[r] = fd.row_acc[row0 + r];
        rdeg[r] = fd.brow_ptr[node + 1] - fd.brow_ptr[node];
    }
    const RGTab<ORDER, NQ> &tb = tab<ORDER, NQ>();
    for (int i = tid; i < NQ * NB * 3; i += NT) rg[i] = T(tb.rg[i]);
    if (HESS)
        for (int i = tid; i < NB * NB * 6; i += NT) rgg[i] = T(tb.rgg[i]);
    __syncthreads();

    const int lane = tid & 63, wave = tid >> 6;
    const int gi = lane / NB, b = lane - gi * NB;
    const bool in_group = gi < GPW;
    const int group = wave * GPW + gi;
    T rgb[NQ][3];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int k = 0; k < 3; ++k) rgb[q][k] = in_group ? rg[(q * NB + b) * 3 + k] : T(0);

    double eacc = 0;
    for (int c = 0; c < nchunks; ++c) {
        const int cbeg = c * CH, cn = min(CH, nel - cbeg);
        if (c > 0) __syncthreads(); // staging of the previous chunk consumed
        // ---- phase A: tasks (element, q), q uniform per wavefront
        if (wave < (CH * NQ) / 64) {
            const int q = wave % NQ;
            const int kk = (wave / NQ) * 64 + lane;
            if (kk < cn) {
                const int32_t ev = fd.elems[el0 + cbeg + kk];
                const int e = ev & 0x7fffffff;
                const bool own = ev < 0;
                T *s = stg + kk * SE;
                if constexpr (NQ == 1) {
                    setup_point<T, ORDER, NQ, MAT, MODE, 0>(m, e, x, v, s, own, eacc);
                } else {
                    switch (q) {
                    case 0: setup_point<T, ORDER, NQ, MAT, MODE, 0>(m, e, x, v, s, own, eacc); break;
                    case 1: setup_point<T, ORDER, NQ, MAT, MODE, 1>(m, e, x, v, s, own, eacc); break;
                    case 2: setup_point<T, ORDER, NQ, MAT, MODE, 2>(m, e, x, v, s, own, eacc); break;
                    default: setup_point<T, ORDER, NQ, MAT, MODE, 3>(m, e, x, v, s, own, eacc); break;
                    }
                }
            }
        }
        __syncthreads();
        // ---- phase B: items of this chunk, per lane group
        if (in_group) {
            const int it0 = fd.cg[cg0 + c * NG + group], it1 = fd.cg[cg0 + c * NG + group + 1];
            for (int it = it0; it < it1; ++it) {
                const int hdr = fd.hdr[it];
                const int kk = hdr & 1023, a = (hdr >> 10) & 15, r = hdr >> 14;
                const T *s = stg + kk * SE;
                if (b < 3) {
                    T gs = T(0);
#pragma unroll
                    for (int q = 0; q < NQ; ++q)
#pragma unroll
                        for (int k = 0; k < 3; ++k) gs += rg[(q * NB + a) * 3 + k] * s[q * SQ + 11 + b * 3 + k];
                    gacc[3 * r + b] += gs;
                }
                if (HESS) {
                    const int pos = fd.pos[int64_t(it) * NB + b];
                    T blk[9];
#pragma unroll
                    for (int i = 0; i < 9; ++i) blk[i] = T(0);
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        const T *sq = s + q * SQ;
                        T kb[3], ka[3];
#pragma unroll
                        for (int mm = 0; mm < 3; ++mm)
                            kb[mm] = sq[mm * 3 + 0] * rgb[q][0] + sq[mm * 3 + 1] * rgb[q][1] + sq[mm * 3 + 2] * rgb[q][2];
#pragma unroll
                        for (int mm = 0; mm < 3; ++mm) ka[mm] = __shfl(kb[mm], gi * NB + a);
                        const T cl = sq[9], cc = sq[10];
#pragma unroll
                        for (int mm = 0; mm < 3; ++mm)
#pragma unroll
                            for (int nn = 0; nn < 3; ++nn) blk[mm * 3 + nn] += cl * ka[mm] * kb[nn] + cc * kb[mm] * ka[nn];
                    }
                    const T *g6 = s + NQ * SQ;
                    const T *cg = rgg + (a * NB + b) * 6;
                    const T sab = g6[0] * cg[0] + g6[1] * cg[1] + g6[2] * cg[2] + g6[3] * cg[3] + g6[4] * cg[4] + g6[5] * cg[5];
                    blk[0] += sab;
                    blk[4] += sab;
                    blk[8] += sab;
                    const int deg = rdeg[r];
                    T *ar = acc + racc[r] + 3 * pos;
#pragma unroll
                    for (int mm = 0; mm < 3; ++mm)
#pragma unroll
                        for (int nn = 0; nn < 3; ++nn) ar[mm * 3 * deg + nn] += blk[mm * 3 + nn];
                }
            }
        }
    }
    __syncthreads();

    // ---- write-out: one wavefront per row, each row one contiguous segment
    if (HESS) {
        for (int r = wave; r < nrows; r += NT / 64) {
            const int node = fd.rows[row0 + r];
            const int64_t goff = 9 * int64_t(fd.brow_ptr[node]);
            const int len = 9 * rdeg[r], a0 = racc[r];
            for (int j = lane; j < len; j += 64) values[goff + j] = acc[a0 + j];
        }
    }
    for (int i = tid; i < 3 * nrows; i += NT) out_vec[3 * int64_t(fd.rows[row0 + i / 3]) + i % 3] = double(gacc[i]);

    if (MODE != MODE_HV) {
        // fixed-shape reduction of the energy partials
        double s = eacc;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) red[wave] = s;
        __syncthreads();
        if (tid == 0) {
            double t = 0;
            for (int w = 0; w < NT / 64; ++w) t += red[w];
            epart[blockIdx.x] = t;
        }
    }
}

constexpr int FUSED_THREADS = 512;
constexpr int LDS_LIMIT = 160 * 1024;

template <typename T, int ORDER, int NQ> FusedLds lds_layout(int max_acc, int max_rows, bool hess)
{
    constexpr int NB = NLocal<ORDER>::value;
    auto al = [](int v) { return (v + 15) & ~15; };
    FusedLds L{};
    int off = 0;
    L.acc = off;
    off = al(off + (hess ? max_acc : 0) * int(sizeof(T)));
    L.gacc = off;
    off = al(off + 3 * max_rows * int(sizeof(T)));
    L.stg = off;
    off = al(off + chunk_for<T, NQ>(FUSED_THREADS) * stage_stride<NQ>() * int(sizeof(T)));
    L.rg = off;
    off = al(off + NQ * NB * 3 * int(sizeof(T)));
    L.rgg = off;
    off = al(off + (hess ? NB * NB * 6 : 0) * int(sizeof(T)));
    L.racc = off;
    off = al(off + max_rows * 4);
    L.rdeg = off;
    off = al(off + max_rows * 4);
    L.red = off;
    off = al(off + (FUSED_THREADS / 64) * 8);
    L.total = off;
    return L;
}

FusedDev device_view(const Handle &h)
{
    return FusedDev{h.patch_desc.as<int32_t>(), h.patch_rows.as<int32_t>(), h.patch_row_acc.as<int32_t>(),
                    h.patch_elems.as<int32_t>(), h.patch_cg.as<int32_t>(), h.patch_hdr.as<int32_t>(),
                    h.patch_pos.as<uint8_t>(), h.brow_ptr.as<int32_t>()};
}

template <typename T, int ORDER, int NQ, int MAT, int MODE>
void launch(Handle &h, const double *x, const double *v, void *values, double *vec, double *energy)
{
    const bool hess = MODE == MODE_EGH;
    const FusedLds L = lds_layout<T, ORDER, NQ>(h.fused_max_acc, h.fused_max_rows, hess);
    if (L.total > LDS_LIMIT) fail(PFGPU_ERR_UNSUPPORTED, "fused kernel LDS layout exceeds 160 KB");
    auto kern = k_fused<T, ORDER, NQ, MAT, MODE, FUSED_THREADS>;
    PF_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern), hipFuncAttributeMaxDynamicSharedMemorySize, L.total));
    h.partial.reserve(sizeof(double) * size_t(h.n_patches));
    kern<<<h.n_patches, FUSED_THREADS, L.total, h.stream>>>(h.mesh<T>(), x, v, device_view(h), L, static_cast<T *>(values),
                                                             vec, h.partial.as<double>());
    PF_LAUNCH_CHECK();
    if (energy && MODE != MODE_HV) sum_doubles(h, h.partial.as<double>(), h.n_patches, energy);
}

template <typename T, int ORDER, int NQ> FusedShape shape_for()
{
    constexpr int NB = NLocal<ORDER>::value;
    FusedShape s{};
    s.threads = FUSED_THREADS;
    s.groups = (FUSED_THREADS / 64) * (64 / NB);
    s.chunk = chunk_for<T, NQ>(FUSED_THREADS);
    s.max_rows = 256;
    const FusedLds L = lds_layout<T, ORDER, NQ>(0, s.max_rows, true);
    s.acc_capacity = (LDS_LIMIT - L.total) / int(sizeof(T));
    return s;
}

} // namespace

void ensure_fused_tables(Handle &h)
{
    if (h.fused_ready) return;
    FusedShape s{};
    dispatch(h, [&]<typename T, int O, int Q, int M>() { s = shape_for<T, O, Q>(); });
    FusedTables t = build_fused_tables(h.pat, h.n_local, h.h_elem_nodes.data(), h.n_elements, h.h_vertices.data(),
                                       h.h_elem_vertices.data(), s);
    h.patch_desc = upload(t.desc, h.stream);
    h.patch_rows = upload(t.rows, h.stream);
    h.patch_row_acc = upload(t.row_acc, h.stream);
    h.patch_elems = upload(t.elems, h.stream);
    h.patch_cg = upload(t.cg, h.stream);
    h.patch_hdr = upload(t.hdr, h.stream);
    h.patch_pos = upload(t.pos, h.stream);
    if (!h.brow_ptr.get()) h.brow_ptr = upload(h.pat.brow_ptr, h.stream);
    h.n_patches = t.n_patches;
    h.fused_max_acc = t.max_acc;
    h.fused_max_rows = t.max_rows;
    h.fused_visits = t.visits;
    h.fused_ready = true;
}

void fused_gradient(Handle &h, const double *x, double *g, double *energy)
{
    ensure_fused_tables(h);
    dispatch(h, [&]<typename T, int O, int Q, int M>() { launch<T, O, Q, M, MODE_EG>(h, x, nullptr, nullptr, g, energy); });
}

void fused_energy_gradient_hessian(Handle &h, const double *x, double *energy, double *g, void *values)
{
    ensure_fused_tables(h);
    if (!values) {
        if (g) {
            fused_gradient(h, x, g, energy);
        } else if (energy) {
            plain_energy(h, x, energy, nullptr);
        }
        return;
    }
    double *gv = g;
    if (!gv) {
        h.loc.reserve(sizeof(double) * size_t(h.n_dofs));
        gv = h.loc.as<double>();
    }
    dispatch(h, [&]<typename T, int O, int Q, int M>() { launch<T, O, Q, M, MODE_EGH>(h, x, nullptr, values, gv, energy); });
}

void fused_hvp(Handle &h, const double *x, const double *v, double *y)
{
    ensure_fused_tables(h);
    dispatch(h, [&]<typename T, int O, int Q, int M>() { launch<T, O, Q, M, MODE_HV>(h, x, v, nullptr, y, nullptr); });
}

} // namespace pfgpu
