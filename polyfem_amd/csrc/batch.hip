// Mixed batches: one launch per kernel over element groups of different
// order and material (P1/P2 x NeoHookean/linear elasticity).
//
// A batch handle holds one sub-handle per (order, material) group, each the
// concatenation of its problems' meshes (disjoint node ranges), and a global
// node space that concatenates the groups in order.  The element kernels of
// the batched Newton-CG (energy, gradient, Hv, diagonal blocks of the
// preconditioner) run as one launch each: workgroup b serves the group whose
// range of workgroups holds b and calls that group's element code (order,
// quadrature and material are compile-time in each branch); the node-row
// gathers likewise run as one launch over the concatenated row plans.  The
// global vectors (x, gradient, CG vectors) are the groups' dof ranges back to
// back, so every vector kernel of the solver is one launch already.
// Summation orders are those of the single-group kernels: results are
// reproducible.
#include "dispatch.hpp"

namespace pfgpu {

namespace {

constexpr int MAX_GROUPS = 4;

// variant of a group: (order, material) of the default quadrature
enum { V_P1_NH = 0, V_P1_LE = 1, V_P2_NH = 2, V_P2_LE = 3 };

template <typename T> struct GroupView {
    DevMesh<T> m;
    int variant;
    int block0;       // first element workgroup of the group in a launch
    int chunk0;       // first row-gather workgroup of the group in a launch
    int n_chunks;     // the group's row-plan chunks
    int64_t dof0;     // first global dof of the group's nodes
    int64_t elem0;    // first global element of the group
    T *loc;           // element-local results (3 or 6 per local node)
    const int32_t *chunk, *slot, *rmeta; // the group's row plan
    SegSkip skip;     // problems whose CG has finished (group-local indices)
};
template <typename T> struct Groups {
    GroupView<T> g[MAX_GROUPS];
    int n;
};

template <typename T> __device__ __forceinline__ int group_of_block(const Groups<T> &G, int b)
{
    int k = 0;
#pragma unroll
    for (int i = 1; i < MAX_GROUPS; ++i)
        if (i < G.n && b >= G.g[i].block0) k = i;
    return k;
}
template <typename T> __device__ __forceinline__ int group_of_chunk(const Groups<T> &G, int c)
{
    int k = 0;
#pragma unroll
    for (int i = 1; i < MAX_GROUPS; ++i)
        if (i < G.n && c >= G.g[i].chunk0) k = i;
    return k;
}

// f.template operator()<ORDER, NQ, MAT>() for a group's variant
template <typename F> __device__ __forceinline__ void on_variant(int v, F &&f)
{
    switch (v) {
    case V_P1_NH: f.template operator()<1, 1, MAT_NH>(); break;
    case V_P1_LE: f.template operator()<1, 1, MAT_LE>(); break;
    case V_P2_NH: f.template operator()<2, 4, MAT_NH>(); break;
    default: f.template operator()<2, 4, MAT_LE>(); break;
    }
}

__device__ __forceinline__ double wg_sum(double v)
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

// element energies (per_elem, global element index) and per-workgroup partials
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_multi_energy(Groups<T> G, const double *__restrict__ x,
                                                        double *__restrict__ per_elem, double *__restrict__ partial)
{
    const GroupView<T> &gv = G.g[group_of_block(G, blockIdx.x)];
    const int e = (int(blockIdx.x) - gv.block0) * BLOCK + threadIdx.x;
    double val = 0;
    if (e < gv.m.n_elements) {
        on_variant(gv.variant, [&]<int O, int Q, int M>() {
            Elem<T, O, Q> el;
            load_geometry(gv.m, e, el);
            gather(gv.m, e, x + gv.dof0, el.U);
            const T w = el.det * T(quad_weight(Q, 0));
            T acc = T(0);
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                Kin<T> k;
                field_grad<T, O, Q>(q, el.U, el.jit, k.H);
                kinematics(k);
                acc += energy_density<T, M>(k, el.lam, el.mu, el.p3) * w;
            }
            val = double(acc);
        });
        if (per_elem) per_elem[gv.elem0 + e] = val;
    }
    const double s = wg_sum(val);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// element-local gradients into the group's loc[(e NB + a) 3 + d]
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_multi_gradient(Groups<T> G, const double *__restrict__ x)
{
    const GroupView<T> &gv = G.g[group_of_block(G, blockIdx.x)];
    const int e = (int(blockIdx.x) - gv.block0) * BLOCK + threadIdx.x;
    if (e >= gv.m.n_elements) return;
    on_variant(gv.variant, [&]<int O, int Q, int M>() {
        constexpr int NB = NLocal<O>::value;
        Elem<T, O, Q> el;
        load_geometry(gv.m, e, el);
        gather(gv.m, e, x + gv.dof0, el.U);
        const T w = el.det * T(quad_weight(Q, 0));
        T g[NB][3];
#pragma unroll
        for (int a = 0; a < NB; ++a) g[a][0] = g[a][1] = g[a][2] = T(0);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            Kin<T> k;
            field_grad<T, O, Q>(q, el.U, el.jit, k.H);
            kinematics(k);
            T P[9];
            pk1<T, M>(k, el.lam, el.mu, el.p3, P);
            contract_add<T, O, Q>(q, P, el.jit, w, g);
        }
        T *out = gv.loc + int64_t(e) * NB * 3;
#pragma unroll
        for (int a = 0; a < NB; ++a)
#pragma unroll
            for (int d = 0; d < 3; ++d) out[a * 3 + d] = g[a][d];
    });
}

// element-local H(x) v into the group's loc (elements of finished problems skip)
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_multi_hvp(Groups<T> G, const double *__restrict__ x, const double *__restrict__ v)
{
    const GroupView<T> &gv = G.g[group_of_block(G, blockIdx.x)];
    const int e = (int(blockIdx.x) - gv.block0) * BLOCK + threadIdx.x;
    if (e >= gv.m.n_elements || gv.skip.element(e)) return;
    on_variant(gv.variant, [&]<int O, int Q, int M>() {
        constexpr int NB = NLocal<O>::value;
        Elem<T, O, Q> el;
        load_geometry(gv.m, e, el);
        gather(gv.m, e, x + gv.dof0, el.U);
        T V[NB][3];
        gather(gv.m, e, v + gv.dof0, V);
        const T w = el.det * T(quad_weight(Q, 0));
        T y[NB][3];
#pragma unroll
        for (int a = 0; a < NB; ++a) y[a][0] = y[a][1] = y[a][2] = T(0);
        // the arithmetic at a raised wave priority, so that other wavefronts'
        // node gathers and stores run around it (232 -> 192 us per batched
        // CG iteration on the 256-problem batch)
        __builtin_amdgcn_s_setprio(1);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            Kin<T> k;
            field_grad<T, O, Q>(q, el.U, el.jit, k.H);
            if (M == MAT_NH) kinematics(k);
            T Gv[9], A[9];
            field_grad<T, O, Q>(q, V, el.jit, Gv);
            dstress<T, M>(k, el.lam, el.mu, el.p3, Gv, A);
            contract_add<T, O, Q>(q, A, el.jit, w, y);
        }
        __builtin_amdgcn_s_setprio(0);
        T *out = gv.loc + int64_t(e) * NB * 3;
#pragma unroll
        for (int a = 0; a < NB; ++a)
#pragma unroll
            for (int d = 0; d < 3; ++d) out[a * 3 + d] = y[a][d];
    });
}

// the element's diagonal blocks H_e(a, a) (00 11 22 01 02 12) into the
// group's loc[(e NB + a) 6 + i]: H_e(a,a) = w sum_q [mu (G_a.G_a) I + cl k_a k_a^T + cc k_a k_a^T]
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_multi_diag(Groups<T> G, const double *__restrict__ x)
{
    const GroupView<T> &gv = G.g[group_of_block(G, blockIdx.x)];
    const int e = (int(blockIdx.x) - gv.block0) * BLOCK + threadIdx.x;
    if (e >= gv.m.n_elements) return;
    on_variant(gv.variant, [&]<int O, int Q, int M>() {
        constexpr int NB = NLocal<O>::value;
        Elem<T, O, Q> el;
        load_geometry(gv.m, e, el);
        if (M != MAT_LE) gather(gv.m, e, x + gv.dof0, el.U);
        const T w = el.det * T(quad_weight(Q, 0));
        T d6[NB][6];
#pragma unroll
        for (int a = 0; a < NB; ++a)
#pragma unroll
            for (int i = 0; i < 6; ++i) d6[a][i] = T(0);
        T g6[6];
        metric(el.jit, g6);
#pragma unroll
        for (int a = 0; a < NB; ++a) {
            const T gd = w * el.mu * grad_dot<O, Q>(a, a, g6);
            d6[a][0] = d6[a][1] = d6[a][2] = gd;
        }
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            Kin<T> k;
            if (M != MAT_LE) field_grad<T, O, Q>(q, el.U, el.jit, k.H);
            if (M == MAT_NH) kinematics(k);
            T Mq[9], cl, cc;
            hess_qp<T, M>(k, el.jit, el.lam, el.mu, w, Mq, cl, cc);
#pragma unroll
            for (int a = 0; a < NB; ++a) {
                T ka[3];
                khat<T, O, Q>(q, a, Mq, ka);
                const T s = cl + cc;
                d6[a][0] += s * ka[0] * ka[0];
                d6[a][1] += s * ka[1] * ka[1];
                d6[a][2] += s * ka[2] * ka[2];
                d6[a][3] += s * ka[0] * ka[1];
                d6[a][4] += s * ka[0] * ka[2];
                d6[a][5] += s * ka[1] * ka[2];
            }
        }
        T *out = gv.loc + int64_t(e) * NB * 6;
#pragma unroll
        for (int a = 0; a < NB; ++a)
#pragma unroll
            for (int i = 0; i < 6; ++i) out[a * 6 + i] = d6[a][i];
    });
}

// Node-row gather over the groups' row plans (layout of k_gather_rows_seg:
// chunk c covers rows chunk[c] .. chunk[c+1]-1, whose contributions sit in
// the slots c SEG .. (c+1) SEG - 1, rmeta[i] = q0 | count << 10).  W = 3:
// FP64 sums into out (3 per node, global dofs; rows of finished problems
// become 0, or are left as they are when the whole chunk's problems have
// finished); W = 6: the diagonal blocks, in T, into diag (9 per node).
constexpr int SEG = BLOCK;
template <typename T, int W>
__global__ __launch_bounds__(BLOCK) void k_multi_gather(Groups<T> G, double *__restrict__ out, T *__restrict__ diag)
{
    __shared__ double sv[SEG * W];
    const GroupView<T> &gv = G.g[group_of_chunk(G, blockIdx.x)];
    const int c = int(blockIdx.x) - gv.chunk0;
    if (c >= gv.n_chunks) return;
    const int t = threadIdx.x;
    const int i0 = gv.chunk[c], i1 = gv.chunk[c + 1];
    const int64_t node0 = gv.dof0 / 3;
    if constexpr (W == 3) {
        // every row's problem finished: nothing is read or written (the PCG
        // reads no row of a finished problem; the few problems a chunk spans
        // are tested by every thread alike)
        if (gv.skip.flag) {
            bool all = true;
            for (int s = gv.skip.node_seg[i0]; s <= gv.skip.node_seg[i1 - 1]; ++s) all = all && gv.skip.flag[s] != 0.0;
            if (all) return;
        }
    }
    const int32_t sl = gv.slot[int64_t(c) * SEG + t];
    if (sl >= 0) {
        const T *l = gv.loc + int64_t(sl) * W;
#pragma unroll
        for (int k = 0; k < W; ++k) sv[t * W + k] = double(l[k]);
    }
    __syncthreads();
    for (int i = i0 + t; i < i1; i += BLOCK) {
        const int32_t mt = gv.rmeta[i];
        const int q0 = mt & 1023, q1 = q0 + (mt >> 10);
        double s[W];
#pragma unroll
        for (int k = 0; k < W; ++k) s[k] = 0;
        const bool zero = W == 3 && gv.skip.node(i);
        if (!zero)
            for (int k = q0; k < q1; ++k)
#pragma unroll
                for (int j = 0; j < W; ++j) s[j] += sv[k * W + j];
        if constexpr (W == 3) {
            double *o = out + 3 * (node0 + i);
            o[0] = s[0];
            o[1] = s[1];
            o[2] = s[2];
        } else {
            T *o = diag + 9 * (node0 + i);
            o[0] = T(s[0]), o[1] = T(s[3]), o[2] = T(s[4]);
            o[3] = T(s[3]), o[4] = T(s[1]), o[5] = T(s[5]);
            o[6] = T(s[4]), o[7] = T(s[5]), o[8] = T(s[2]);
        }
    }
}

template <typename T> Groups<T> views(Handle &h, int loc_width, bool chunks)
{
    Groups<T> G{};
    G.n = int(h.groups.size());
    int blocks = 0, chunk0 = 0;
    int64_t dof0 = 0, elem0 = 0;
    for (int k = 0; k < G.n; ++k) {
        Handle &s = *h.groups[size_t(k)];
        GroupView<T> &v = G.g[k];
        v.m = s.mesh<T>();
        v.variant = (s.order == 2 ? 2 : 0) + (s.material == MAT_LE ? 1 : 0);
        v.block0 = blocks;
        blocks += int(grid_for(s.n_elements));
        v.chunk0 = chunk0;
        v.n_chunks = s.n_row_chunks;
        chunk0 += s.n_row_chunks;
        v.dof0 = dof0;
        v.elem0 = elem0;
        s.loc.reserve(h.value_size() * size_t(s.n_elements) * s.n_local * size_t(loc_width));
        v.loc = s.loc.as<T>();
        v.chunk = s.row_chunk.as<int32_t>();
        v.slot = s.row_slot.as<int32_t>();
        v.rmeta = s.row_meta.as<int32_t>();
        if (h.skip.flag)
            v.skip = SegSkip{h.skip.elem_seg + elem0, h.skip.node_seg + dof0 / 3, h.skip.flag};
        dof0 += s.n_dofs;
        elem0 += s.n_elements;
    }
    (void)chunks;
    return G;
}

int element_blocks(const Handle &h)
{
    int b = 0;
    for (const auto &s : h.groups) b += int(grid_for(s->n_elements));
    return b;
}
int gather_blocks(const Handle &h)
{
    int b = 0;
    for (const auto &s : h.groups) b += s->n_row_chunks;
    return b;
}

template <typename F> void on_precision(const Handle &h, F &&f)
{
    if (h.precision == PFGPU_MIXED)
        f.template operator()<float>();
    else
        f.template operator()<double>();
}

} // namespace

void mixed_prepare(Handle &h)
{
    for (auto &s : h.groups) ensure_row_chunks(*s);
}

void mixed_energy(Handle &h, const double *x, double *energy, double *per_element)
{
    mixed_prepare(h);
    const int nb = element_blocks(h);
    h.partial.reserve(sizeof(double) * size_t(nb));
    on_precision(h, [&]<typename T>() {
        k_multi_energy<T><<<nb, BLOCK, 0, h.stream>>>(views<T>(h, 3, false), x, per_element, h.partial.as<double>());
    });
    PF_LAUNCH_CHECK();
    if (energy) sum_doubles(h, h.partial.as<double>(), nb, energy);
}

void mixed_gradient(Handle &h, const double *x, double *g, double *energy)
{
    mixed_prepare(h);
    if (energy) mixed_energy(h, x, energy, nullptr);
    on_precision(h, [&]<typename T>() {
        const Groups<T> G = views<T>(h, 3, true);
        k_multi_gradient<T><<<element_blocks(h), BLOCK, 0, h.stream>>>(G, x);
        k_multi_gather<T, 3><<<gather_blocks(h), BLOCK, 0, h.stream>>>(G, g, nullptr);
    });
    PF_LAUNCH_CHECK();
}

void mixed_hvp(Handle &h, const double *x, const double *v, double *y)
{
    mixed_prepare(h);
    on_precision(h, [&]<typename T>() {
        const Groups<T> G = views<T>(h, 3, true);
        k_multi_hvp<T><<<element_blocks(h), BLOCK, 0, h.stream>>>(G, x, v);
        k_multi_gather<T, 3><<<gather_blocks(h), BLOCK, 0, h.stream>>>(G, y, nullptr);
    });
    PF_LAUNCH_CHECK();
}

void mixed_node_diagonal(Handle &h, const double *x, void *diag)
{
    mixed_prepare(h);
    on_precision(h, [&]<typename T>() {
        const Groups<T> G = views<T>(h, 6, true);
        k_multi_diag<T><<<element_blocks(h), BLOCK, 0, h.stream>>>(G, x);
        k_multi_gather<T, 6><<<gather_blocks(h), BLOCK, 0, h.stream>>>(G, nullptr, static_cast<T *>(diag));
    });
    PF_LAUNCH_CHECK();
}

} // namespace pfgpu
