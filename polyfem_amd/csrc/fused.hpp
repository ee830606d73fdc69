// Patch tables of the fused assembly kernel (see fused_tables.cpp).
#pragma once

#include "elem.hpp"

namespace pfgpu {

// ---- launch shape of the fused kernel (shared with the host table builder
// checks): one persistent workgroup of NT threads per CU; the first
// producer_waves() wavefronts stage chunks, the others accumulate.
constexpr int NT = 1024;            // workgroup size
constexpr int NWAVE = NT / 64;
constexpr int LDS_LIMIT = 160 * 1024;
// Staging record of one element (value type T, 16-byte aligned sections):
//   K[q][b][3]   k_b = F^{-T} G_b at point q (J^{-T} G_b for linear elasticity)
//   C[2 NQ]      cl[q] = w lambda, cc[q] = w (mu - lambda ln J_q)
//   Q[q][d][3]   w P_q J^{-1}: gradient row of node a = sum_q sum_k RG(q,a,k) Q[q][d][k]
//   G[6]         w mu J^{-T}J^{-1} (00 11 22 01 02 12)
constexpr int align4(int v) { return (v + 3) & ~3; }
template <int ORDER, int NQ> struct Stage {
    static constexpr int NB = NLocal<ORDER>::value;
    static constexpr int K = 0;
    static constexpr int C = align4(K + NQ * NB * 3);
    static constexpr int Q = align4(C + 2 * NQ);
    static constexpr int G = align4(Q + 9 * NQ);
    static constexpr int SIZE = align4(G + 6);
};

// elements per chunk: one producer thread per (element, q), 4 producer waves
// in FP32 and 2 in FP64 (whose staging is twice as large)
template <typename T, int NQ> constexpr int chunk_elems() { return (sizeof(T) == 4 ? 256 : 128) / NQ; }
template <typename T, int NQ> constexpr int producer_waves() { return chunk_elems<T, NQ>() * NQ / 64; }
template <typename T, int ORDER> constexpr int item_cap() { return (ORDER == 2 ? 512 : 1024) / (sizeof(T) == 4 ? 1 : 2); }

// producer thread t -> (quadrature point, element in chunk); the point is
// uniform across a wavefront whenever the chunk is a multiple of 64 elements
template <typename T, int NQ> __host__ __device__ constexpr void producer_task(int t, int &q, int &kk)
{
    constexpr int CH = chunk_elems<T, NQ>();
    static_assert(CH * NQ == producer_waves<T, NQ>() * 64, "one producer thread per (element, point)");
    q = t / CH;
    kk = t - q * CH;
}
template <typename T, int ORDER, int NQ> constexpr int n_groups()
{
    return (NWAVE - producer_waves<T, NQ>()) * (64 / NLocal<ORDER>::value);
}


// Launch shape the tables are built for.
struct FusedShape {
    int groups;       // lane groups that accumulate rows (rows are dealt to groups)
    int chunk;        // elements staged in LDS at a time (at most)
    int item_cap;     // (row, element) items per chunk (at most)
    int acc_capacity; // accumulator entries (values of owned rows) per patch
    int max_rows;     // owned rows per patch
};

// desc fields per patch
enum { PD_ROW0, PD_NROWS, PD_EL0, PD_NEL, PD_CG0, PD_NCH, PD_ACC, PD_CHUNK_BASE, PD_CS0, PD_STRIDE = 12 };

struct FusedTables {
    FusedShape shape{};
    int n_patches = 0;
    int max_rows = 0;
    int max_acc = 0;
    int max_unit_items = 0;
    int64_t n_chunks = 0;         // chunks over all patches (one energy partial each)
    int64_t visits = 0;           // sum over patches of incident elements
    std::vector<int32_t> desc;    // PD_STRIDE ints per patch
    std::vector<int32_t> rows;    // owned node ids
    std::vector<int32_t> row_acc; // accumulator offset of each owned row
    std::vector<int32_t> elems;   // incident element ids; bit 31: the patch sums its energy
    std::vector<int32_t> chunk_start; // per patch n_chunks+1 element offsets (patch-local)
    std::vector<int32_t> cg;      // per patch, per (chunk, group): first item; + sentinel
    std::vector<int32_t> hdr;     // item: element-in-chunk | a << 10 | row << 14
    std::vector<uint8_t> pos;     // item x n_local: column-block position in the row
};

FusedTables build_fused_tables(const Pattern &pat, int n_local, const int32_t *elem_nodes, int n_elements,
                               const double *vertices, const int32_t *elem_vertices, const FusedShape &shape);

// Device view of the tables.
struct FusedDev {
    const int32_t *desc, *rows, *row_acc, *elems, *chunk_start, *cg, *hdr;
    const uint8_t *pos;
    const int32_t *brow_ptr;
};

} // namespace pfgpu
