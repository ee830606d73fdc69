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
// elements per chunk: one producer thread per (element, quadrature point),
// 2 producer waves in FP32 and 1 in FP64 (whose staging is twice as large)
template <typename T, int NQ> constexpr int chunk_elems() { return (sizeof(T) == 4 ? 128 : 64) / NQ; }
// Producer thread t works on quadrature point t / chunk, so a wavefront holds
// 64 / chunk points of each element (one when the chunk has 64 elements or
// more).  The element gradient is summed over the points of one wavefront
// with cross-lane shuffles, leaving grad_parts() partial sums per element.
template <typename T, int NQ> constexpr int grad_parts()
{
    return chunk_elems<T, NQ>() >= 64 ? NQ : NQ * chunk_elems<T, NQ>() / 64;
}

// Staging record of one element (value type T; sections 16-byte aligned, and
// every 3-vector padded to 4 so that it is one LDS vector load):
//   K[q][b][4]   k_b = F^{-T} G_b at point q (J^{-T} G_b for linear elasticity)
//   C[q][2]      (w lambda, w (mu - lambda ln J_q))
//   R[h][a][4]   partial sums over points of the element gradient row of node a
//   G[8]         w mu J^{-T}J^{-1} (00 11 22 01 02 12), 2 pad
// The record length is 4 (mod 8) words so that the records of consecutive
// elements start on different LDS banks.
constexpr int align4(int v) { return (v + 3) & ~3; }
template <int ORDER, int NQ, int NPART> struct StageT {
    static constexpr int NB = NLocal<ORDER>::value;
    static constexpr int K = 0;
    static constexpr int C = K + NQ * NB * 4;
    static constexpr int R = align4(C + 2 * NQ);
    static constexpr int G = R + NPART * NB * 4;
    static constexpr int SIZE = (G + 8) % 8 == 4 ? G + 8 : G + 12;
};
template <typename T, int ORDER, int NQ> using Stage = StageT<ORDER, NQ, grad_parts<T, NQ>()>;
template <typename T, int NQ> constexpr int producer_waves() { return chunk_elems<T, NQ>() * NQ / 64; }
// (row, element) items per chunk: every element node may be an owned row
template <typename T, int ORDER, int NQ> constexpr int item_cap() { return chunk_elems<T, NQ>() * NLocal<ORDER>::value; }
// consumer lane group: one lane per (column block b, block row m) of an item
template <int ORDER> constexpr int group_lanes() { return 3 * NLocal<ORDER>::value; }
template <int ORDER> constexpr int groups_per_wave() { return 64 / group_lanes<ORDER>(); }
// bytes of column positions stored per item (n_local used, padded for vector copies)
__host__ __device__ constexpr int pos_stride(int n_local) { return n_local <= 4 ? 4 : 16; }

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
    return (NWAVE - producer_waves<T, NQ>()) * groups_per_wave<ORDER>();
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
    std::vector<int32_t> vnodes;  // per listed element: its n_local node ids
    std::vector<int32_t> chunk_start; // per patch n_chunks+1 element offsets (patch-local); the
                                      // elements of a chunk are contiguous in elems
    std::vector<int32_t> cg;      // per patch, per (chunk, group): first item; + sentinel
    std::vector<int32_t> hdr;     // item: element-in-chunk | a << 10 | row << 14
    std::vector<uint8_t> pos;     // item x pos_stride(n_local): column-block position in the row
};

FusedTables build_fused_tables(const Pattern &pat, int n_local, const int32_t *elem_nodes, int n_elements,
                               const double *vertices, const int32_t *elem_vertices, const FusedShape &shape);

// Device view of the tables.
struct FusedDev {
    const int32_t *desc, *rows, *row_acc, *elems, *vnodes, *chunk_start, *cg, *hdr;
    const uint8_t *pos;
    const int32_t *brow_ptr;
};

} // namespace pfgpu
