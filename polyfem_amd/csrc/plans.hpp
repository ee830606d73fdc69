// Host-side gather plans (plans.cpp), shared by the library and the CPU
// checks of the plans.
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace pfgpu {

// Row-unit plan of the factored gather for workgroups of gs threads: a unit
// is a node incidence (node_inc order: rows in node order, each row's
// elements in element order); chunk_row[c] .. chunk_row[c+1]-1 are the rows
// of chunk c (at most gs units and gs - 1 rows); udst[u n_local + b] is the
// stage position, within u's chunk, of u's block (row, node_e(b)); bq0[blk]
// is the first stage position of pattern block blk within its chunk (a
// block's contributions are consecutive, in unit order).
//
// The sums of a chunk's blocks are split into pieces: a block of n
// contributions is one piece when n <= UNIT_PIECE, else ceil(n / UNIT_PIECE)
// pieces of UNIT_PIECE consecutive contributions (the last shorter); a split
// block is the sum of its piece sums in piece order, each piece summed in
// unit order.  The pieces depend on n alone, so the summation order of a
// block does not depend on gs.  lanes[c (n_lanes + 1) + t] is the first
// piece of thread t of the n_lanes summing threads in chunk c (threads take
// consecutive runs of pieces, balanced by contributions; entry n_lanes is the
// end), packed by unit_lane(): the block
// (chunk-local), the piece within the block and the block's row
// (chunk-local).  A thread's run holds the first piece of at most two split
// blocks.  False when a row has more than min(gs, 255) units.
constexpr int UNIT_PIECE = 8;
constexpr uint32_t unit_lane(int block, int piece, int row)
{
    return uint32_t(block) | uint32_t(piece) << 12 | uint32_t(row) << 20;
}
struct UnitPlanHost {
    std::vector<int32_t> chunk_row;
    std::vector<uint16_t> udst, bq0;
    std::vector<uint32_t> lanes;
};
bool make_unit_plan(const Pattern &pat, int n_local, const int32_t *elem_nodes, int gs, int n_lanes,
                    UnitPlanHost &out);

// Chunks of consecutive outputs 0 .. n-1, output i gathering the
// contributions ptr[i] .. ptr[i+1]-1: a chunk holds at most cap
// contributions; c = chunk starts plus n.  False when one output alone has
// more than cap.
bool segment_chunks(const int32_t *ptr, int64_t n, int cap, std::vector<int32_t> &c);

// Row plan of the node-row gathers (k_gather_rows_seg): chunk[c] ..
// chunk[c+1]-1 are the rows of chunk c (at most ROW_SEG incidences);
// slot[c ROW_SEG + k] is the element-local vector index (node_inc code) of
// the chunk's k-th incidence, -1 past the last; rmeta[i] = q0 | count << 10,
// row i's first slot in its chunk and its incidence count.  False when a
// node has more than ROW_SEG incidences.
constexpr int ROW_SEG = 256;
struct RowPlanHost {
    std::vector<int32_t> chunk, slot, rmeta;
};
bool make_row_plan(const Pattern &pat, RowPlanHost &out);

// Per-contribution plan of the block and per-contribution factored gathers
// (k_gather_blocks_seg, k_gather_factored) for workgroups of gs threads:
// ptr[b] .. ptr[b+1]-1 are pattern block b's contributions (pattern order,
// n_blocks blocks); chunk[c] .. chunk[c+1]-1 are the blocks of chunk c (at
// most gs contributions), stored in the slots c gs .. (c+1) gs - 1 (-1 when
// unused) as words[k]; meta[b] = q0 | count << 10 | deg << 20 (first slot,
// count, the degree of b's row).  With keys (by element): each chunk's slots
// are ordered by keys[k] (ties in block order) and dst[c gs + i] is slot i's
// block-order position.  False when a block has more than gs contributions
// or a row more than 4095 blocks.
struct ContribPlanHost {
    std::vector<int32_t> chunk, slot, meta;
    std::vector<uint16_t> dst;
};
bool make_contrib_plan(const Pattern &pat, const std::vector<int32_t> &ptr, const std::vector<int32_t> &words,
                       const std::vector<int32_t> *keys, int gs, ContribPlanHost &out);

} // namespace pfgpu
