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

} // namespace pfgpu
