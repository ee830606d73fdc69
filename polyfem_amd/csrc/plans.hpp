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
// of chunk c (at most gs units and gs rows); udst[u n_local + b] is the stage
// position, within u's chunk, of u's block (row, node_e(b)); bq0[blk] is the
// first stage position of pattern block blk within its chunk (a block's
// contributions are consecutive, in unit order).  False when a row has more
// than min(gs, 255) units.
struct UnitPlanHost {
    std::vector<int32_t> chunk_row;
    std::vector<uint16_t> udst, bq0;
};
bool make_unit_plan(const Pattern &pat, int n_local, const int32_t *elem_nodes, int gs, UnitPlanHost &out);

} // namespace pfgpu
