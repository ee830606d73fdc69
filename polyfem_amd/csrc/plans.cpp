// Host-side construction of the row-unit gather plan of the factored
// Hessian route (the kernel is k_gather_units in plain.hip).  Host code
// only, so that tests can build and check the plans on the CPU.
#include <algorithm>
#include <thread>

#include "plans.hpp"

namespace pfgpu {

bool make_unit_plan(const Pattern &pat, int n_local, const int32_t *elem_nodes, int gs, UnitPlanHost &out)
{
    const std::vector<int32_t> &np = pat.node_ptr, &ni = pat.node_inc, &bp = pat.brow_ptr, &bc = pat.bcol;
    const int NB = n_local, N = pat.n_nodes;
    // chunks of whole rows: at most gs units and gs rows each
    std::vector<int32_t> c{0};
    int units = 0, rows = 0;
    for (int i = 0; i < N; ++i) {
        const int ui = np[size_t(i) + 1] - np[size_t(i)];
        if (ui > std::min(gs, 255)) return false;
        if (units + ui > gs || rows + 1 > gs) {
            c.push_back(i);
            units = 0;
            rows = 0;
        }
        units += ui;
        rows += 1;
    }
    c.push_back(N);
    const int nc = int(c.size()) - 1;
    std::vector<uint16_t> q0(size_t(pat.n_blocks)), dst(size_t(np[size_t(N)]) * NB);
    // per chunk: every block's contributions are consecutive stage positions
    // (blocks in pattern order, contributions in unit order); a unit's block
    // b goes to the next free position of block (row, node_e(b))
    auto fill = [&](int c0, int c1) {
        std::vector<int32_t> cnt;
        for (int ci = c0; ci < c1; ++ci) {
            int run = 0;
            for (int i = c[size_t(ci)]; i < c[size_t(ci) + 1]; ++i) {
                const int32_t r0 = bp[size_t(i)], deg = bp[size_t(i) + 1] - r0;
                const int32_t *cols = bc.data() + r0;
                auto pos = [&](int32_t j) { return int32_t(std::lower_bound(cols, cols + deg, j) - cols); };
                cnt.assign(size_t(deg), 0);
                for (int32_t u = np[size_t(i)]; u < np[size_t(i) + 1]; ++u) {
                    const int32_t e = ni[size_t(u)] / NB;
                    for (int b = 0; b < NB; ++b) cnt[size_t(pos(elem_nodes[size_t(e) * NB + b]))]++;
                }
                for (int32_t k = 0; k < deg; ++k) {
                    q0[size_t(r0 + k)] = uint16_t(run);
                    run += cnt[size_t(k)];
                    cnt[size_t(k)] = 0;
                }
                for (int32_t u = np[size_t(i)]; u < np[size_t(i) + 1]; ++u) {
                    const int32_t e = ni[size_t(u)] / NB;
                    for (int b = 0; b < NB; ++b) {
                        const int32_t k = pos(elem_nodes[size_t(e) * NB + b]);
                        dst[size_t(u) * NB + b] = uint16_t(q0[size_t(r0 + k)] + cnt[size_t(k)]++);
                    }
                }
            }
        }
    };
    const int nt = int(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t) pool.emplace_back(fill, int(int64_t(nc) * t / nt), int(int64_t(nc) * (t + 1) / nt));
    for (std::thread &th : pool) th.join();
    out.chunk_row = std::move(c);
    out.udst = std::move(dst);
    out.bq0 = std::move(q0);
    return true;
}

} // namespace pfgpu
