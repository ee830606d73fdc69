// Host-side construction of the row-unit gather plan of the factored
// Hessian route (the kernel is k_gather_units in plain.hip).  Host code
// only, so that tests can build and check the plans on the CPU.
#include <algorithm>
#include <atomic>
#include <thread>

#include "plans.hpp"

namespace pfgpu {

namespace {

// a piece of a chunk's block sums: block (chunk-local), piece within the
// block, the block's row (chunk-local), contributions, first piece of a
// split block
struct Piece {
    int32_t block, piece, row, len;
    bool head;
};
// the cost of a piece in contributions read: its reads plus the storing of
// its sum (one block, to the values or to the stage)
constexpr int PIECE_COST = 3;

// greedy runs of consecutive pieces of at most bound cost and at most two
// heads each: the number of runs (and their first pieces)
int piece_runs(const std::vector<Piece> &p, int bound, std::vector<int> *first)
{
    int n = 0, cur = 0, heads = 0;
    for (size_t i = 0; i < p.size(); ++i) {
        const int c = p[i].len + PIECE_COST;
        if (n == 0 || cur + c > bound || (p[i].head && heads == 2)) {
            ++n;
            cur = 0;
            heads = 0;
            if (first) first->push_back(int(i));
        }
        cur += c;
        heads += p[i].head ? 1 : 0;
    }
    return n;
}

// the threads' runs of pieces in one chunk: the least maximal run cost for
// which gs threads suffice, written as gs + 1 lane entries
void chunk_lanes(const std::vector<Piece> &p, int nb, int nr, int gs, uint32_t *lanes)
{
    int lo = 0, hi = 0;
    for (const Piece &x : p) {
        lo = std::max(lo, x.len + PIECE_COST);
        hi += x.len + PIECE_COST;
    }
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (piece_runs(p, mid, nullptr) <= gs)
            hi = mid;
        else
            lo = mid + 1;
    }
    std::vector<int> first;
    piece_runs(p, lo, &first);
    for (int t = 0; t <= gs; ++t) {
        if (t < int(first.size())) {
            const Piece &x = p[size_t(first[size_t(t)])];
            lanes[t] = unit_lane(x.block, x.piece, x.row);
        } else {
            lanes[t] = unit_lane(nb, 0, nr);
        }
    }
}

} // namespace

bool make_unit_plan(const Pattern &pat, int n_local, const int32_t *elem_nodes, int gs, int n_lanes,
                    UnitPlanHost &out)
{
    const std::vector<int32_t> &np = pat.node_ptr, &ni = pat.node_inc, &bp = pat.brow_ptr, &bc = pat.bcol;
    const int NB = n_local, N = pat.n_nodes;
    // chunks of whole rows: at most gs units and gs - 1 rows each (a thread
    // per row boundary)
    std::vector<int32_t> c{0};
    int units = 0, rows = 0;
    for (int i = 0; i < N; ++i) {
        const int ui = np[size_t(i) + 1] - np[size_t(i)];
        if (ui > std::min(gs, 255)) return false;
        if (units + ui > gs || rows + 1 > gs - 1) {
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
    std::vector<uint32_t> lanes(size_t(nc) * size_t(n_lanes + 1));
    std::atomic<bool> fits{true};
    // per chunk: every block's contributions are consecutive stage positions
    // (blocks in pattern order, contributions in unit order); a unit's block
    // b goes to the next free position of block (row, node_e(b))
    auto fill = [&](int c0, int c1) {
        std::vector<int32_t> cnt;
        std::vector<Piece> pieces;
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
            // the pieces of the chunk's block sums, in block order
            const int i0 = c[size_t(ci)], i1 = c[size_t(ci) + 1], B0 = bp[size_t(i0)], nb = bp[size_t(i1)] - B0;
            if (nb > 4095 || i1 - i0 > 4095) {
                fits = false;
                continue;
            }
            pieces.clear();
            for (int i = i0; i < i1; ++i)
                for (int32_t blk = bp[size_t(i)]; blk < bp[size_t(i) + 1]; ++blk) {
                    const int k = blk - B0, start = q0[size_t(blk)];
                    const int n = (k + 1 < nb ? int(q0[size_t(blk) + 1]) : run) - start;
                    const int pc = n <= UNIT_PIECE ? 1 : (n + UNIT_PIECE - 1) / UNIT_PIECE;
                    for (int j = 0; j < pc; ++j)
                        pieces.push_back(Piece{k, j, i - i0, std::min(UNIT_PIECE, n - j * UNIT_PIECE), j == 0 && pc > 1});
                }
            chunk_lanes(pieces, nb, i1 - i0, n_lanes, lanes.data() + size_t(ci) * size_t(n_lanes + 1));
        }
    };
    const int nt = int(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t) pool.emplace_back(fill, int(int64_t(nc) * t / nt), int(int64_t(nc) * (t + 1) / nt));
    for (std::thread &th : pool) th.join();
    if (!fits) return false;
    out.chunk_row = std::move(c);
    out.udst = std::move(dst);
    out.bq0 = std::move(q0);
    out.lanes = std::move(lanes);
    return true;
}

bool segment_chunks(const int32_t *ptr, int64_t n, int cap, std::vector<int32_t> &c)
{
    c.assign(1, 0);
    int64_t k0 = ptr[0];
    for (int64_t i = 0; i < n; ++i) {
        if (ptr[i + 1] - ptr[i] > cap) return false;
        if (ptr[i + 1] - k0 > cap) {
            c.push_back(int32_t(i));
            k0 = ptr[i];
        }
    }
    c.push_back(int32_t(n));
    return true;
}

bool make_row_plan(const Pattern &pat, RowPlanHost &out)
{
    const std::vector<int32_t> &ptr = pat.node_ptr, &inc = pat.node_inc;
    std::vector<int32_t> c;
    if (!segment_chunks(ptr.data(), pat.n_nodes, ROW_SEG, c)) return false;
    const int nc = int(c.size()) - 1;
    std::vector<int32_t> slot(size_t(nc) * ROW_SEG, -1), rmeta(size_t(pat.n_nodes));
    for (int ci = 0; ci < nc; ++ci) {
        const int32_t k0 = ptr[size_t(c[size_t(ci)])];
        for (int32_t i = c[size_t(ci)]; i < c[size_t(ci) + 1]; ++i) {
            rmeta[size_t(i)] = (ptr[size_t(i)] - k0) | ((ptr[size_t(i) + 1] - ptr[size_t(i)]) << 10);
            for (int32_t k = ptr[size_t(i)]; k < ptr[size_t(i) + 1]; ++k)
                slot[size_t(ci) * ROW_SEG + size_t(k - k0)] = inc[size_t(k)];
        }
    }
    out.chunk = std::move(c);
    out.slot = std::move(slot);
    out.rmeta = std::move(rmeta);
    return true;
}

bool make_contrib_plan(const Pattern &pat, const std::vector<int32_t> &ptr, const std::vector<int32_t> &words,
                       const std::vector<int32_t> *keys, int gs, ContribPlanHost &out)
{
    std::vector<int32_t> c;
    if (!segment_chunks(ptr.data(), pat.n_blocks, gs, c)) return false;
    const int nc = int(c.size()) - 1;
    std::vector<int32_t> slot(size_t(nc) * gs, -1), meta(size_t(pat.n_blocks));
    for (int i = 0; i < pat.n_nodes; ++i) {
        const int32_t deg = pat.brow_ptr[size_t(i) + 1] - pat.brow_ptr[size_t(i)];
        if (deg > 4095) return false;
        for (int32_t b = pat.brow_ptr[size_t(i)]; b < pat.brow_ptr[size_t(i) + 1]; ++b) meta[size_t(b)] = deg << 20;
    }
    std::vector<uint16_t> dst(keys ? size_t(nc) * gs : 0, 0);
    // chunks are independent: fill them on the host's threads
    auto fill = [&](int c0, int c1) {
        std::vector<int64_t> kv;
        std::vector<int32_t> w;
        for (int ci = c0; ci < c1; ++ci) {
            const int32_t k0 = ptr[size_t(c[size_t(ci)])], n = ptr[size_t(c[size_t(ci) + 1])] - k0;
            int32_t *sl = slot.data() + size_t(ci) * gs;
            for (int32_t b = c[size_t(ci)]; b < c[size_t(ci) + 1]; ++b) {
                meta[size_t(b)] |= (ptr[size_t(b)] - k0) | ((ptr[size_t(b) + 1] - ptr[size_t(b)]) << 10);
                for (int32_t k = ptr[size_t(b)]; k < ptr[size_t(b) + 1]; ++k) sl[k - k0] = words[size_t(k)];
            }
            if (!keys) continue;
            // the chunk's slots ordered by key (element), ties in block order,
            // with the block-order position of every contribution
            kv.resize(size_t(n));
            w.assign(sl, sl + n);
            for (int32_t i = 0; i < n; ++i) kv[size_t(i)] = (int64_t((*keys)[size_t(k0 + i)]) << 16) | i;
            std::sort(kv.begin(), kv.end());
            for (int32_t i = 0; i < n; ++i) {
                const int32_t o = int32_t(kv[size_t(i)] & 0xffff);
                sl[i] = w[size_t(o)];
                dst[size_t(ci) * gs + size_t(i)] = uint16_t(o);
            }
        }
    };
    const int nt = int(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t) pool.emplace_back(fill, int(int64_t(nc) * t / nt), int(int64_t(nc) * (t + 1) / nt));
    for (std::thread &th : pool) th.join();
    out.chunk = std::move(c);
    out.slot = std::move(slot);
    out.meta = std::move(meta);
    out.dst = std::move(dst);
    return true;
}

} // namespace pfgpu
