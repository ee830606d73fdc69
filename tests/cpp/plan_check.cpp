// CPU check of the host-built index tables (no GPU), meant to run under
// AddressSanitizer / UndefinedBehaviorSanitizer.
//
// Builds the pattern (pattern.cpp), the row-unit plans of the factored
// gather (64 and 256 units with one thread each, 64 units with two), the row
// plan of the node-row gathers and the per-contribution plan (plans.cpp) and
// the fused kernel's patch tables (fused_tables.cpp) exactly as the library
// does, then checks every index the kernels derive from them:
//  1. chunks cover the rows in order, with at most GS units and GS rows;
//  2. every unit's n_local stage positions lie inside the chunk's stage
//     (units x n_local <= GS x n_local slots) and inside the range of the
//     pattern block (row, node_e(b)) that bq0 gives; the positions of a
//     chunk are a bijection onto [0, units x n_local);
//  3. within a block, contributions are staged in unit (element) order, and
//     each block's count equals a brute-force count of the row's elements
//     that contain the column node;
//  4. the threads' runs of pieces, walked exactly as the kernel walks them,
//     read every staged contribution once, write every block of the chunk
//     once (whole, or through the piece sums of a split block: pieces of
//     UNIT_PIECE, in order), stay inside the stage, the chunk's blocks and
//     rows, and hold at most two first pieces of split blocks;
//  5. the row plan of the node-row gathers lists every row's incidences in
//     order, within ROW_SEG slots per chunk;
//  6. the per-contribution plan of the factored gather holds every block's
//     contributions once, in element order within a chunk, with the right
//     block metadata.
// Input: a binary mesh file (order, E, N, Nv, elem_nodes, elem_vertices,
// vertices) written by tests/test_plans.py.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../polyfem_amd/csrc/fused.hpp"
#include "../../polyfem_amd/csrc/plans.hpp"

using namespace pfgpu;

static long n_fail = 0;
#define CHECK(c, ...)                                                                                          \
    do {                                                                                                       \
        if (!(c)) {                                                                                            \
            if (n_fail < 20) {                                                                                 \
                std::fprintf(stderr, "CHECK failed %s:%d: ", __FILE__, __LINE__);                              \
                std::fprintf(stderr, __VA_ARGS__);                                                             \
                std::fprintf(stderr, "\n");                                                                    \
            }                                                                                                  \
            n_fail++;                                                                                          \
        }                                                                                                      \
    } while (0)

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: plan_check mesh.bin\n");
        return 2;
    }
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int32_t hdr[4];
    if (std::fread(hdr, 4, 4, f) != 4) return 2;
    const int order = hdr[0], E = hdr[1], N = hdr[2], Nv = hdr[3], NB = order == 1 ? 4 : 10;
    std::vector<int32_t> en(size_t(E) * NB), ev(size_t(E) * 4);
    std::vector<double> vx(size_t(Nv) * 3);
    if (std::fread(en.data(), 4, en.size(), f) != en.size() || std::fread(ev.data(), 4, ev.size(), f) != ev.size() ||
        std::fread(vx.data(), 8, vx.size(), f) != vx.size())
        return 2;
    std::fclose(f);

    const Pattern pat = build_pattern(E, NB, en.data(), N, Nv, ev.data());
    const auto &np = pat.node_ptr, &ni = pat.node_inc, &bp = pat.brow_ptr, &bc = pat.bcol;
    // (units per chunk, summing threads): one thread per unit, or two per
    // unit (each forming half of its blocks)
    const int configs[4][2] = {{64, 64}, {128, 128}, {256, 256}, {64, 128}};
    for (const auto &cfg : configs) {
        const int gs = cfg[0], nl = cfg[1];
        UnitPlanHost u;
        if (!make_unit_plan(pat, NB, en.data(), gs, nl, u)) {
            std::printf("gs %d: a row has too many units (the per-contribution gather serves)\n", gs);
            continue;
        }
        const int nc = int(u.chunk_row.size()) - 1;
        CHECK(u.chunk_row.front() == 0 && u.chunk_row.back() == N, "chunk rows do not cover the nodes");
        CHECK(u.udst.size() == size_t(np[N]) * NB && u.bq0.size() == size_t(pat.n_blocks) &&
                  u.lanes.size() == size_t(nc) * size_t(nl + 1),
              "plan sizes");
        std::vector<uint8_t> seen;
        for (int c = 0; c < nc; ++c) {
            const int i0 = u.chunk_row[c], i1 = u.chunk_row[c + 1];
            CHECK(i1 > i0, "empty chunk %d", c);
            const int u0 = np[i0], nu = np[i1] - u0, nr = i1 - i0;
            CHECK(nu <= gs && nr <= gs, "chunk %d: %d units, %d rows > %d", c, nu, nr, gs);
            const int B0 = bp[i0], nb = bp[i1] - B0, total = nu * NB;
            CHECK(nb <= total, "chunk %d: %d blocks > %d contributions", c, nb, total);
            if (nb > 0) CHECK(u.bq0[B0] == 0, "chunk %d: first block does not start at 0", c);
            for (int k = 0; k < nb; ++k) {
                const int q0 = u.bq0[B0 + k], q1 = k + 1 < nb ? u.bq0[B0 + k + 1] : total;
                CHECK(q0 <= q1 && q1 <= total, "chunk %d block %d: range [%d, %d) of %d", c, k, q0, q1, total);
            }
            seen.assign(size_t(total), 0);
            for (int i = i0; i < i1; ++i) {
                const int r0 = bp[i], deg = bp[i + 1] - r0;
                for (int t = 1; t < deg; ++t) CHECK(bc[r0 + t - 1] < bc[r0 + t], "row %d: columns not sorted", i);
                std::vector<int> last(size_t(deg), -1), count(size_t(deg), 0);
                for (int uu = np[i]; uu < np[i + 1]; ++uu) {
                    const int e = ni[uu] / NB, a = ni[uu] % NB;
                    CHECK(e >= 0 && e < E && en[size_t(e) * NB + a] == i, "unit %d is not an incidence of row %d", uu, i);
                    if (uu > np[i]) CHECK(ni[uu - 1] / NB < e, "row %d: units not in element order", i);
                    for (int b = 0; b < NB; ++b) {
                        const int j = en[size_t(e) * NB + b];
                        const int32_t *cb = bc.data() + r0;
                        const int p = int(std::lower_bound(cb, cb + deg, j) - cb);
                        const int k = p < deg && cb[p] == j ? p : -1;
                        CHECK(k >= 0, "row %d lacks column %d", i, j);
                        if (k < 0) continue;
                        const int d = u.udst[size_t(uu) * NB + b];
                        const int q0 = u.bq0[r0 + k], q1 = (r0 + k + 1 - B0) < nb ? u.bq0[r0 + k + 1] : total;
                        CHECK(d >= 0 && d < total, "unit %d block %d: stage %d outside [0, %d)", uu, b, d, total);
                        CHECK(d >= q0 && d < q1, "unit %d block %d: stage %d outside its block [%d, %d)", uu, b, d, q0, q1);
                        CHECK(d > last[size_t(k)], "block (%d, %d): not in unit order", i, j);
                        last[size_t(k)] = d;
                        count[size_t(k)]++;
                        if (d >= 0 && d < total) {
                            CHECK(!seen[size_t(d)], "chunk %d: stage %d used twice", c, d);
                            seen[size_t(d)] = 1;
                        }
                    }
                }
                for (int t = 0; t < deg; ++t) {
                    const int q0 = u.bq0[r0 + t], q1 = (r0 + t + 1 - B0) < nb ? u.bq0[r0 + t + 1] : total;
                    CHECK(q1 - q0 == count[size_t(t)], "block (%d, %d): count %d, staged %d", i, bc[r0 + t], q1 - q0,
                          count[size_t(t)]);
                }
            }
            for (int d = 0; d < total; ++d) CHECK(seen[size_t(d)], "chunk %d: stage %d never written", c, d);
            CHECK(nr <= gs - 1, "chunk %d: %d rows > %d", c, nr, gs - 1);
            // 4. the kernel's walk over the lane entries
            const uint32_t *ln = u.lanes.data() + size_t(c) * size_t(nl + 1);
            std::vector<int> read(size_t(total), 0), written(size_t(nb), 0), partial(size_t(total), 0);
            std::vector<int> srb(size_t(nr) + 1);
            for (int r = 0; r <= nr; ++r) srb[size_t(r)] = bp[i0 + r] - B0;
            auto bq = [&](int k) { return int(u.bq0[size_t(B0 + k)]); };
            auto bend = [&](int k) { return k + 1 < nb ? bq(k + 1) : total; };
            auto row_of = [&](int k) { return int(std::upper_bound(srb.begin(), srb.end(), k) - srb.begin()) - 1; };
            std::vector<std::pair<int, int>> heads;
            for (int t = 0; t < nl; ++t) {
                int k = int(ln[t] & 0xfffu), j = int((ln[t] >> 12) & 0xffu), rr = int(ln[t] >> 20);
                const int ke = int(ln[t + 1] & 0xfffu), je = int((ln[t + 1] >> 12) & 0xffu);
                if (!((k != ke || j != je) && k < nb)) continue;
                CHECK(rr >= 0 && rr < nr && rr == row_of(k), "chunk %d lane %d: row %d of block %d", c, t, rr, k);
                int q0 = bq(k), q1 = bend(k), rend = srb[size_t(rr) + 1], nh = 0, steps = 0;
                while (true) {
                    const int s = q0 + j * UNIT_PIECE, pe = std::min(q1, s + UNIT_PIECE);
                    CHECK(s >= 0 && pe <= total && (s < pe || (j == 0 && q0 == q1)), "chunk %d lane %d: piece [%d, %d)", c, t, s, pe);
                    for (int x = s; x < pe && x < total; ++x) read[size_t(x)]++;
                    const bool whole = pe == q1;
                    CHECK(rr == row_of(k), "chunk %d lane %d: block %d tracked in row %d", c, t, k, rr);
                    if (j == 0 && whole) {
                        written[size_t(k)]++;
                    } else {
                        if (s < total) partial[size_t(s)]++;
                        if (j == 0) {
                            heads.emplace_back(k, rr);
                            ++nh;
                        }
                    }
                    if (whole) {
                        ++k;
                        j = 0;
                    } else {
                        ++j;
                    }
                    if ((k == ke && j == je) || k >= nb || ++steps > 100000) break;
                    if (j == 0) {
                        while (k >= rend) rend = srb[size_t(++rr) + 1];
                        q0 = q1;
                        q1 = bend(k);
                    }
                }
                CHECK(nh <= 2, "chunk %d lane %d: %d split-block heads", c, t, nh);
            }
            for (const auto &hd : heads) {
                const int k = hd.first, q0 = bq(k), q1 = bend(k);
                CHECK(q1 - q0 > UNIT_PIECE, "chunk %d: block %d of %d contributions split", c, k, q1 - q0);
                for (int s = q0; s < q1; s += UNIT_PIECE) CHECK(partial[size_t(s)] == 1, "chunk %d: piece %d of block %d", c, s, k);
                written[size_t(k)]++;
            }
            for (int d = 0; d < total; ++d) CHECK(read[size_t(d)] == 1, "chunk %d: stage %d read %d times", c, d, read[size_t(d)]);
            for (int k = 0; k < nb; ++k) CHECK(written[size_t(k)] == 1, "chunk %d: block %d written %d times", c, k, written[size_t(k)]);
        }
        std::printf("gs %d lanes %d: %d chunks checked\n", gs, nl, nc);
    }
    // 5. the row plan of the node-row gathers
    {
        RowPlanHost r;
        CHECK(make_row_plan(pat, r), "row plan refused");
        const int nc = int(r.chunk.size()) - 1;
        CHECK(r.chunk.front() == 0 && r.chunk.back() == N, "row chunks do not cover the nodes");
        for (int c = 0; c < nc; ++c) {
            const int i0 = r.chunk[c], i1 = r.chunk[c + 1], k0 = np[i0], tot = np[i1] - k0;
            CHECK(i1 > i0 && tot <= ROW_SEG, "row chunk %d: rows [%d, %d), %d incidences", c, i0, i1, tot);
            for (int i = i0; i < i1; ++i) {
                const int q0 = r.rmeta[i] & 1023, cnt = r.rmeta[i] >> 10;
                CHECK(q0 == np[i] - k0 && cnt == np[i + 1] - np[i], "row %d: meta %d/%d", i, q0, cnt);
                for (int j = 0; j < cnt; ++j)
                    CHECK(r.slot[size_t(c) * ROW_SEG + q0 + j] == ni[np[i] + j], "row %d incidence %d", i, j);
            }
            for (int k = tot; k < ROW_SEG; ++k) CHECK(r.slot[size_t(c) * ROW_SEG + k] == -1, "row chunk %d slot %d used", c, k);
        }
        std::printf("row plan: %d chunks checked\n", nc);
    }
    // 6. the per-contribution plan of the factored gather (element-ordered slots)
    {
        std::vector<int64_t> ptr64;
        std::vector<int32_t> codes, block_row;
        build_block_contributions(pat, NB, en.data(), ptr64, codes, block_row);
        const std::vector<int32_t> ptr(ptr64.begin(), ptr64.end());
        std::vector<int32_t> keys(codes.size());
        for (size_t k = 0; k < codes.size(); ++k) keys[k] = codes[k] / (NB * NB);
        for (int gs : {128, 256, 512}) {
            ContribPlanHost cp;
            if (!make_contrib_plan(pat, ptr, codes, &keys, gs, cp)) {
                std::printf("contribution plan gs %d refused\n", gs);
                continue;
            }
            const int nc = int(cp.chunk.size()) - 1;
            CHECK(cp.chunk.front() == 0 && cp.chunk.back() == pat.n_blocks, "contribution chunks do not cover the blocks");
            for (int c = 0; c < nc; ++c) {
                const int b0 = cp.chunk[c], b1 = cp.chunk[c + 1], k0 = ptr[b0], n = ptr[b1] - k0;
                CHECK(b1 > b0 && n <= gs, "contribution chunk %d: %d contributions > %d", c, n, gs);
                std::vector<int> hit(size_t(n), 0);
                for (int i = 0; i < n; ++i) {
                    const int o = cp.dst[size_t(c) * gs + i];
                    CHECK(o >= 0 && o < n, "chunk %d slot %d: position %d", c, i, o);
                    if (o < 0 || o >= n) continue;
                    hit[size_t(o)]++;
                    CHECK(cp.slot[size_t(c) * gs + i] == codes[size_t(k0 + o)], "chunk %d slot %d: word", c, i);
                    if (i > 0) {
                        const int po = cp.dst[size_t(c) * gs + i - 1];
                        const int ka = keys[size_t(k0 + po)], kb = keys[size_t(k0 + o)];
                        CHECK(ka < kb || (ka == kb && po < o), "chunk %d slots %d/%d not in element order", c, i - 1, i);
                    }
                }
                for (int o = 0; o < n; ++o) CHECK(hit[size_t(o)] == 1, "chunk %d: position %d taken %d times", c, o, hit[size_t(o)]);
                for (int k = n; k < gs; ++k) CHECK(cp.slot[size_t(c) * gs + k] == -1, "chunk %d slot %d used", c, k);
                for (int b = b0; b < b1; ++b) {
                    const int q0 = cp.meta[b] & 1023, cnt = (cp.meta[b] >> 10) & 1023, deg = int(uint32_t(cp.meta[b]) >> 20);
                    const int row = block_row[size_t(b)];
                    CHECK(q0 == ptr[b] - k0 && cnt == ptr[b + 1] - ptr[b] && deg == bp[row + 1] - bp[row],
                          "block %d: meta %d/%d/%d", b, q0, cnt, deg);
                }
            }
            std::printf("contribution plan gs %d: %d chunks checked\n", gs, nc);
        }
    }
    // the fused kernel's tables, built for the FP32 P1/P2 launch shape (the
    // sanitizers watch the builder; fused_check verifies their content)
    {
        FusedShape s{};
        s.groups = order == 1 ? 70 : 28;
        s.chunk = order == 1 ? 128 : 32;
        s.item_cap = s.chunk * NB;
        s.max_rows = 256;
        s.acc_capacity = 30000;
        const FusedTables t = build_fused_tables(pat, NB, en.data(), E, vx.data(), ev.data(), s);
        CHECK(t.n_patches > 0, "no patches");
        std::printf("fused tables: %d patches\n", t.n_patches);
    }
    std::printf("%s (%ld failures)\n", n_fail ? "FAIL" : "OK", n_fail);
    return n_fail ? 1 : 0;
}
