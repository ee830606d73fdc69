// Host-side construction of the patch tables of the fused assembly kernels.
//
// A patch is a set of nodes whose Hessian rows (and gradient entries) one
// workgroup owns.  The workgroup visits every element incident to one of its
// nodes (elements on the patch border are visited by several workgroups) and
// accumulates the owned rows in LDS; no other workgroup writes those rows, so
// the result needs neither atomics nor a second pass.
//
// Nodes are grouped in Morton order of their positions so that patches are
// compact and the border (re-visited) elements few.  Inside a patch the rows
// are dealt to lane groups (one group = 3 n_local lanes, one lane per row of
// each 3x3 column block of an element) so that each group owns whole rows of a
// chunk; a group processes its (row, element) items in a fixed order, which
// fixes the summation order.
#include <algorithm>
#include <array>
#include <numeric>
#include <thread>

#include "fused.hpp"

namespace pfgpu {

namespace {

uint64_t spread3(uint64_t v)
{
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}

// reference coordinates of the P2 nodes (autogen/auto_p_bases.cpp:1437-1448)
const double P2_NODES[10][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {.5, 0, 0},
                                {.5, .5, 0}, {0, .5, 0}, {0, 0, .5}, {.5, 0, .5}, {0, .5, .5}};

} // namespace

FusedTables build_fused_tables(const Pattern &pat, int n_local, const int32_t *elem_nodes, int n_elements,
                               const double *vertices, const int32_t *elem_vertices, const FusedShape &shape)
{
    const int N = pat.n_nodes;
    FusedTables t;
    t.shape = shape;

    // node positions from their first incident element
    std::vector<double> X(size_t(N) * 3, 0.0);
    std::vector<char> seen(N, 0);
    for (int e = 0; e < n_elements; ++e)
        for (int a = 0; a < n_local; ++a) {
            const int node = elem_nodes[size_t(e) * n_local + a];
            if (seen[node]) continue;
            seen[node] = 1;
            const double *xi = P2_NODES[a];
            const double w[4] = {1 - xi[0] - xi[1] - xi[2], xi[0], xi[1], xi[2]};
            for (int j = 0; j < 4; ++j) {
                const double *V = vertices + 3 * size_t(elem_vertices[size_t(e) * 4 + j]);
                for (int c = 0; c < 3; ++c) X[size_t(node) * 3 + c] += w[j] * V[c];
            }
        }
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < N; ++i)
        for (int c = 0; c < 3; ++c) {
            lo[c] = std::min(lo[c], X[size_t(i) * 3 + c]);
            hi[c] = std::max(hi[c], X[size_t(i) * 3 + c]);
        }
    std::vector<uint64_t> code(N);
    for (int i = 0; i < N; ++i) {
        uint64_t q[3];
        for (int c = 0; c < 3; ++c) {
            const double s = hi[c] > lo[c] ? (X[size_t(i) * 3 + c] - lo[c]) / (hi[c] - lo[c]) : 0.0;
            q[c] = uint64_t(std::min(2097151.0, std::max(0.0, s * 2097151.0)));
        }
        code[i] = spread3(q[0]) | spread3(q[1]) << 1 | spread3(q[2]) << 2;
    }
    std::vector<int32_t> order(N);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return code[a] < code[b]; });

    // greedy patches under the accumulator budget
    std::vector<int32_t> patch_of(N);
    std::vector<int32_t> patch_start{0};
    {
        int64_t used = 0;
        int rows = 0;
        for (int k = 0; k < N; ++k) {
            const int node = order[k];
            const int64_t need = 9 * int64_t(pat.brow_ptr[node + 1] - pat.brow_ptr[node]);
            if (rows > 0 && (used + need > shape.acc_capacity || rows + 1 > shape.max_rows)) {
                patch_start.push_back(k);
                used = 0;
                rows = 0;
            }
            if (need > shape.acc_capacity) fail(PFGPU_ERR_UNSUPPORTED, "a Hessian row does not fit the LDS accumulator");
            used += need;
            rows++;
            patch_of[node] = int32_t(patch_start.size() - 1);
        }
        patch_start.push_back(N);
    }
    const int P = int(patch_start.size()) - 1;
    t.n_patches = P;

    // per-patch construction, independent across patches
    struct Local {
        std::vector<int32_t> rows, row_acc, elems, chunk_start, cg, hdr;
        std::vector<uint8_t> pos;
        int n_chunks = 0, acc = 0, max_unit_items = 0;
    };
    std::vector<Local> loc(P);
    const int G = shape.groups;
    const int CH = shape.chunk;
    const int PS = pos_stride(n_local);
    auto build_one = [&](int p) {
        Local &L = loc[p];
        const int k0 = patch_start[p], k1 = patch_start[p + 1];
        L.rows.assign(order.begin() + k0, order.begin() + k1);
        std::sort(L.rows.begin(), L.rows.end());
        const int R = int(L.rows.size());
        // accumulator offsets: rows concatenated, each laid out as its global segment
        L.row_acc.resize(R);
        int acc = 0;
        for (int r = 0; r < R; ++r) {
            L.row_acc[r] = acc;
            acc += 9 * (pat.brow_ptr[L.rows[r] + 1] - pat.brow_ptr[L.rows[r]]);
        }
        L.acc = acc;
        // incident elements, in element order
        std::vector<int32_t> el;
        for (int r = 0; r < R; ++r)
            for (int k = pat.node_ptr[L.rows[r]]; k < pat.node_ptr[L.rows[r] + 1]; ++k) el.push_back(pat.node_inc[k] / n_local);
        std::sort(el.begin(), el.end());
        el.erase(std::unique(el.begin(), el.end()), el.end());
        const int nel = int(el.size());
        // owned rows of each element
        std::vector<std::vector<int>> erows(nel);
        for (int r = 0; r < R; ++r)
            for (int k = pat.node_ptr[L.rows[r]]; k < pat.node_ptr[L.rows[r] + 1]; ++k) {
                const int e = pat.node_inc[k] / n_local;
                erows[std::lower_bound(el.begin(), el.end(), e) - el.begin()].push_back(r);
            }
        int n_items = 0;
        for (const auto &v : erows) n_items += int(v.size());
        // Chunks: at most CH elements and item_cap items each.  A row's items
        // in one chunk run serially on one lane group, so elements are dealt
        // to chunks (in element order, each to the chunk where its busiest row
        // has the fewest items so far, then the fewest items overall) to
        // spread every row's items over the chunks.
        std::vector<int> chunk_of(nel);
        int nch = std::max((nel + CH - 1) / CH, (n_items + shape.item_cap - 1) / shape.item_cap);
        nch = std::max(nch, 1);
        for (;; ++nch) {
            std::vector<int> load(size_t(nch) * R, 0), cnt(nch, 0), itc(nch, 0);
            bool ok = true;
            for (int k = 0; k < nel && ok; ++k) {
                int best = -1, bkey0 = 0, bkey1 = 0;
                for (int c = 0; c < nch; ++c) {
                    if (cnt[c] >= CH || itc[c] + int(erows[k].size()) > shape.item_cap) continue;
                    int key0 = 0;
                    for (int r : erows[k]) key0 = std::max(key0, load[size_t(c) * R + r]);
                    if (best < 0 || key0 < bkey0 || (key0 == bkey0 && itc[c] < bkey1)) {
                        best = c;
                        bkey0 = key0;
                        bkey1 = itc[c];
                    }
                }
                if (best < 0) {
                    ok = false;
                    break;
                }
                chunk_of[k] = best;
                cnt[best]++;
                itc[best] += int(erows[k].size());
                for (int r : erows[k]) load[size_t(best) * R + r]++;
            }
            if (ok) {
                for (int c = 0; c < nch; ++c) L.max_unit_items = std::max(L.max_unit_items, itc[c]);
                break;
            }
        }
        // drop empty chunks, then lay the elements out chunk by chunk
        {
            std::vector<int> used(nch, 0), remap(nch, -1);
            for (int k = 0; k < nel; ++k) used[chunk_of[k]] = 1;
            int n = 0;
            for (int c = 0; c < nch; ++c)
                if (used[c]) remap[c] = n++;
            for (int k = 0; k < nel; ++k) chunk_of[k] = remap[chunk_of[k]];
            nch = n;
        }
        std::vector<int> perm(nel);
        std::iota(perm.begin(), perm.end(), 0);
        std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) { return chunk_of[x] < chunk_of[y]; });
        std::vector<int> slot_of(nel); // element index in the sorted list -> position in the chunked list
        for (int k = 0; k < nel; ++k) slot_of[perm[k]] = k;
        L.elems.resize(nel);
        L.chunk_start.assign(nch + 1, 0);
        for (int k = 0; k < nel; ++k) {
            const int e = el[perm[k]];
            // the element's energy is summed by the patch owning its first node
            const bool own = patch_of[elem_nodes[size_t(e) * n_local]] == p;
            L.elems[k] = e | (own ? int32_t(0x80000000u) : 0);
            L.chunk_start[chunk_of[perm[k]] + 1]++;
        }
        for (int c = 0; c < nch; ++c) L.chunk_start[c + 1] += L.chunk_start[c];
        L.n_chunks = nch;
        // items (row, element), each tagged with its chunk
        struct Item { int chunk, group, elem_k, a, r; };
        std::vector<Item> items;
        items.reserve(n_items);
        for (int r = 0; r < R; ++r) {
            const int node = L.rows[r];
            for (int k = pat.node_ptr[node]; k < pat.node_ptr[node + 1]; ++k) {
                const int e = pat.node_inc[k] / n_local, a = pat.node_inc[k] % n_local;
                const int ek = int(std::lower_bound(el.begin(), el.end(), e) - el.begin());
                items.push_back({chunk_of[ek], -1, slot_of[ek], a, r});
            }
        }
        // Per chunk, rows -> groups: a row needs one owner only within a chunk
        // (chunks are separated by barriers).  Longest row first onto the least
        // loaded group, then groups renumbered by decreasing load so that the
        // groups sharing a wavefront carry similar loads.
        {
            std::vector<std::vector<int>> row_items_in_chunk(L.n_chunks, std::vector<int>(R, 0));
            for (const Item &it : items) row_items_in_chunk[it.chunk][it.r]++;
            std::vector<std::vector<int>> group_of(L.n_chunks, std::vector<int>(R, -1));
            for (int c = 0; c < L.n_chunks; ++c) {
                const std::vector<int> &cnt = row_items_in_chunk[c];
                std::vector<int> ridx;
                for (int r = 0; r < R; ++r)
                    if (cnt[r] > 0) ridx.push_back(r);
                std::stable_sort(ridx.begin(), ridx.end(), [&](int a, int b) { return cnt[a] > cnt[b]; });
                std::vector<int> load(G, 0);
                for (int r : ridx) {
                    const int g = int(std::min_element(load.begin(), load.end()) - load.begin());
                    group_of[c][r] = g;
                    load[g] += cnt[r];
                }
                std::vector<int> gorder(G), rank(G);
                std::iota(gorder.begin(), gorder.end(), 0);
                std::stable_sort(gorder.begin(), gorder.end(), [&](int a, int b) { return load[a] > load[b]; });
                for (int k = 0; k < G; ++k) rank[gorder[k]] = k;
                for (int r : ridx) group_of[c][r] = rank[group_of[c][r]];
            }
            for (Item &it : items) it.group = group_of[it.chunk][it.r];
        }
        std::stable_sort(items.begin(), items.end(), [](const Item &x, const Item &y) {
            if (x.chunk != y.chunk) return x.chunk < y.chunk;
            if (x.group != y.group) return x.group < y.group;
            return x.elem_k < y.elem_k;
        });
        L.cg.assign(size_t(L.n_chunks) * G + 1, 0);
        for (const Item &it : items) L.cg[size_t(it.chunk) * G + it.group + 1]++;
        for (size_t i = 1; i < L.cg.size(); ++i) L.cg[i] += L.cg[i - 1];
        L.hdr.resize(items.size());
        L.pos.assign(items.size() * size_t(PS), 0);
        for (size_t i = 0; i < items.size(); ++i) {
            const Item &it = items[i];
            const int kk = it.elem_k - L.chunk_start[it.chunk];
            L.hdr[i] = kk | (it.a << 10) | (it.r << 14);
            const int node = L.rows[it.r];
            const int32_t *c0 = pat.bcol.data() + pat.brow_ptr[node], *c1 = pat.bcol.data() + pat.brow_ptr[node + 1];
            const int e = el[perm[it.elem_k]];
            for (int b = 0; b < n_local; ++b) {
                const int j = elem_nodes[size_t(e) * n_local + b];
                const int64_t ps = std::lower_bound(c0, c1, j) - c0;
                if (ps > 255) fail(PFGPU_ERR_UNSUPPORTED, "a node has more than 256 block neighbours");
                L.pos[i * PS + b] = uint8_t(ps);
            }
        }
    };
    {
        const int T = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        for (int w = 0; w < T; ++w)
            th.emplace_back([&, w] {
                for (int p = w; p < P; p += T) build_one(p);
            });
        for (auto &x : th) x.join();
    }

    // concatenate
    t.desc.assign(size_t(P) * PD_STRIDE, 0);
    int64_t nrow = 0, nel = 0, ncg = 0, nit = 0, ncs = 0;
    for (int p = 0; p < P; ++p) {
        nrow += loc[p].rows.size();
        nel += loc[p].elems.size();
        ncg += loc[p].cg.size();
        nit += loc[p].hdr.size();
        ncs += loc[p].chunk_start.size();
    }
    if (nit >= (int64_t(1) << 31) || nel >= (int64_t(1) << 31)) fail(PFGPU_ERR_UNSUPPORTED, "patch tables exceed 32-bit offsets");
    t.rows.reserve(nrow);
    t.row_acc.reserve(nrow);
    t.elems.reserve(nel);
    t.vnodes.reserve(nel * n_local);
    t.chunk_start.reserve(ncs);
    t.cg.reserve(ncg);
    t.hdr.reserve(nit);
    t.pos.reserve(nit * pos_stride(n_local));
    for (int p = 0; p < P; ++p) {
        Local &L = loc[p];
        int32_t *d = t.desc.data() + size_t(p) * PD_STRIDE;
        d[PD_ROW0] = int32_t(t.rows.size());
        d[PD_NROWS] = int32_t(L.rows.size());
        d[PD_EL0] = int32_t(t.elems.size());
        d[PD_NEL] = int32_t(L.elems.size());
        d[PD_CG0] = int32_t(t.cg.size());
        d[PD_NCH] = L.n_chunks;
        d[PD_ACC] = L.acc;
        d[PD_CHUNK_BASE] = int32_t(t.n_chunks);
        d[PD_CS0] = int32_t(t.chunk_start.size());
        const int32_t item_base = int32_t(t.hdr.size());
        t.rows.insert(t.rows.end(), L.rows.begin(), L.rows.end());
        t.row_acc.insert(t.row_acc.end(), L.row_acc.begin(), L.row_acc.end());
        t.elems.insert(t.elems.end(), L.elems.begin(), L.elems.end());
        for (int32_t ev : L.elems) {
            const int32_t *en = elem_nodes + size_t(ev & 0x7fffffff) * n_local;
            t.vnodes.insert(t.vnodes.end(), en, en + n_local);
        }
        t.chunk_start.insert(t.chunk_start.end(), L.chunk_start.begin(), L.chunk_start.end());
        for (int32_t v : L.cg) t.cg.push_back(v + item_base);
        t.hdr.insert(t.hdr.end(), L.hdr.begin(), L.hdr.end());
        t.pos.insert(t.pos.end(), L.pos.begin(), L.pos.end());
        t.max_rows = std::max(t.max_rows, int(L.rows.size()));
        t.max_acc = std::max(t.max_acc, L.acc);
        t.max_unit_items = std::max(t.max_unit_items, L.max_unit_items);
        t.n_chunks += L.n_chunks;
        t.visits += int64_t(L.elems.size());
        L = Local();
    }
    return t;
}

} // namespace pfgpu
