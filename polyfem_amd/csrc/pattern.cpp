// Host-side construction of the Hessian sparsity pattern and of the maps that
// scatter element-local results into it.
//
// The reference builds its pattern on the first Hessian assembly from
// triplets (SparseMatrixCache::prune / get_matrix, utils/MatrixCache.cpp:115-212):
// every element adds all (3n)^2 entries, duplicates are summed, explicit
// zeros are kept and the inner indices are sorted.  The result is therefore
// node adjacency (nodes sharing an element) times a full 3x3 block, which is
// what this file builds directly, without triplets.
#include <algorithm>
#include <thread>

#include "common.hpp"

namespace pfgpu {

namespace {

template <typename F> void parallel_for(int n, F &&f)
{
    const int T = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 4096 || T == 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) {
        const int s = int(int64_t(t) * n / T), e = int(int64_t(t + 1) * n / T);
        th.emplace_back([&f, s, e] { f(s, e); });
    }
    for (auto &t : th) t.join();
}

void incidence(int n_elements, int n_local, const int32_t *conn, int n_items, std::vector<int32_t> &ptr,
               std::vector<int32_t> &inc)
{
    ptr.assign(size_t(n_items) + 1, 0);
    const int64_t total = int64_t(n_elements) * n_local;
    for (int64_t k = 0; k < total; ++k) {
        const int32_t v = conn[k];
        if (v < 0 || v >= n_items) fail(PFGPU_ERR_INVALID_ARGUMENT, "element table references id " + std::to_string(v) + " outside [0, " + std::to_string(n_items) + ")");
        ptr[size_t(v) + 1]++;
    }
    for (int i = 0; i < n_items; ++i) ptr[i + 1] += ptr[i];
    inc.resize(size_t(total));
    std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
    // element order, so each list is sorted by element
    for (int64_t k = 0; k < total; ++k) inc[size_t(fill[conn[k]]++)] = int32_t(k);
}

} // namespace

Pattern build_pattern(int n_elements, int n_local, const int32_t *elem_nodes, int n_nodes, int n_vertices,
                      const int32_t *elem_vertices)
{
    Pattern p;
    p.n_nodes = n_nodes;
    incidence(n_elements, n_local, elem_nodes, n_nodes, p.node_ptr, p.node_inc);
    incidence(n_elements, 4, elem_vertices, n_vertices, p.vert_ptr, p.vert_inc);

    // conforming meshes: a node appears at most once per element
    for (int e = 0; e < n_elements; ++e)
        for (int a = 0; a < n_local; ++a)
            for (int b = a + 1; b < n_local; ++b)
                if (elem_nodes[int64_t(e) * n_local + a] == elem_nodes[int64_t(e) * n_local + b])
                    fail(PFGPU_ERR_INVALID_ARGUMENT, "element " + std::to_string(e) + " repeats a node");

    // block adjacency, per node sorted
    std::vector<int32_t> deg(n_nodes, 0);
    std::vector<std::vector<int32_t>> rows(n_nodes);
    parallel_for(n_nodes, [&](int s, int e) {
        std::vector<int32_t> tmp;
        for (int i = s; i < e; ++i) {
            tmp.clear();
            for (int k = p.node_ptr[i]; k < p.node_ptr[i + 1]; ++k) {
                const int el = p.node_inc[k] / n_local;
                for (int b = 0; b < n_local; ++b) tmp.push_back(elem_nodes[int64_t(el) * n_local + b]);
            }
            std::sort(tmp.begin(), tmp.end());
            tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
            rows[i] = tmp;
            deg[i] = int32_t(tmp.size());
        }
    });
    p.brow_ptr.assign(size_t(n_nodes) + 1, 0);
    for (int i = 0; i < n_nodes; ++i) p.brow_ptr[i + 1] = p.brow_ptr[i] + deg[i];
    p.n_blocks = p.brow_ptr[n_nodes];
    if (9 * p.n_blocks >= (int64_t(1) << 31)) fail(PFGPU_ERR_UNSUPPORTED, "Hessian has more than 2^31 non-zeros");
    p.bcol.resize(size_t(p.n_blocks));
    parallel_for(n_nodes, [&](int s, int e) {
        for (int i = s; i < e; ++i) {
            std::copy(rows[i].begin(), rows[i].end(), p.bcol.begin() + p.brow_ptr[i]);
            std::vector<int32_t>().swap(rows[i]);
        }
    });
    p.nnz = 9 * p.n_blocks;
    // scalar rows 3i+m: 3 deg(i) entries each, the three rows of node i adjacent
    p.outer.resize(size_t(3) * n_nodes + 1);
    for (int i = 0; i < n_nodes; ++i)
        for (int m = 0; m < 3; ++m) p.outer[3 * i + m] = 9 * p.brow_ptr[i] + m * 3 * deg[i];
    p.outer[3 * size_t(n_nodes)] = int32_t(p.nnz);
    return p;
}

void build_block_contributions(const Pattern &pat, int n_local, const int32_t *elem_nodes,
                               std::vector<int64_t> &ptr, std::vector<int32_t> &contrib,
                               std::vector<int32_t> &block_row)
{
    const int N = pat.n_nodes;
    if (int64_t(pat.node_inc.size()) * n_local >= (int64_t(1) << 31))
        fail(PFGPU_ERR_UNSUPPORTED, "too many element-local blocks for 32-bit contribution ids");
    std::vector<int32_t> count(size_t(pat.n_blocks), 0);
    block_row.resize(size_t(pat.n_blocks));
    // every block gets one contribution per element containing both nodes
    parallel_for(N, [&](int s, int e) {
        for (int i = s; i < e; ++i) {
            const int32_t *c0 = pat.bcol.data() + pat.brow_ptr[i], *c1 = pat.bcol.data() + pat.brow_ptr[i + 1];
            for (int32_t b = pat.brow_ptr[i]; b < pat.brow_ptr[i + 1]; ++b) block_row[b] = i;
            for (int k = pat.node_ptr[i]; k < pat.node_ptr[i + 1]; ++k) {
                const int el = pat.node_inc[k] / n_local;
                for (int b = 0; b < n_local; ++b) {
                    const int32_t j = elem_nodes[int64_t(el) * n_local + b];
                    const int64_t pos = std::lower_bound(c0, c1, j) - pat.bcol.data();
                    count[size_t(pos)]++;
                }
            }
        }
    });
    ptr.assign(size_t(pat.n_blocks) + 1, 0);
    for (int64_t b = 0; b < pat.n_blocks; ++b) ptr[b + 1] = ptr[b] + count[b];
    contrib.resize(size_t(ptr[pat.n_blocks]));
    parallel_for(N, [&](int s, int e) {
        std::vector<int64_t> fill;
        for (int i = s; i < e; ++i) {
            const int32_t *c0 = pat.bcol.data() + pat.brow_ptr[i], *c1 = pat.bcol.data() + pat.brow_ptr[i + 1];
            fill.assign(ptr.begin() + pat.brow_ptr[i], ptr.begin() + pat.brow_ptr[i + 1]);
            for (int k = pat.node_ptr[i]; k < pat.node_ptr[i + 1]; ++k) {
                const int32_t ea = pat.node_inc[k]; // e * n_local + a
                const int el = ea / n_local;
                for (int b = 0; b < n_local; ++b) {
                    const int32_t j = elem_nodes[int64_t(el) * n_local + b];
                    const int64_t pos = std::lower_bound(c0, c1, j) - c0;
                    contrib[size_t(fill[pos]++)] = ea * n_local + b;
                }
            }
        }
    });
}

} // namespace pfgpu
