// CPU check of the fused kernel's patch tables and indexing (no GPU).
//
// Builds the pattern and the patch tables exactly as the library does, then
//  1. checks every index the kernel derives from the tables against the
//     bounds of the buffers it addresses (LDS accumulator, staging slots,
//     element lists, column positions);
//  2. checks that every (row, incident element) pair is scheduled exactly
//     once and every node is owned by exactly one patch;
//  3. replays phase A / phase B / write-out serially in double precision and
//     compares the resulting Hessian values and gradient with a direct
//     element-by-element assembly.
// Input: a binary mesh file written by the test (see tests/test_fused_tables.py).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../polyfem_amd/csrc/elem.hpp"
#include "../../polyfem_amd/csrc/fused.hpp"

using namespace pfgpu;

static int n_fail = 0;
#define CHECK(c, ...)                                                                                          \
    do {                                                                                                       \
        if (!(c)) {                                                                                            \
            if (n_fail < 20) { std::fprintf(stderr, "CHECK failed %s:%d: ", __FILE__, __LINE__); std::fprintf(stderr, __VA_ARGS__); std::fprintf(stderr, "\n"); } \
            n_fail++;                                                                                          \
        }                                                                                                      \
    } while (0)

struct Mesh {
    int order, E, N, Nv, nb;
    std::vector<int32_t> en, ev;
    std::vector<double> vx, x, lam, mu;
};

static Mesh read_mesh(const char *path)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) { std::perror(path); std::exit(2); }
    Mesh m;
    int hdr[4];
    if (std::fread(hdr, 4, 4, f) != 4) std::exit(2);
    m.order = hdr[0]; m.E = hdr[1]; m.N = hdr[2]; m.Nv = hdr[3];
    m.nb = m.order == 1 ? 4 : 10;
    m.en.resize(size_t(m.E) * m.nb); m.ev.resize(size_t(m.E) * 4);
    m.vx.resize(size_t(m.Nv) * 3); m.x.resize(size_t(m.N) * 3); m.lam.resize(m.E); m.mu.resize(m.E);
    size_t ok = std::fread(m.en.data(), 4, m.en.size(), f) + std::fread(m.ev.data(), 4, m.ev.size(), f) +
                std::fread(m.vx.data(), 8, m.vx.size(), f) + std::fread(m.x.data(), 8, m.x.size(), f) +
                std::fread(m.lam.data(), 8, m.lam.size(), f) + std::fread(m.mu.data(), 8, m.mu.size(), f);
    std::fclose(f);
    if (ok != m.en.size() + m.ev.size() + m.vx.size() + m.x.size() + m.lam.size() + m.mu.size()) std::exit(2);
    return m;
}

// geometry: J^{-T} row-major and det, as the library's host setup computes it
static void geometry(const Mesh &m, int e, double *jit, double &det)
{
    const double *X0 = &m.vx[3 * size_t(m.ev[4 * e])];
    double t[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) t[r * 3 + c] = m.vx[3 * size_t(m.ev[4 * e + r + 1]) + c] - X0[c];
    const double c00 = t[4] * t[8] - t[5] * t[7], c01 = t[5] * t[6] - t[3] * t[8], c02 = t[3] * t[7] - t[4] * t[6];
    const double c10 = t[2] * t[7] - t[1] * t[8], c11 = t[0] * t[8] - t[2] * t[6], c12 = t[1] * t[6] - t[0] * t[7];
    const double c20 = t[1] * t[5] - t[2] * t[4], c21 = t[2] * t[3] - t[0] * t[5], c22 = t[0] * t[4] - t[1] * t[3];
    det = t[0] * c00 + t[1] * c01 + t[2] * c02;
    const double id = 1.0 / det, J[9] = {c00, c01, c02, c10, c11, c12, c20, c21, c22};
    for (int k = 0; k < 9; ++k) jit[k] = J[k] * id;
}

// the kernel's per-point staging record (NeoHookean), layout Stage<ORDER, NQ>
template <int ORDER, int NQ>
static void stage_point(const Mesh &m, int e, int q, double *s)
{
    using S = Stage<ORDER, NQ>;
    constexpr int NB = NLocal<ORDER>::value;
    double jit[9], det;
    geometry(m, e, jit, det);
    const double w = det * quad_weight(NQ), lam = m.lam[e], mu = m.mu[e];
    double A[9] = {0};
    for (int a = 0; a < NB; ++a)
        for (int k = 0; k < 3; ++k)
            for (int d = 0; d < 3; ++d) A[d * 3 + k] += m.x[3 * size_t(m.en[e * NB + a]) + d] * RG<ORDER, NQ>(q, a, k);
    double F[9];
    for (int d = 0; d < 3; ++d)
        for (int c = 0; c < 3; ++c) F[d * 3 + c] = A[d * 3] * jit[c] + A[d * 3 + 1] * jit[3 + c] + A[d * 3 + 2] * jit[6 + c] + (d == c);
    const double J = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) + F[2] * (F[3] * F[7] - F[4] * F[6]);
    const double cof[9] = {F[4] * F[8] - F[5] * F[7], F[5] * F[6] - F[3] * F[8], F[3] * F[7] - F[4] * F[6],
                           F[2] * F[7] - F[1] * F[8], F[0] * F[8] - F[2] * F[6], F[1] * F[6] - F[0] * F[7],
                           F[1] * F[5] - F[2] * F[4], F[2] * F[3] - F[0] * F[5], F[0] * F[4] - F[1] * F[3]};
    double FmT[9], P[9];
    const double lj = std::log(J);
    for (int i = 0; i < 9; ++i) {
        FmT[i] = cof[i] / J;
        P[i] = mu * (F[i] - FmT[i]) + lam * lj * FmT[i];
    }
    double M[9];
    for (int mm = 0; mm < 3; ++mm)
        for (int kk = 0; kk < 3; ++kk)
            M[mm * 3 + kk] = FmT[mm * 3] * jit[kk * 3] + FmT[mm * 3 + 1] * jit[kk * 3 + 1] + FmT[mm * 3 + 2] * jit[kk * 3 + 2];
    for (int b = 0; b < NB; ++b)
        for (int mm = 0; mm < 3; ++mm) {
            double v = 0;
            for (int k = 0; k < 3; ++k) v += M[mm * 3 + k] * RG<ORDER, NQ>(q, b, k);
            s[S::K + (q * NB + b) * 3 + mm] = v;
        }
    s[S::C + q] = w * lam;
    s[S::C + NQ + q] = w * (mu - lam * lj);
    for (int d = 0; d < 3; ++d)
        for (int kk = 0; kk < 3; ++kk)
            s[S::Q + q * 9 + d * 3 + kk] = w * (P[d * 3] * jit[kk * 3] + P[d * 3 + 1] * jit[kk * 3 + 1] + P[d * 3 + 2] * jit[kk * 3 + 2]);
    if (q == 0) {
        auto dot = [&](int k, int l) { return jit[k * 3] * jit[l * 3] + jit[k * 3 + 1] * jit[l * 3 + 1] + jit[k * 3 + 2] * jit[l * 3 + 2]; };
        const double g6[6] = {dot(0, 0), dot(1, 1), dot(2, 2), dot(0, 1), dot(0, 2), dot(1, 2)};
        for (int i = 0; i < 6; ++i) s[S::G + i] = w * mu * g6[i];
    }
}

// block (a, b) and gradient row a of one staged element, as the kernel's consumer lanes form them
template <int ORDER, int NQ>
static void consume_item(const double *sp, int a, int b, double *blk, double *grow)
{
    using S = Stage<ORDER, NQ>;
    constexpr int NB = NLocal<ORDER>::value;
    for (int i = 0; i < 9; ++i) blk[i] = 0;
    for (int q = 0; q < NQ; ++q) {
        const double *ka = sp + S::K + (q * NB + a) * 3, *kb = sp + S::K + (q * NB + b) * 3;
        const double cl = sp[S::C + q], cc = sp[S::C + NQ + q];
        for (int mm = 0; mm < 3; ++mm)
            for (int nn = 0; nn < 3; ++nn) blk[mm * 3 + nn] += cl * ka[mm] * kb[nn] + cc * kb[mm] * ka[nn];
    }
    const double sab = grad_dot<ORDER, NQ>(a, b, sp + S::G);
    for (int d = 0; d < 3; ++d) blk[d * 4] += sab;
    for (int d = 0; d < 3; ++d) {
        double gs = 0;
        for (int q = 0; q < NQ; ++q)
            for (int k = 0; k < 3; ++k) gs += RG<ORDER, NQ>(q, a, k) * sp[S::Q + q * 9 + d * 3 + k];
        grow[d] = gs;
    }
}

// direct assembly of the same quantities, element by element
template <int ORDER, int NQ>
static void direct(const Mesh &m, const Pattern &pat, std::vector<double> &vals, std::vector<double> &g)
{
    constexpr int NB = NLocal<ORDER>::value;
    vals.assign(size_t(pat.nnz), 0.0);
    g.assign(size_t(m.N) * 3, 0.0);
    std::vector<double> s(Stage<ORDER, NQ>::SIZE);
    for (int e = 0; e < m.E; ++e) {
        for (int q = 0; q < NQ; ++q) stage_point<ORDER, NQ>(m, e, q, s.data());
        for (int a = 0; a < NB; ++a) {
            const int i = m.en[e * NB + a];
            const int deg = pat.brow_ptr[i + 1] - pat.brow_ptr[i];
            for (int b = 0; b < NB; ++b) {
                const int j = m.en[e * NB + b];
                double blk[9], grow[3];
                consume_item<ORDER, NQ>(s.data(), a, b, blk, grow);
                if (b == 0)
                    for (int d = 0; d < 3; ++d) g[3 * size_t(i) + d] += grow[d];
                int pos = -1;
                for (int k = pat.brow_ptr[i]; k < pat.brow_ptr[i + 1]; ++k)
                    if (pat.bcol[k] == j) pos = k - pat.brow_ptr[i];
                for (int mm = 0; mm < 3; ++mm)
                    for (int nn = 0; nn < 3; ++nn) vals[9 * size_t(pat.brow_ptr[i]) + mm * 3 * deg + 3 * pos + nn] += blk[mm * 3 + nn];
            }
        }
    }
}

template <int ORDER, int NQ>
static int run(const Mesh &m, int groups, int chunk, int item_cap, int acc_capacity, int max_rows)
{
    constexpr int NB = NLocal<ORDER>::value, SE = Stage<ORDER, NQ>::SIZE;
    Pattern pat = build_pattern(m.E, NB, m.en.data(), m.N, m.Nv, m.ev.data());
    FusedShape s{groups, chunk, item_cap, acc_capacity, max_rows};
    FusedTables t = build_fused_tables(pat, NB, m.en.data(), m.E, m.vx.data(), m.ev.data(), s);
    const int G = s.groups;
    std::printf("patches %d max_rows %d max_acc %d max_unit_items %d chunks %ld visits/E %.3f\n", t.n_patches,
                t.max_rows, t.max_acc, t.max_unit_items, long(t.n_chunks), double(t.visits) / m.E);
    CHECK(t.max_unit_items <= item_cap, "chunk item cap");
    int64_t chunk_sum = 0;
    std::vector<int> owner(m.N, -1);
    std::vector<double> vals(size_t(pat.nnz), -1e300), grad(size_t(m.N) * 3, -1e300);
    std::vector<long> seen_pairs(size_t(m.E) * NB, 0);
    for (int p = 0; p < t.n_patches; ++p) {
        const int32_t *D = &t.desc[PD_STRIDE * size_t(p)];
        const int row0 = D[PD_ROW0], nrows = D[PD_NROWS], el0 = D[PD_EL0], nel = D[PD_NEL], cg0 = D[PD_CG0];
        const int nch = D[PD_NCH], nacc = D[PD_ACC], cs0 = D[PD_CS0];
        CHECK(nrows <= t.max_rows && nacc <= t.max_acc && nacc <= acc_capacity, "patch %d size", p);
        CHECK(D[PD_CHUNK_BASE] == chunk_sum, "chunk base");
        chunk_sum += nch;
        CHECK(t.chunk_start[cs0] == 0 && t.chunk_start[cs0 + nch] == nel, "patch %d chunk range", p);
        CHECK(cg0 + nch * G < int(t.cg.size()), "patch %d cg range", p);
        for (int r = 0; r < nrows; ++r) {
            const int node = t.rows[row0 + r];
            CHECK(node >= 0 && node < m.N, "row node");
            CHECK(owner[node] == -1, "node %d owned twice", node);
            owner[node] = p;
            const int deg = pat.brow_ptr[node + 1] - pat.brow_ptr[node];
            CHECK(t.row_acc[row0 + r] + 9 * deg <= nacc, "row acc bound");
            if (r > 0) CHECK(t.row_acc[row0 + r] == t.row_acc[row0 + r - 1] + 9 * (pat.brow_ptr[t.rows[row0 + r - 1] + 1] - pat.brow_ptr[t.rows[row0 + r - 1]]), "row acc contiguous");
        }
        for (int k = 1; k < nel; ++k) CHECK((t.elems[el0 + k] & 0x7fffffff) > (t.elems[el0 + k - 1] & 0x7fffffff), "elements sorted");
        // replay
        std::vector<double> acc(nacc, 0.0), gacc(3 * size_t(nrows), 0.0), stg(size_t(chunk) * SE, std::nan(""));
        for (int c = 0; c < nch; ++c) {
            const int cbeg = t.chunk_start[cs0 + c], cn = t.chunk_start[cs0 + c + 1] - cbeg;
            CHECK(cn > 0 && cn <= chunk, "chunk size");
            CHECK(t.cg[cg0 + (c + 1) * G] - t.cg[cg0 + c * G] <= item_cap, "chunk items");
            std::fill(stg.begin(), stg.end(), std::nan(""));
            for (int kk = 0; kk < cn; ++kk) {
                const int e = t.elems[el0 + cbeg + kk] & 0x7fffffff;
                CHECK(e >= 0 && e < m.E, "element id");
                for (int q = 0; q < NQ; ++q) stage_point<ORDER, NQ>(m, e, q, &stg[size_t(kk) * SE]);
            }
            // within a chunk every row is accumulated by exactly one lane group
            std::vector<int> row_group(nrows, -1);
            for (int grp = 0; grp < G; ++grp)
                for (int it = t.cg[cg0 + c * G + grp]; it < t.cg[cg0 + c * G + grp + 1]; ++it) {
                    const int r = t.hdr[it] >> 14;
                    if (r < nrows) {
                        CHECK(row_group[r] == -1 || row_group[r] == grp, "row %d in two groups of one chunk", r);
                        row_group[r] = grp;
                    }
                }
            for (int grp = 0; grp < G; ++grp) {
                const int it0 = t.cg[cg0 + c * G + grp], it1 = t.cg[cg0 + c * G + grp + 1];
                CHECK(it0 <= it1 && it1 <= int(t.hdr.size()), "item range");
                for (int it = it0; it < it1; ++it) {
                    const int hdr = t.hdr[it];
                    const int kk = hdr & 1023, a = (hdr >> 10) & 15, r = hdr >> 14;
                    CHECK(kk < cn && a < NB && r < nrows, "item fields %d %d %d", kk, a, r);
                    const int e = t.elems[el0 + cbeg + kk] & 0x7fffffff;
                    const int node = t.rows[row0 + r];
                    CHECK(m.en[e * NB + a] == node, "item element does not contain its row node");
                    seen_pairs[size_t(e) * NB + a]++;
                    const int deg = pat.brow_ptr[node + 1] - pat.brow_ptr[node];
                    const double *sp = &stg[size_t(kk) * SE];
                    for (int b = 0; b < NB; ++b) {
                        const int pos = t.pos[size_t(it) * NB + b];
                        CHECK(pos < deg && pat.bcol[pat.brow_ptr[node] + pos] == m.en[e * NB + b], "column position");
                        double blk[9], grow[3];
                        consume_item<ORDER, NQ>(sp, a, b, blk, grow);
                        for (int mm = 0; mm < 3; ++mm)
                            for (int nn = 0; nn < 3; ++nn) {
                                const int ai = t.row_acc[row0 + r] + 3 * pos + mm * 3 * deg + nn;
                                CHECK(ai >= t.row_acc[row0 + r] && ai < t.row_acc[row0 + r] + 9 * deg, "acc index in row segment");
                                acc[ai] += blk[mm * 3 + nn];
                            }
                        if (b < 3) gacc[3 * size_t(r) + b] += grow[b];
                    }
                    for (int i = 0; i < Stage<ORDER, NQ>::G + 6; ++i)
                        CHECK(!std::isnan(sp[i]) || (i >= Stage<ORDER, NQ>::C + 2 * NQ && i < Stage<ORDER, NQ>::Q) ||
                                  (i >= Stage<ORDER, NQ>::Q + 9 * NQ && i < Stage<ORDER, NQ>::G) ||
                                  (i >= NQ * NB * 3 && i < Stage<ORDER, NQ>::C),
                              "staging slot read before written");
                }
            }
        }
        for (int r = 0; r < nrows; ++r) {
            const int node = t.rows[row0 + r];
            const int deg = pat.brow_ptr[node + 1] - pat.brow_ptr[node];
            for (int j = 0; j < 9 * deg; ++j) vals[9 * size_t(pat.brow_ptr[node]) + j] = acc[t.row_acc[row0 + r] + j];
            for (int d = 0; d < 3; ++d) grad[3 * size_t(node) + d] = gacc[3 * size_t(r) + d];
        }
    }
    for (int i = 0; i < m.N; ++i) CHECK(owner[i] >= 0, "node %d not owned", i);
    for (size_t k = 0; k < seen_pairs.size(); ++k) CHECK(seen_pairs[k] == 1, "pair (e,a)=%zu scheduled %ld times", k, seen_pairs[k]);
    std::vector<double> rv, rg;
    direct<ORDER, NQ>(m, pat, rv, rg);
    double ev = 0, sv = 0, eg = 0, sg = 0;
    for (size_t i = 0; i < rv.size(); ++i) { ev = std::max(ev, std::fabs(vals[i] - rv[i])); sv = std::max(sv, std::fabs(rv[i])); }
    for (size_t i = 0; i < rg.size(); ++i) { eg = std::max(eg, std::fabs(grad[i] - rg[i])); sg = std::max(sg, std::fabs(rg[i])); }
    std::printf("replay vs direct: hessian rel %.3e gradient rel %.3e\n", ev / sv, eg / sg);
    CHECK(ev <= 1e-12 * sv && eg <= 1e-12 * sg, "replay differs from direct assembly");
    return n_fail;
}

// every producer thread maps to a distinct (point, element) and all pairs of a chunk are covered
template <typename T, int NQ> static void check_producer_map()
{
    constexpr int CH = chunk_elems<T, NQ>(), WA = producer_waves<T, NQ>();
    std::vector<int> seen(size_t(CH) * NQ, 0);
    for (int t = 0; t < WA * 64; ++t) {
        int q = -1, kk = -1;
        producer_task<T, NQ>(t, q, kk);
        CHECK(q >= 0 && q < NQ && kk >= 0 && kk < CH, "producer task %d -> (%d, %d) out of range", t, q, kk);
        if (q >= 0 && q < NQ && kk >= 0 && kk < CH) seen[size_t(kk) * NQ + q]++;
    }
    for (size_t i = 0; i < seen.size(); ++i) CHECK(seen[i] == 1, "producer pair %zu covered %d times", i, seen[i]);
    CHECK(WA < NWAVE, "consumers left");
}

int main(int argc, char **argv)
{
    check_producer_map<float, 1>();
    check_producer_map<float, 4>();
    check_producer_map<double, 1>();
    check_producer_map<double, 4>();
    if (argc < 7) { std::fprintf(stderr, "usage: fused_check mesh.bin groups chunk item_cap acc_capacity max_rows\n"); return 2; }
    Mesh m = read_mesh(argv[1]);
    const int groups = std::atoi(argv[2]), chunk = std::atoi(argv[3]), icap = std::atoi(argv[4]);
    const int cap = std::atoi(argv[5]), max_rows = std::atoi(argv[6]);
    int f = m.order == 1 ? run<1, 1>(m, groups, chunk, icap, cap, max_rows) : run<2, 4>(m, groups, chunk, icap, cap, max_rows);
    std::printf(f ? "FAIL (%d)\n" : "OK\n", f);
    return f ? 1 : 0;
}
