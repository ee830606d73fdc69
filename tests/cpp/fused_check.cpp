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
#include <algorithm>
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

// the kernel's per-point staging record (NeoHookean), layout StageT<ORDER, NQ, NPART>;
// the gradient partial sums R must be zeroed before the first point
template <int ORDER, int NQ, int NPART>
static void stage_point(const Mesh &m, int e, int q, double *s)
{
    using S = StageT<ORDER, NQ, NPART>;
    constexpr int NB = NLocal<ORDER>::value;
    double jit[9], det;
    geometry(m, e, jit, det);
    const double w = det * quad_weight(NQ, 0), lam = m.lam[e], mu = m.mu[e];
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
    for (int b = 0; b < NB; ++b) {
        for (int mm = 0; mm < 3; ++mm) {
            double v = 0;
            for (int k = 0; k < 3; ++k) v += M[mm * 3 + k] * RG<ORDER, NQ>(q, b, k);
            s[S::K + (q * NB + b) * 4 + mm] = v;
        }
        s[S::K + (q * NB + b) * 4 + 3] = 0;
    }
    s[S::C + 2 * q] = w * lam;
    s[S::C + 2 * q + 1] = w * (mu - lam * lj);
    // point q adds to partial sum h = q / (NQ / NPART) of the element gradient
    const int h = q / (NQ / NPART);
    for (int d = 0; d < 3; ++d) {
        double Qm[3];
        for (int kk = 0; kk < 3; ++kk)
            Qm[kk] = w * (P[d * 3] * jit[kk * 3] + P[d * 3 + 1] * jit[kk * 3 + 1] + P[d * 3 + 2] * jit[kk * 3 + 2]);
        for (int a = 0; a < NB; ++a) {
            double g = 0;
            for (int kk = 0; kk < 3; ++kk) g += RG<ORDER, NQ>(q, a, kk) * Qm[kk];
            s[S::R + (h * NB + a) * 4 + d] += g;
        }
    }
    if (q == 0) {
        auto dot = [&](int k, int l) { return jit[k * 3] * jit[l * 3] + jit[k * 3 + 1] * jit[l * 3 + 1] + jit[k * 3 + 2] * jit[l * 3 + 2]; };
        const double g6[6] = {dot(0, 0), dot(1, 1), dot(2, 2), dot(0, 1), dot(0, 2), dot(1, 2)};
        for (int i = 0; i < 6; ++i) s[S::G + i] = w * mu * g6[i];
        s[S::G + 6] = s[S::G + 7] = 0;
    }
}

// block (a, b) and gradient row a of one staged element, as the kernel's
// consumer lanes form them (lane (b, mm) forms block row mm)
template <int ORDER, int NQ, int NPART>
static void consume_item(const double *sp, int a, int b, double *blk, double *grow)
{
    using S = StageT<ORDER, NQ, NPART>;
    constexpr int NB = NLocal<ORDER>::value;
    for (int i = 0; i < 9; ++i) blk[i] = 0;
    for (int q = 0; q < NQ; ++q) {
        const double *ka = sp + S::K + (q * NB + a) * 4, *kb = sp + S::K + (q * NB + b) * 4;
        const double cl = sp[S::C + 2 * q], cc = sp[S::C + 2 * q + 1];
        for (int mm = 0; mm < 3; ++mm)
            for (int nn = 0; nn < 3; ++nn) blk[mm * 3 + nn] += cl * ka[mm] * kb[nn] + cc * kb[mm] * ka[nn];
    }
    const double sab = grad_dot<ORDER, NQ>(a, b, sp + S::G);
    for (int d = 0; d < 3; ++d) blk[d * 4] += sab;
    for (int d = 0; d < 3; ++d) {
        double gs = 0;
        for (int h = 0; h < NPART; ++h) gs += sp[S::R + (h * NB + a) * 4 + d];
        grow[d] = gs;
    }
}

// a staging record before its points are staged: NaN except the gradient sums
template <int ORDER, int NQ, int NPART> static void clear_record(double *s)
{
    using S = StageT<ORDER, NQ, NPART>;
    for (int i = 0; i < S::SIZE; ++i) s[i] = i >= S::R && i < S::G ? 0.0 : std::nan("");
}

// direct assembly of the same quantities, element by element
template <int ORDER, int NQ, int NPART>
static void direct(const Mesh &m, const Pattern &pat, std::vector<double> &vals, std::vector<double> &g)
{
    constexpr int NB = NLocal<ORDER>::value;
    vals.assign(size_t(pat.nnz), 0.0);
    g.assign(size_t(m.N) * 3, 0.0);
    std::vector<double> s(StageT<ORDER, NQ, NPART>::SIZE);
    for (int e = 0; e < m.E; ++e) {
        clear_record<ORDER, NQ, NPART>(s.data());
        for (int q = 0; q < NQ; ++q) stage_point<ORDER, NQ, NPART>(m, e, q, s.data());
        for (int a = 0; a < NB; ++a) {
            const int i = m.en[e * NB + a];
            const int deg = pat.brow_ptr[i + 1] - pat.brow_ptr[i];
            for (int b = 0; b < NB; ++b) {
                const int j = m.en[e * NB + b];
                double blk[9], grow[3];
                consume_item<ORDER, NQ, NPART>(s.data(), a, b, blk, grow);
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

template <int ORDER, int NQ, int NPART>
static int run(const Mesh &m, int groups, int chunk, int item_cap, int acc_capacity, int max_rows)
{
    using S = StageT<ORDER, NQ, NPART>;
    constexpr int NB = NLocal<ORDER>::value, SE = S::SIZE;
    CHECK(SE % 8 == 4, "record length 4 mod 8 words");
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
        {
            std::vector<int> pe;
            for (int k = 0; k < nel; ++k) pe.push_back(t.elems[el0 + k] & 0x7fffffff);
            std::sort(pe.begin(), pe.end());
            for (int k = 1; k < nel; ++k) CHECK(pe[k] > pe[k - 1], "patch %d lists an element twice", p);
        }
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
                for (int a = 0; a < NB; ++a) CHECK(t.vnodes[size_t(el0 + cbeg + kk) * NB + a] == m.en[size_t(e) * NB + a], "visit node list");
                clear_record<ORDER, NQ, NPART>(&stg[size_t(kk) * SE]);
                for (int q = 0; q < NQ; ++q) stage_point<ORDER, NQ, NPART>(m, e, q, &stg[size_t(kk) * SE]);
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
                        const int pos = t.pos[size_t(it) * pos_stride(NB) + b];
                        CHECK(pos < deg && pat.bcol[pat.brow_ptr[node] + pos] == m.en[e * NB + b], "column position");
                        double blk[9], grow[3];
                        consume_item<ORDER, NQ, NPART>(sp, a, b, blk, grow);
                        for (int mm = 0; mm < 3; ++mm)
                            for (int nn = 0; nn < 3; ++nn) {
                                const int ai = t.row_acc[row0 + r] + 3 * pos + mm * 3 * deg + nn;
                                CHECK(ai >= t.row_acc[row0 + r] && ai < t.row_acc[row0 + r] + 9 * deg, "acc index in row segment");
                                acc[ai] += blk[mm * 3 + nn];
                            }
                        if (b == 0)
                            for (int d = 0; d < 3; ++d) gacc[3 * size_t(r) + d] += grow[d];
                    }
                    for (int i = 0; i < S::G + 8; ++i)
                        CHECK(!std::isnan(sp[i]) || (i >= S::C + 2 * NQ && i < S::R), "staging slot read before written");
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
    direct<ORDER, NQ, NPART>(m, pat, rv, rg);
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

// consumer lane groups fit a wavefront and the LDS copy of the item tables is
// whole words
template <typename T, int ORDER, int NQ> static void check_groups()
{
    constexpr int NB = NLocal<ORDER>::value, GPW = groups_per_wave<ORDER>(), GL = group_lanes<ORDER>();
    constexpr int CH = chunk_elems<T, NQ>(), IC = item_cap<T, ORDER, NQ>();
    using S = Stage<T, ORDER, NQ>;
    CHECK(GPW * GL <= 64 && GPW >= 1, "groups per wave");
    CHECK(pos_stride(NB) >= NB && (pos_stride(NB) == 4 || pos_stride(NB) == 16), "pos stride");
    CHECK(CH <= 1024, "element-in-chunk field is 10 bits");
    CHECK(IC >= CH * NB, "item cap below chunk x n_local");
    CHECK(S::SIZE % 8 == 4 && S::C % 4 == 0 && S::R % 4 == 0 && S::G % 4 == 0, "staging sections 16-byte aligned");
    constexpr int NPART = grad_parts<T, NQ>();
    CHECK(NPART >= 1 && NQ % NPART == 0, "gradient partial sums");
}

int main(int argc, char **argv)
{
    check_producer_map<float, 1>();
    check_producer_map<float, 4>();
    check_producer_map<double, 1>();
    check_producer_map<double, 4>();
    check_groups<float, 1, 1>();
    check_groups<float, 2, 4>();
    check_groups<double, 1, 1>();
    check_groups<double, 2, 4>();
    if (argc < 7) { std::fprintf(stderr, "usage: fused_check mesh.bin groups chunk item_cap acc_capacity max_rows\n"); return 2; }
    Mesh m = read_mesh(argv[1]);
    const int groups = std::atoi(argv[2]), chunk = std::atoi(argv[3]), icap = std::atoi(argv[4]);
    const int cap = std::atoi(argv[5]), max_rows = std::atoi(argv[6]);
    // partial gradient sums per element, as the kernel's producers leave them
    const int np = chunk >= 64 ? 4 : std::max(1, 4 * chunk / 64);
    int f = m.order == 1 ? run<1, 1, 1>(m, groups, chunk, icap, cap, max_rows)
            : np == 1    ? run<2, 4, 1>(m, groups, chunk, icap, cap, max_rows)
            : np == 2    ? run<2, 4, 2>(m, groups, chunk, icap, cap, max_rows)
                         : run<2, 4, 4>(m, groups, chunk, icap, cap, max_rows);
    std::printf(f ? "FAIL (%d)\n" : "OK\n", f);
    return f ? 1 : 0;
}
