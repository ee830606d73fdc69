// Consistent (row-sum or HRZ lumped) mass matrix of P1/P2 tets (Mass::assemble, assembler/Mass.cpp:4-20):
// M_ab = rho int phi_a phi_b, block (a, b) = M_ab I_3.  The reference
// integrates with the quadrature of order 2p (AssemblerUtils.cpp:204-211),
// which is exact for these degree-2p integrands; here the element matrix is
// the exact integral, from the barycentric-monomial formula
//   int_T l0^a l1^b l2^c l3^d = 6 |T| a! b! c! d! / (a + b + c + d + 3)!.
#include <algorithm>
#include <array>
#include <map>

#include "handle.hpp"

namespace pfgpu {

namespace {

using Mono = std::array<int, 4>; // barycentric exponents
using Poly = std::map<Mono, double>;

Poly mul(const Poly &a, const Poly &b)
{
    Poly r;
    for (const auto &[ma, ca] : a)
        for (const auto &[mb, cb] : b) {
            Mono m{ma[0] + mb[0], ma[1] + mb[1], ma[2] + mb[2], ma[3] + mb[3]};
            r[m] += ca * cb;
        }
    return r;
}

double fact(int n)
{
    double f = 1;
    for (int i = 2; i <= n; ++i) f *= i;
    return f;
}

// integral over a tet of unit volume
double integrate(const Poly &p)
{
    double s = 0;
    for (const auto &[m, c] : p)
        s += c * 6.0 * fact(m[0]) * fact(m[1]) * fact(m[2]) * fact(m[3]) / fact(m[0] + m[1] + m[2] + m[3] + 3);
    return s;
}

// Lagrange basis in barycentric coordinates, PolyFEM's local node order
// (autogen/auto_p_bases.cpp:1437-1448): v0..v3, e01, e12, e20, e03, e13, e23
std::vector<Poly> basis(int order)
{
    auto lam = [](int i) {
        Mono m{0, 0, 0, 0};
        m[i] = 1;
        return Poly{{m, 1.0}};
    };
    std::vector<Poly> b;
    if (order == 1) {
        for (int i = 0; i < 4; ++i) b.push_back(lam(i));
        return b;
    }
    for (int i = 0; i < 4; ++i) {
        Poly p = mul(lam(i), lam(i));
        for (auto &kv : p) kv.second *= 2.0;
        p[Mono{i == 0, i == 1, i == 2, i == 3}] -= 1.0; // l_i (2 l_i - 1)
        b.push_back(p);
    }
    const int E[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
    for (const auto &e : E) {
        Poly p = mul(lam(e[0]), lam(e[1]));
        for (auto &kv : p) kv.second *= 4.0;
        b.push_back(p);
    }
    return b;
}

} // namespace

std::vector<double> reference_mass(int order)
{
    const std::vector<Poly> b = basis(order);
    const int n = int(b.size());
    std::vector<double> m(size_t(n) * n);
    for (int a = 0; a < n; ++a)
        for (int c = 0; c < n; ++c) m[size_t(a) * n + c] = integrate(mul(b[a], b[c]));
    return m;
}

void build_mass(Handle &h, const double *rho, double rho_uniform)
{
    const int NB = h.n_local;
    std::vector<double> R = reference_mass(h.order);
    if (h.lump_mass == PFGPU_MASS_HRZ) {
        // HRZMass::assemble (Mass.cpp:48-86): the element's diagonal scaled
        // so that it keeps the element's total mass; the same factor for
        // every element, as the reference element matrix only scales
        double all = 0, diag = 0;
        for (int a = 0; a < NB; ++a)
            for (int c = 0; c < NB; ++c) {
                all += R[size_t(a) * NB + c];
                if (a == c) diag += R[size_t(a) * NB + c];
            }
        for (int a = 0; a < NB; ++a)
            for (int c = 0; c < NB; ++c) R[size_t(a) * NB + c] = a == c ? R[size_t(a) * NB + c] * all / diag : 0.0;
    }
    std::vector<double> vals(size_t(h.pat.n_blocks), 0.0);
    for (int e = 0; e < h.n_elements; ++e) {
        // |T| from the P1 geometry
        const int32_t *ev = h.h_elem_vertices.data() + size_t(e) * 4;
        const double *X0 = h.h_vertices.data() + 3 * size_t(ev[0]);
        double t[9];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) t[r * 3 + c] = h.h_vertices[3 * size_t(ev[r + 1]) + c] - X0[c];
        const double det = t[0] * (t[4] * t[8] - t[5] * t[7]) - t[1] * (t[3] * t[8] - t[5] * t[6]) +
                           t[2] * (t[3] * t[7] - t[4] * t[6]);
        const double r = rho ? rho[e] : rho_uniform;
        const double scale = r * det / 6.0;
        const int32_t *en = h.h_elem_nodes.data() + size_t(e) * NB;
        for (int a = 0; a < NB; ++a) {
            const int i = en[a];
            const int32_t *c0 = h.pat.bcol.data() + h.pat.brow_ptr[i], *c1 = h.pat.bcol.data() + h.pat.brow_ptr[i + 1];
            for (int c = 0; c < NB; ++c) {
                const int64_t pos = std::lower_bound(c0, c1, en[c]) - h.pat.bcol.data();
                vals[size_t(pos)] += scale * R[size_t(a) * NB + c];
            }
        }
    }
    if (h.lump_mass == PFGPU_MASS_ROW_SUM) {
        // utils::lump_matrix (MatrixUtils.cpp:84-101): every row's sum on its
        // diagonal, nothing off it (State.cpp:1542-1545)
        for (int i = 0; i < h.n_nodes; ++i) {
            const int32_t r0 = h.pat.brow_ptr[i], r1 = h.pat.brow_ptr[i + 1];
            double sum = 0;
            for (int32_t b = r0; b < r1; ++b) {
                sum += vals[size_t(b)];
                vals[size_t(b)] = 0.0;
            }
            const int64_t d = std::lower_bound(h.pat.bcol.data() + r0, h.pat.bcol.data() + r1, i) - h.pat.bcol.data();
            vals[size_t(d)] = sum;
        }
    }
    upload_into(h.mass, vals, h.stream);
    if (!h.brow_ptr.get()) h.brow_ptr = upload(h.pat.brow_ptr, h.stream);
    if (!h.pat_bcol.get()) h.pat_bcol = upload(h.pat.bcol, h.stream);
    h.has_mass = true;
}

} // namespace pfgpu
