/*
 * FP64 CPU oracle for PolyFEM's nonlinear-elasticity assembly path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the CPU
 * baseline timed beside it.  Nothing in the product links or loads this file.
 *
 * Every routine restates the reference's arithmetic in the reference's own
 * order of operations where that order is visible in the source (Eigen
 * expression evaluation order is not pinned, so parity is tolerance-based).
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

struct orc_problem {
    int order, n_local, n_quad, material;
    int n_elements, n_nodes, n_vertices;
    const int32_t *elem_nodes;
    const double *vertices;
    const int32_t *elem_vertices;
    const double *lambda, *mu;
    double qp[64 * 3], qw[64];
    double ref_grad[64 * 10 * 3]; /* [q][a][c] */
    double geo_grad[64 * 4 * 3];  /* P1 geometric basis [q][j][c] */
    double *jac_it;               /* [e][q][9] row-major */
    double *det;                  /* [e][q] */
};

/* ---------------------------------------------------------------- tables */

/* Tet rules (autogen/auto_tetrahedron.ipp:1-24): order 1 is the centroid,
 * order 2 the Hammer-Marlowe-Stroud 4-point rule, orders 3 and 4 the
 * Xiao-Gimbutas rules of 6 and 11 points (rows x, y, z, w); weights sum to 1
 * and are divided by 6 (quadrature/TetQuadrature.cpp:43-55).  Orders above 4
 * return 0 points. */
static const double XG_3[6][4] = {
        {0.16200149169852446, 0.18385035049209775, 0.01271836631368145, 0.12232200275734507},
        {0.010905212211189244, 0.2815238021235462, 0.3621268299455338, 0.12806641271074692},
        {0.19011700243928392, 0.011403329444557169, 0.35862072046688387, 0.13256802714444524},
        {0.170816925164989, 0.15281814309092734, 0.6384932999617267, 0.14062440966040324},
        {0.15868516322744058, 0.5856628056552158, 0.13084716895209647, 0.22441516691755742},
        {0.5712260521491151, 0.1469183900871696, 0.14037280579421071, 0.2520039808095023},
};
static const double XG_4[11][4] = {
        {0.04049050672759043, 0.013560701879802881, 0.7712547326953761, 0.03925109092483996},
        {0.752508507009655, 0.06809937093820666, 0.09798720364927911, 0.0552733691559369},
        {0.0672232948933834, 0.035183929773598716, 0.15636389323939529, 0.05539379887157637},
        {0.419266313879513, 0.04778143555908666, 0.47961101102565507, 0.05993318514655953},
        {0.45076587609127683, 0.05945661629943383, 0.05682401712793367, 0.06946996593763537},
        {0.1294113737889104, 0.3301904148374645, 0.0023910074574393647, 0.07616271524555836},
        {0.1215419913339278, 0.3064939884296903, 0.5629727601430461, 0.07942668006802531},
        {0.09720464458758327, 0.68439041545304, 0.1118007673973831, 0.1064680341554901},
        {0.029569495206479612, 0.3179035602133946, 0.32329398483747895, 0.11023423242849766},
        {0.43271023904776856, 0.3538232392092971, 0.10962240533194123, 0.15497611601624608},
        {0.24027666492807262, 0.12680172591539202, 0.32847320672203845, 0.19341081204963445},
};

int orc_quadrature(int quad_order, double *points, double *weights)
{
    if (quad_order == 3 || quad_order == 4) {
        const int n = quad_order == 3 ? 6 : 11;
        for (int q = 0; q < n; ++q) {
            const double *r = quad_order == 3 ? XG_3[q] : XG_4[q];
            points[3 * q] = r[0];
            points[3 * q + 1] = r[1];
            points[3 * q + 2] = r[2];
            weights[q] = r[3] / 6.0;
        }
        return n;
    }
    if (quad_order <= 1) {
        points[0] = points[1] = points[2] = 0.25;
        weights[0] = 1.0 / 6.0;
        return 1;
    }
    if (quad_order == 2) {
        const double a = 0.13819660112501053195188660538406111299991607666015625;
        const double b = 0.58541019662496840414434018384781666100025177001953125;
        const double pts[12] = {a, a, b, a, b, a, b, a, a, a, a, a};
        for (int i = 0; i < 12; ++i) points[i] = pts[i];
        for (int q = 0; q < 4; ++q) weights[q] = 0.25 / 6.0;
        return 4;
    }
    return 0;
}

/* P1/P2 tet basis gradients on the reference element
 * (autogen/auto_p_bases.cpp:1100-1123 for P1, :1260-1330 for P2). */
void orc_basis_grad(int order, const double xi[3], double *g)
{
    const double x = xi[0], y = xi[1], z = xi[2];
    if (order == 1) {
        const double t[12] = {-1, -1, -1, 1, 0, 0, 0, 1, 0, 0, 0, 1};
        memcpy(g, t, sizeof(t));
        return;
    }
    const double s = 4 * x + 4 * y + 4 * z - 3;
    const double t[30] = {
        s, s, s,
        4 * x - 1, 0, 0,
        0, 4 * y - 1, 0,
        0, 0, 4 * z - 1,
        4 * (-2 * x - y - z + 1), -4 * x, -4 * x,
        4 * y, 4 * x, 0,
        -4 * y, 4 * (-x - 2 * y - z + 1), -4 * y,
        -4 * z, -4 * z, 4 * (-x - y - 2 * z + 1),
        4 * z, 0, 4 * x,
        0, 4 * z, 4 * y};
    memcpy(g, t, sizeof(t));
}

/* Basis values (auto_p_bases.cpp:1050-1079 for P1, :1178-1240 for P2). */
void orc_basis_val(int order, const double xi[3], double *v)
{
    const double x = xi[0], y = xi[1], z = xi[2];
    if (order == 1) {
        v[0] = 1 - x - y - z;
        v[1] = x;
        v[2] = y;
        v[3] = z;
        return;
    }
    const double l = x + y + z - 1;
    v[0] = l * (2 * x + 2 * y + 2 * z - 1);
    v[1] = x * (2 * x - 1);
    v[2] = y * (2 * y - 1);
    v[3] = z * (2 * z - 1);
    v[4] = -4 * x * l;
    v[5] = 4 * x * y;
    v[6] = -4 * y * l;
    v[7] = -4 * z * l;
    v[8] = 4 * x * z;
    v[9] = 4 * y * z;
}

/* ------------------------------------------------------------- 3x3 algebra */

/* Eigen's 3x3 determinant: cofactor expansion along the first row. */
static double det3(const double *m)
{
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
}

/* inverse via the adjugate */
static void inv3(const double *m, double *out)
{
    const double c00 = m[4] * m[8] - m[5] * m[7];
    const double c01 = m[5] * m[6] - m[3] * m[8];
    const double c02 = m[3] * m[7] - m[4] * m[6];
    const double c10 = m[2] * m[7] - m[1] * m[8];
    const double c11 = m[0] * m[8] - m[2] * m[6];
    const double c12 = m[1] * m[6] - m[0] * m[7];
    const double c20 = m[1] * m[5] - m[2] * m[4];
    const double c21 = m[2] * m[3] - m[0] * m[5];
    const double c22 = m[0] * m[4] - m[1] * m[3];
    const double d = m[0] * c00 + m[1] * c01 + m[2] * c02;
    const double id = 1.0 / d;
    out[0] = c00 * id; out[1] = c10 * id; out[2] = c20 * id;
    out[3] = c01 * id; out[4] = c11 * id; out[5] = c21 * id;
    out[6] = c02 * id; out[7] = c12 * id; out[8] = c22 * id;
}

static void transpose3(const double *m, double *t)
{
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) t[c * 3 + r] = m[r * 3 + c];
}

static void cross3(const double *x, const double *y, double *z)
{
    z[0] = x[1] * y[2] - x[2] * y[1];
    z[1] = x[2] * y[0] - x[0] * y[2];
    z[2] = x[0] * y[1] - x[1] * y[0];
}

/* ------------------------------------------------------------ the problem */

orc_problem *orc_create(int order, int quad_order, int material,
                        int n_elements, const int32_t *elem_nodes, int n_nodes,
                        int n_vertices, const double *vertices, const int32_t *elem_vertices,
                        const double *lambda, const double *mu)
{
    if (order != 1 && order != 2) return NULL;
    orc_problem *p = (orc_problem *)calloc(1, sizeof(orc_problem));
    p->order = order;
    p->n_local = order == 1 ? 4 : 10;
    if (quad_order <= 0) quad_order = order == 1 ? 1 : 2; /* max(2(p-1),1) */
    p->n_quad = orc_quadrature(quad_order, p->qp, p->qw);
    p->material = material;
    p->n_elements = n_elements;
    p->n_nodes = n_nodes;
    p->n_vertices = n_vertices;
    p->elem_nodes = elem_nodes;
    p->vertices = vertices;
    p->elem_vertices = elem_vertices;
    p->lambda = lambda;
    p->mu = mu;
    for (int q = 0; q < p->n_quad; ++q) {
        orc_basis_grad(order, p->qp + 3 * q, p->ref_grad + q * p->n_local * 3);
        orc_basis_grad(1, p->qp + 3 * q, p->geo_grad + q * 12);
    }
    const int nq = p->n_quad;
    p->jac_it = (double *)malloc(sizeof(double) * 9 * nq * (size_t)n_elements);
    p->det = (double *)malloc(sizeof(double) * nq * (size_t)n_elements);
    /* ElementAssemblyValues::finalize3d (ElementAssemblyValues.cpp:73-103):
     * tmp.row(r) = sum_j dphi_j/dxi_r * node_j, det = |tmp|, jac_it = tmp^{-T} */
    for (int e = 0; e < n_elements; ++e) {
        for (int q = 0; q < nq; ++q) {
            double tmp[9] = {0};
            for (int j = 0; j < 4; ++j) {
                const double *X = vertices + 3 * (size_t)elem_vertices[4 * (size_t)e + j];
                const double *g = p->geo_grad + q * 12 + j * 3;
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) tmp[r * 3 + c] += g[r] * X[c];
            }
            double inv[9];
            p->det[(size_t)e * nq + q] = det3(tmp);
            inv3(tmp, inv);
            transpose3(inv, p->jac_it + ((size_t)e * nq + q) * 9);
        }
    }
    return p;
}

void orc_destroy(orc_problem *p)
{
    if (!p) return;
    free(p->jac_it);
    free(p->det);
    free(p);
}

int orc_n_quad(const orc_problem *p) { return p->n_quad; }
int orc_n_local(const orc_problem *p) { return p->n_local; }

void orc_geometry(const orc_problem *p, int e, double *jac_it, double *det)
{
    memcpy(jac_it, p->jac_it + (size_t)e * p->n_quad * 9, sizeof(double) * 9 * p->n_quad);
    memcpy(det, p->det + (size_t)e * p->n_quad, sizeof(double) * p->n_quad);
}

/* local displacement U(a,d) (NeoHookeanElasticity.cpp:343-355) */
static void local_disp(const orc_problem *p, int e, const double *x, double *U)
{
    const int n = p->n_local;
    for (int a = 0; a < n; ++a) {
        const int32_t node = p->elem_nodes[(size_t)e * n + a];
        for (int d = 0; d < 3; ++d) U[a * 3 + d] = x[3 * (size_t)node + d];
    }
}

/* grad_t_m = grad * jac_it, rows a (ElementAssemblyValues.cpp:101-102) */
static void phys_grad(const orc_problem *p, int e, int q, double *G)
{
    const double *J = p->jac_it + ((size_t)e * p->n_quad + q) * 9;
    const double *g = p->ref_grad + q * p->n_local * 3;
    for (int a = 0; a < p->n_local; ++a)
        for (int c = 0; c < 3; ++c)
            G[a * 3 + c] = g[a * 3] * J[0 * 3 + c] + g[a * 3 + 1] * J[1 * 3 + c] + g[a * 3 + 2] * J[2 * 3 + c];
}

static double da_of(const orc_problem *p, int e, int q)
{
    return p->det[(size_t)e * p->n_quad + q] * p->qw[q];
}

/* --------------------------------------------------------- NeoHookean */

/* NeoHookeanElasticity::compute_energy_aux<double,n,3> (NeoHookeanElasticity.cpp:337-386):
 * F = (U^T grad) jac_it + I, psi = mu/2 (|F|^2 - 3 - 2 ln J) + lambda/2 ln^2 J */
static double nh_energy(const orc_problem *p, int e, const double *x)
{
    const int n = p->n_local;
    double U[30];
    local_disp(p, e, x, U);
    const double lambda = p->lambda[e], mu = p->mu[e];
    double energy = 0;
    for (int q = 0; q < p->n_quad; ++q) {
        const double *g = p->ref_grad + q * n * 3;
        const double *J = p->jac_it + ((size_t)e * p->n_quad + q) * 9;
        double A[9] = {0}; /* U^T grad : A(d,k) = sum_a U(a,d) g(a,k) */
        for (int a = 0; a < n; ++a)
            for (int d = 0; d < 3; ++d)
                for (int k = 0; k < 3; ++k) A[d * 3 + k] += U[a * 3 + d] * g[a * 3 + k];
        double F[9];
        for (int d = 0; d < 3; ++d)
            for (int c = 0; c < 3; ++c)
                F[d * 3 + c] = A[d * 3 + 0] * J[0 * 3 + c] + A[d * 3 + 1] * J[1 * 3 + c] + A[d * 3 + 2] * J[2 * 3 + c] + (d == c ? 1.0 : 0.0);
        const double detF = det3(F);
        const double ld = log(detF);
        double sq = 0;
        for (int i = 0; i < 9; ++i) sq += F[i] * F[i];
        const double val = mu / 2 * (sq - 3 - 2 * ld) + lambda / 2 * ld * ld;
        energy += val * da_of(p, e, q);
    }
    return energy;
}

/* deformation gradient from delF_delU = grad * jac_it (NeoHookeanElasticity.cpp:489-496) */
static void nh_defgrad(const orc_problem *p, int e, int q, const double *U, double *G, double *F)
{
    const int n = p->n_local;
    phys_grad(p, e, q, G);
    for (int d = 0; d < 3; ++d)
        for (int c = 0; c < 3; ++c) {
            double s = 0;
            for (int a = 0; a < n; ++a) s += U[a * 3 + d] * G[a * 3 + c];
            F[d * 3 + c] = s + (d == c ? 1.0 : 0.0);
        }
}

/* delJ_delF = [v x w, w x u, u x v] columns (NeoHookeanElasticity.cpp:509-527) */
static void cofactor_cols(const double *F, double *C)
{
    double u[3] = {F[0], F[3], F[6]}, v[3] = {F[1], F[4], F[7]}, w[3] = {F[2], F[5], F[8]};
    double c0[3], c1[3], c2[3];
    cross3(v, w, c0);
    cross3(w, u, c1);
    cross3(u, v, c2);
    for (int r = 0; r < 3; ++r) {
        C[r * 3 + 0] = c0[r];
        C[r * 3 + 1] = c1[r];
        C[r * 3 + 2] = c2[r];
    }
}

/* compute_energy_aux_gradient_fast<n,3> (NeoHookeanElasticity.cpp:453-545) */
static void nh_gradient(const orc_problem *p, int e, const double *x, double *out)
{
    const int n = p->n_local;
    double U[30], G[30], F[9], C[9];
    local_disp(p, e, x, U);
    const double lambda = p->lambda[e], mu = p->mu[e];
    double acc[30] = {0}; /* acc(a,d) */
    for (int q = 0; q < p->n_quad; ++q) {
        nh_defgrad(p, e, q, U, G, F);
        const double J = det3(F);
        const double ld = log(J);
        cofactor_cols(F, C);
        double P[9];
        for (int i = 0; i < 9; ++i) P[i] = mu * F[i] - mu * (1 / J) * C[i] + lambda * ld * (1 / J) * C[i];
        const double da = da_of(p, e, q);
        /* gradient = delF_delU * P^T : (a,d) = sum_c G(a,c) P(d,c) */
        for (int a = 0; a < n; ++a)
            for (int d = 0; d < 3; ++d) {
                double s = 0;
                for (int c = 0; c < 3; ++c) s += G[a * 3 + c] * P[d * 3 + c];
                acc[a * 3 + d] += s * da;
            }
    }
    memcpy(out, acc, sizeof(double) * 3 * n); /* flattened a*3+d (:540-544) */
}

/* hat(x) y = x cross y (NeoHookeanElasticity.cpp:422-437) */
static void hat3(const double *x, double *h)
{
    h[0] = 0; h[1] = -x[2]; h[2] = x[1];
    h[3] = x[2]; h[4] = 0; h[5] = -x[0];
    h[6] = -x[1]; h[7] = x[0]; h[8] = 0;
}

static void put_block(double *M, int r0, int c0, const double *b, double s)
{
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) M[(r0 + r) * 9 + c0 + c] = s * b[r * 3 + c];
}

/* compute_energy_hessian_aux_fast<n,3> (NeoHookeanElasticity.cpp:547-658):
 * H9 = mu I + (mu + lambda(1-ln J))/J^2 g g^T + (lambda ln J - mu)/J d2J,
 * local H = D^T H9 D with D(c*3+j, a*3+j) = grad_t_m(a,c). */
static void nh_hessian(const orc_problem *p, int e, const double *x, double *H)
{
    const int n = p->n_local, N = 3 * n;
    double U[30], G[30], F[9], C[9];
    local_disp(p, e, x, U);
    const double lambda = p->lambda[e], mu = p->mu[e];
    memset(H, 0, sizeof(double) * N * N);
    for (int q = 0; q < p->n_quad; ++q) {
        nh_defgrad(p, e, q, U, G, F);
        const double J = det3(F);
        const double ld = log(J);
        cofactor_cols(F, C);
        double u[3] = {F[0], F[3], F[6]}, v[3] = {F[1], F[4], F[7]}, w[3] = {F[2], F[5], F[8]};
        double d2[81] = {0}, hb[9];
        hat3(v, hb); put_block(d2, 0, 6, hb, 1.0); put_block(d2, 6, 0, hb, -1.0);
        hat3(w, hb); put_block(d2, 0, 3, hb, -1.0); put_block(d2, 3, 0, hb, 1.0);
        hat3(u, hb); put_block(d2, 3, 6, hb, -1.0); put_block(d2, 6, 3, hb, 1.0);
        double gj[9]; /* column-major flatten of delJ_delF */
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) gj[c * 3 + r] = C[r * 3 + c];
        const double s1 = (mu + lambda * (1 - ld)) / (J * J);
        const double s2 = (lambda * ld - mu) / J;
        double H9[81];
        for (int i = 0; i < 9; ++i)
            for (int j = 0; j < 9; ++j)
                H9[i * 9 + j] = (i == j ? mu : 0.0) + s1 * gj[i] * gj[j] + s2 * d2[i * 9 + j];
        /* D (9 x N): column a*3+j has G(a, c) at row c*3+j */
        double T[9 * 30]; /* H9 * D */
        for (int i = 0; i < 9; ++i)
            for (int a = 0; a < n; ++a)
                for (int j = 0; j < 3; ++j) {
                    double s = 0;
                    for (int c = 0; c < 3; ++c) s += H9[i * 9 + c * 3 + j] * G[a * 3 + c];
                    T[i * N + a * 3 + j] = s;
                }
        const double da = da_of(p, e, q);
        for (int a = 0; a < n; ++a)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < N; ++k) {
                    double s = 0;
                    for (int c = 0; c < 3; ++c) s += G[a * 3 + c] * T[(c * 3 + j) * N + k];
                    H[(a * 3 + j) * N + k] += s * da;
                }
    }
}

/* --------------------------------------------------- linear elasticity */

/* disp_grad = (sum_a U(a,d) grad(a,c)) * jac_it (ElasticityUtils.hpp:106-135) */
static void le_disp_grad(const orc_problem *p, int e, int q, const double *U, double *D)
{
    const int n = p->n_local;
    const double *g = p->ref_grad + q * n * 3;
    const double *J = p->jac_it + ((size_t)e * p->n_quad + q) * 9;
    double A[9] = {0};
    for (int a = 0; a < n; ++a)
        for (int d = 0; d < 3; ++d)
            for (int c = 0; c < 3; ++c) A[d * 3 + c] += g[a * 3 + c] * U[a * 3 + d];
    for (int d = 0; d < 3; ++d)
        for (int c = 0; c < 3; ++c)
            D[d * 3 + c] = A[d * 3 + 0] * J[0 * 3 + c] + A[d * 3 + 1] * J[1 * 3 + c] + A[d * 3 + 2] * J[2 * 3 + c];
}

/* LinearElasticity::compute_energy_aux (LinearElasticity.cpp:105-134):
 * psi = mu tr(eps^T eps) + lambda/2 tr(eps)^2 */
static double le_energy(const orc_problem *p, int e, const double *x)
{
    double U[30], D[9];
    local_disp(p, e, x, U);
    const double lambda = p->lambda[e], mu = p->mu[e];
    double energy = 0;
    for (int q = 0; q < p->n_quad; ++q) {
        le_disp_grad(p, e, q, U, D);
        double eps[9], tr = 0, ee = 0;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) eps[r * 3 + c] = (D[r * 3 + c] + D[c * 3 + r]) / 2;
        for (int i = 0; i < 9; ++i) ee += eps[i] * eps[i];
        tr = eps[0] + eps[4] + eps[8];
        energy += (mu * ee + lambda / 2 * tr * tr) * da_of(p, e, q);
    }
    return energy;
}

/* Gradient of the energy above, which the reference takes by DScalar1
 * forward autodiff (LinearElasticity.cpp:70-86): dpsi/d(grad u) = 2 mu eps +
 * lambda tr(eps) I, contracted with grad_t_m. */
static void le_gradient(const orc_problem *p, int e, const double *x, double *out)
{
    const int n = p->n_local;
    double U[30], D[9], G[30];
    local_disp(p, e, x, U);
    const double lambda = p->lambda[e], mu = p->mu[e];
    memset(out, 0, sizeof(double) * 3 * n);
    for (int q = 0; q < p->n_quad; ++q) {
        le_disp_grad(p, e, q, U, D);
        phys_grad(p, e, q, G);
        double S[9];
        const double tr = D[0] + D[4] + D[8];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) S[r * 3 + c] = mu * (D[r * 3 + c] + D[c * 3 + r]) + (r == c ? lambda * tr : 0.0);
        const double da = da_of(p, e, q);
        for (int a = 0; a < n; ++a)
            for (int d = 0; d < 3; ++d) {
                double s = 0;
                for (int c = 0; c < 3; ++c) s += S[d * 3 + c] * G[a * 3 + c];
                out[a * 3 + d] += s * da;
            }
    }
}

/* Stiffness blocks of LinearElasticity::assemble (LinearElasticity.cpp:30-63),
 * scattered as in LinearAssembler::assemble (Assembler.cpp:209-250):
 * K[(i,m),(j,n)] = sum_q da (mu G_i[n] G_j[m] + lambda G_i[m] G_j[n] + delta_mn mu G_i.G_j) */
static void le_hessian(const orc_problem *p, int e, double *H)
{
    const int n = p->n_local, N = 3 * n;
    const double lambda = p->lambda[e], mu = p->mu[e];
    double G[30];
    memset(H, 0, sizeof(double) * N * N);
    for (int q = 0; q < p->n_quad; ++q) {
        phys_grad(p, e, q, G);
        const double da = da_of(p, e, q);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                const double *gi = G + 3 * i, *gj = G + 3 * j;
                const double dot = gi[0] * gj[0] + gi[1] * gj[1] + gi[2] * gj[2];
                for (int m = 0; m < 3; ++m)
                    for (int nn = 0; nn < 3; ++nn) {
                        double v = gi[nn] * gj[m] * mu + gi[m] * gj[nn] * lambda;
                        if (m == nn) v += mu * dot;
                        H[(i * 3 + m) * N + j * 3 + nn] += v * da;
                    }
            }
    }
}

/* --------------------------------------------------------- SaintVenant */

/* SaintVenantElasticity::compute_energy_aux (SaintVenantElasticity.cpp:235-278):
 * strain = (grad u^T grad u + grad u + grad u^T) / 2 (:9-20), Voigt strain
 * (E00, E11, E22, 2E12, 2E02, 2E01), stress = C strain with the isotropic C of
 * ElasticityTensor::set_from_lambda_mu (MatParams.cpp:211-254), stress tensor
 * [[s0 s5 s4] [s5 s1 s3] [s4 s3 s2]], psi = tr(stress strain) / 2. */
static void sv_strain(const double *D, double *Es)
{
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double s = D[r * 3 + c] + D[c * 3 + r];
            for (int k = 0; k < 3; ++k) s += D[k * 3 + r] * D[k * 3 + c];
            Es[r * 3 + c] = 0.5 * s;
        }
}

static void sv_stress_voigt(const double *Es, double lambda, double mu, double *S)
{
    const double eps[6] = {Es[0], Es[4], Es[8], 2 * Es[5], 2 * Es[2], 2 * Es[1]};
    double C[36] = {0};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C[i * 6 + j] = lambda + (i == j ? 2 * mu : 0.0);
    for (int i = 3; i < 6; ++i) C[i * 6 + i] = mu;
    double s[6];
    for (int i = 0; i < 6; ++i) {
        s[i] = 0;
        for (int j = 0; j < 6; ++j) s[i] += C[i * 6 + j] * eps[j];
    }
    const double T[9] = {s[0], s[5], s[4], s[5], s[1], s[3], s[4], s[3], s[2]};
    memcpy(S, T, sizeof(T));
}

static double sv_energy(const orc_problem *p, int e, const double *x)
{
    double U[30], D[9], Es[9], S[9];
    local_disp(p, e, x, U);
    double energy = 0;
    for (int q = 0; q < p->n_quad; ++q) {
        le_disp_grad(p, e, q, U, D);
        sv_strain(D, Es);
        sv_stress_voigt(Es, p->lambda[e], p->mu[e], S);
        double tr = 0; /* tr(S Es) */
        for (int r = 0; r < 3; ++r)
            for (int k = 0; k < 3; ++k) tr += S[r * 3 + k] * Es[k * 3 + r];
        energy += tr * da_of(p, e, q);
    }
    return 0.5 * energy;
}

/* The reference differentiates the energy by forward autodiff
 * (SaintVenantElasticity.cpp:89-131); restated: dpsi/dF = F S (S the stress
 * tensor above), contracted with grad_t_m. */
static void sv_gradient(const orc_problem *p, int e, const double *x, double *out)
{
    const int n = p->n_local;
    double U[30], D[9], G[30], Es[9], S[9];
    local_disp(p, e, x, U);
    memset(out, 0, sizeof(double) * 3 * n);
    for (int q = 0; q < p->n_quad; ++q) {
        le_disp_grad(p, e, q, U, D);
        phys_grad(p, e, q, G);
        sv_strain(D, Es);
        sv_stress_voigt(Es, p->lambda[e], p->mu[e], S);
        double P[9];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double s = S[r * 3 + c];
                for (int k = 0; k < 3; ++k) s += D[r * 3 + k] * S[k * 3 + c];
                P[r * 3 + c] = s;
            }
        const double da = da_of(p, e, q);
        for (int a = 0; a < n; ++a)
            for (int d = 0; d < 3; ++d) {
                double s = 0;
                for (int c = 0; c < 3; ++c) s += P[d * 3 + c] * G[a * 3 + c];
                out[a * 3 + d] += s * da;
            }
    }
}

/* Second derivative of the energy above: d2psi/dF_jc dF_kl =
 * delta_jk S_cl + lambda F_jc F_kl + mu delta_cl (F F^T)_jk + mu F_jl F_kc,
 * assembled as D^T A D with D(c*3+j, a*3+j) = grad_t_m(a,c). */
static void sv_hessian(const orc_problem *p, int e, const double *x, double *H)
{
    const int n = p->n_local, N = 3 * n;
    double U[30], D[9], G[30], Es[9], S[9], F[9], FF[9];
    local_disp(p, e, x, U);
    const double lambda = p->lambda[e], mu = p->mu[e];
    memset(H, 0, sizeof(double) * N * N);
    for (int q = 0; q < p->n_quad; ++q) {
        le_disp_grad(p, e, q, U, D);
        phys_grad(p, e, q, G);
        sv_strain(D, Es);
        sv_stress_voigt(Es, lambda, mu, S);
        for (int i = 0; i < 9; ++i) F[i] = D[i] + (i % 4 == 0 ? 1.0 : 0.0);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) FF[r * 3 + c] = F[r * 3 + 0] * F[c * 3 + 0] + F[r * 3 + 1] * F[c * 3 + 1] + F[r * 3 + 2] * F[c * 3 + 2];
        double A[81];
        for (int c = 0; c < 3; ++c)
            for (int j = 0; j < 3; ++j)
                for (int l = 0; l < 3; ++l)
                    for (int k = 0; k < 3; ++k)
                        A[(c * 3 + j) * 9 + l * 3 + k] = (j == k ? S[c * 3 + l] : 0.0) + lambda * F[j * 3 + c] * F[k * 3 + l] +
                                                         (c == l ? mu * FF[j * 3 + k] : 0.0) + mu * F[j * 3 + l] * F[k * 3 + c];
        double T[9 * 30];
        for (int i = 0; i < 9; ++i)
            for (int a = 0; a < n; ++a)
                for (int j = 0; j < 3; ++j) {
                    double s = 0;
                    for (int c = 0; c < 3; ++c) s += A[i * 9 + c * 3 + j] * G[a * 3 + c];
                    T[i * N + a * 3 + j] = s;
                }
        const double da = da_of(p, e, q);
        for (int a = 0; a < n; ++a)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < N; ++k) {
                    double s = 0;
                    for (int c = 0; c < 3; ++c) s += G[a * 3 + c] * T[(c * 3 + j) * N + k];
                    H[(a * 3 + j) * N + k] += s * da;
                }
    }
}

/* ------------------------------------------------------ public element API */

double orc_element_energy(const orc_problem *p, int e, const double *x)
{
    if (p->material == ORC_SAINT_VENANT) return sv_energy(p, e, x);
    return p->material == ORC_NEOHOOKEAN ? nh_energy(p, e, x) : le_energy(p, e, x);
}

void orc_element_gradient(const orc_problem *p, int e, const double *x, double *g)
{
    if (p->material == ORC_NEOHOOKEAN)
        nh_gradient(p, e, x, g);
    else if (p->material == ORC_SAINT_VENANT)
        sv_gradient(p, e, x, g);
    else
        le_gradient(p, e, x, g);
}

void orc_element_hessian(const orc_problem *p, int e, const double *x, double *H)
{
    if (p->material == ORC_NEOHOOKEAN)
        nh_hessian(p, e, x, H);
    else if (p->material == ORC_SAINT_VENANT)
        sv_hessian(p, e, x, H);
    else
        le_hessian(p, e, H);
}

/* ------------------------------------------------------------ assembly */

/* NLAssembler::assemble_energy (Assembler.cpp:495-531), one thread */
double orc_energy(const orc_problem *p, const double *x)
{
    double s = 0;
    for (int e = 0; e < p->n_elements; ++e) s += orc_element_energy(p, e, x);
    return s;
}

/* NLAssembler::assemble_energy_per_element (Assembler.cpp:533-572) */
void orc_energy_per_element(const orc_problem *p, const double *x, double *out)
{
    for (int e = 0; e < p->n_elements; ++e) out[e] = orc_element_energy(p, e, x);
}

/* NLAssembler::assemble_gradient (Assembler.cpp:574-643): rhs[node*3+m] += val[j*3+m] */
void orc_gradient(const orc_problem *p, const double *x, double *g)
{
    const int n = p->n_local;
    memset(g, 0, sizeof(double) * 3 * (size_t)p->n_nodes);
    double loc[30];
    for (int e = 0; e < p->n_elements; ++e) {
        orc_element_gradient(p, e, x, loc);
        for (int j = 0; j < n; ++j) {
            const size_t node = (size_t)p->elem_nodes[(size_t)e * n + j];
            for (int m = 0; m < 3; ++m) g[node * 3 + m] += loc[j * 3 + m];
        }
    }
}

/* ------------------------------------------------- triplets -> CSC */

typedef struct {
    int64_t key; /* col * n + row */
    double v;
} trip_t;

static int cmp_trip(const void *a, const void *b)
{
    const int64_t ka = ((const trip_t *)a)->key, kb = ((const trip_t *)b)->key;
    return (ka > kb) - (ka < kb);
}

/* Eigen::SparseMatrix::setFromTriplets semantics used by SparseMatrixCache
 * (MatrixCache.cpp:115-212): duplicates summed, inner indices sorted, explicit
 * zeros kept.  Stable counting sort by column, then by row within column. */
int64_t orc_triplets_to_csc(int n, int64_t n_trip, const int32_t *rows, const int32_t *cols,
                            const double *vals, int32_t *outer, int32_t *inner, double *values)
{
    trip_t *t = (trip_t *)malloc(sizeof(trip_t) * (size_t)(n_trip > 0 ? n_trip : 1));
    for (int64_t k = 0; k < n_trip; ++k) {
        t[k].key = (int64_t)cols[k] * n + rows[k];
        t[k].v = vals ? vals[k] : 0.0;
    }
    qsort(t, (size_t)n_trip, sizeof(trip_t), cmp_trip);
    int64_t nnz = 0;
    if (outer)
        for (int c = 0; c <= n; ++c) outer[c] = 0;
    for (int64_t k = 0; k < n_trip;) {
        int64_t k2 = k;
        double s = 0;
        while (k2 < n_trip && t[k2].key == t[k].key) s += t[k2++].v;
        if (inner) {
            inner[nnz] = (int32_t)(t[k].key % n);
            if (values) values[nnz] = s;
        }
        if (outer) outer[t[k].key / n + 1]++;
        nnz++;
        k = k2;
    }
    if (outer)
        for (int c = 0; c < n; ++c) outer[c + 1] += outer[c];
    free(t);
    return nnz;
}

/* Local Hessian scatter order of NLAssembler::assemble_hessian
 * (Assembler.cpp:712-752): for i, j, n, m -> (node_i*3+m, node_j*3+n). */
static int64_t hessian_triplets(const orc_problem *p, const double *x, int32_t *r, int32_t *c, double *v)
{
    const int n = p->n_local, N = 3 * n;
    double *H = v ? (double *)malloc(sizeof(double) * N * N) : NULL;
    int64_t k = 0;
    for (int e = 0; e < p->n_elements; ++e) {
        const int32_t *nodes = p->elem_nodes + (size_t)e * n;
        if (H) orc_element_hessian(p, e, x, H);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                for (int nn = 0; nn < 3; ++nn)
                    for (int m = 0; m < 3; ++m) {
                        r[k] = nodes[i] * 3 + m;
                        c[k] = nodes[j] * 3 + nn;
                        if (H) v[k] = H[(i * 3 + m) * N + j * 3 + nn];
                        ++k;
                    }
    }
    free(H);
    return k;
}

int64_t orc_hessian_nnz(const orc_problem *p)
{
    const int n = p->n_local;
    const int64_t nt = (int64_t)p->n_elements * 9 * n * n;
    int32_t *r = (int32_t *)malloc(sizeof(int32_t) * nt), *c = (int32_t *)malloc(sizeof(int32_t) * nt);
    hessian_triplets(p, NULL, r, c, NULL);
    const int64_t nnz = orc_triplets_to_csc(3 * p->n_nodes, nt, r, c, NULL, NULL, NULL, NULL);
    free(r);
    free(c);
    return nnz;
}

/* NLAssembler::assemble_hessian first call (Assembler.cpp:645-771), project_to_psd = false */
void orc_hessian(const orc_problem *p, const double *x, int32_t *outer, int32_t *inner, double *values)
{
    const int n = p->n_local;
    const int64_t nt = (int64_t)p->n_elements * 9 * n * n;
    int32_t *r = (int32_t *)malloc(sizeof(int32_t) * nt), *c = (int32_t *)malloc(sizeof(int32_t) * nt);
    double *v = (double *)malloc(sizeof(double) * nt);
    hessian_triplets(p, x, r, c, v);
    orc_triplets_to_csc(3 * p->n_nodes, nt, r, c, v, outer, inner, values);
    free(r);
    free(c);
    free(v);
}

/* ------------------------------------------------------- adjoint terms */

/* interpolate_at_local_vals gradient (io/Evaluator.cpp:727-757):
 * grad(d,c) = sum_a f[node_a*3+d] grad_t_m(a,c) */
static void interp_grad(const orc_problem *p, int e, const double *G, const double *f, double *out)
{
    const int n = p->n_local;
    memset(out, 0, sizeof(double) * 9);
    for (int a = 0; a < n; ++a) {
        const size_t node = (size_t)p->elem_nodes[(size_t)e * n + a];
        for (int d = 0; d < 3; ++d)
            for (int c = 0; c < 3; ++c) out[d * 3 + c] += f[node * 3 + d] * G[a * 3 + c];
    }
}

/* force_material_derivative (ElasticForceDerivative.cpp:97-129) with
 * compute_dstress_dmu_dlambda (NeoHookeanElasticity.cpp:747-762,
 * LinearElasticity.cpp:290-299): term[e] = -sum da <dP/dlambda, grad p>,
 * term[e+E] = -sum da <dP/dmu, grad p>. */
void orc_material_vjp(const orc_problem *p, const double *x, const double *adjoint, double *term)
{
    const int E = p->n_elements;
    double G[30], gu[9], gp[9];
    for (int e = 0; e < E; ++e) {
        double tl = 0, tm = 0;
        for (int q = 0; q < p->n_quad; ++q) {
            phys_grad(p, e, q, G);
            interp_grad(p, e, G, x, gu);
            interp_grad(p, e, G, adjoint, gp);
            double dmu[9], dla[9];
            if (p->material == ORC_NEOHOOKEAN) {
                double F[9], Fi[9], FmT[9];
                for (int i = 0; i < 9; ++i) F[i] = gu[i] + (i % 4 == 0 ? 1.0 : 0.0);
                inv3(F, Fi);
                transpose3(Fi, FmT);
                const double ld = log(det3(F));
                for (int i = 0; i < 9; ++i) {
                    dmu[i] = F[i] - FmT[i];
                    dla[i] = ld * FmT[i];
                }
            } else if (p->material == ORC_SAINT_VENANT) {
                /* no counterpart in the reference: the derivatives of P = F S
                 * in lambda and mu, dP/dlambda = tr(E) F, dP/dmu = 2 F E */
                double Es[9], F[9];
                sv_strain(gu, Es);
                for (int i = 0; i < 9; ++i) F[i] = gu[i] + (i % 4 == 0 ? 1.0 : 0.0);
                const double tr = Es[0] + Es[4] + Es[8];
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) {
                        double fe = 0;
                        for (int k = 0; k < 3; ++k) fe += F[r * 3 + k] * Es[k * 3 + c];
                        dmu[r * 3 + c] = 2 * fe;
                        dla[r * 3 + c] = tr * F[r * 3 + c];
                    }
            } else {
                const double tr = gu[0] + gu[4] + gu[8];
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) {
                        dmu[r * 3 + c] = gu[c * 3 + r] + gu[r * 3 + c];
                        dla[r * 3 + c] = r == c ? tr : 0.0;
                    }
            }
            double im = 0, il = 0;
            for (int i = 0; i < 9; ++i) {
                im += dmu[i] * gp[i];
                il += dla[i] * gp[i];
            }
            const double da = da_of(p, e, q);
            tm += -im * da;
            tl += -il * da;
        }
        term[e] = tl;
        term[e + E] = tm;
    }
}

/* compute_stress_grad_multiply_mat (NeoHookeanElasticity.cpp:660-680,
 * LinearElasticity.cpp:253-270) */
static void stress_grad_mul(const orc_problem *p, int e, const double *gu, const double *mat, double *stress, double *res)
{
    const double lambda = p->lambda[e], mu = p->mu[e];
    if (p->material == ORC_NEOHOOKEAN) {
        double F[9], Fi[9], FmT[9];
        for (int i = 0; i < 9; ++i) F[i] = gu[i] + (i % 4 == 0 ? 1.0 : 0.0);
        inv3(F, Fi);
        transpose3(Fi, FmT);
        const double ld = log(det3(F));
        for (int i = 0; i < 9; ++i) stress[i] = mu * (F[i] - FmT[i]) + lambda * ld * FmT[i];
        double fm = 0;
        for (int i = 0; i < 9; ++i) fm += FmT[i] * mat[i];
        double t1[9], t2[9]; /* FmT * mat^T * FmT */
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double s = 0;
                for (int k = 0; k < 3; ++k) s += FmT[r * 3 + k] * mat[c * 3 + k];
                t1[r * 3 + c] = s;
            }
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double s = 0;
                for (int k = 0; k < 3; ++k) s += t1[r * 3 + k] * FmT[k * 3 + c];
                t2[r * 3 + c] = s;
            }
        for (int i = 0; i < 9; ++i) res[i] = mu * mat[i] + t2[i] * (mu - lambda * ld) + lambda * fm * FmT[i];
    } else if (p->material == ORC_SAINT_VENANT) {
        /* no counterpart in the reference: P = F S and its derivative along
         * mat, mat S + F (lambda tr(dE) I + 2 mu dE), dE = (mat^T F + F^T mat) / 2 */
        double Es[9], S[9], F[9], dE[9];
        sv_strain(gu, Es);
        sv_stress_voigt(Es, lambda, mu, S);
        for (int i = 0; i < 9; ++i) F[i] = gu[i] + (i % 4 == 0 ? 1.0 : 0.0);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double s = 0;
                for (int k = 0; k < 3; ++k) s += mat[k * 3 + r] * F[k * 3 + c] + F[k * 3 + r] * mat[k * 3 + c];
                dE[r * 3 + c] = 0.5 * s;
            }
        const double trd = dE[0] + dE[4] + dE[8];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double ps = 0, ms = 0, fd = 0;
                for (int k = 0; k < 3; ++k) {
                    ps += F[r * 3 + k] * S[k * 3 + c];
                    ms += mat[r * 3 + k] * S[k * 3 + c];
                    fd += F[r * 3 + k] * (2 * mu * dE[k * 3 + c] + (k == c ? lambda * trd : 0.0));
                }
                stress[r * 3 + c] = ps;
                res[r * 3 + c] = ms + fd;
            }
    } else {
        const double tr = gu[0] + gu[4] + gu[8], tm = mat[0] + mat[4] + mat[8];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                stress[r * 3 + c] = mu * (gu[r * 3 + c] + gu[c * 3 + r]) + (r == c ? lambda * tr : 0.0);
                res[r * 3 + c] = mu * (mat[r * 3 + c] + mat[c * 3 + r]) + (r == c ? lambda * tm : 0.0);
            }
    }
}

/* force_shape_derivative, elastic branch (ElasticForceDerivative.cpp:208-259) */
void orc_shape_vjp(const orc_problem *p, const double *x, const double *adjoint, double *term)
{
    memset(term, 0, sizeof(double) * 3 * (size_t)p->n_vertices);
    double G[30], gu[9], gp[9];
    for (int e = 0; e < p->n_elements; ++e) {
        for (int q = 0; q < p->n_quad; ++q) {
            phys_grad(p, e, q, G);
            interp_grad(p, e, G, x, gu);
            interp_grad(p, e, G, adjoint, gp);
            const double *J = p->jac_it + ((size_t)e * p->n_quad + q) * 9;
            const double da = da_of(p, e, q);
            for (int v = 0; v < 4; ++v) {
                const double *g = p->geo_grad + q * 12 + v * 3;
                double gv[3];
                for (int c = 0; c < 3; ++c) gv[c] = g[0] * J[0 * 3 + c] + g[1] * J[1 * 3 + c] + g[2] * J[2 * 3 + c];
                const size_t vid = (size_t)p->elem_vertices[(size_t)e * 4 + v];
                for (int d = 0; d < 3; ++d) {
                    double GV[9] = {0};
                    for (int c = 0; c < 3; ++c) GV[d * 3 + c] = gv[c];
                    double mat[9]; /* grad_u * grad_v */
                    for (int r = 0; r < 3; ++r)
                        for (int c = 0; c < 3; ++c) mat[r * 3 + c] = gu[r * 3 + d] * GV[d * 3 + c];
                    double stress[9], fp[9];
                    stress_grad_mul(p, e, gu, mat, stress, fp);
                    const double trv = GV[0] + GV[4] + GV[8];
                    double s = 0;
                    for (int r = 0; r < 3; ++r)
                        for (int c = 0; c < 3; ++c) {
                            double tmp = 0; /* stress * GV^T */
                            for (int k = 0; k < 3; ++k) tmp += stress[r * 3 + k] * GV[c * 3 + k];
                            tmp -= trv * stress[r * 3 + c];
                            s += (fp[r * 3 + c] + tmp) * gp[r * 3 + c];
                        }
                    term[vid * 3 + d] -= s * da;
                }
            }
        }
    }
}

/* ============================================ reference-algorithm baseline */

/* Per-element copy of the geometry/basis cache, as AssemblyValsCache::compute
 * deep-copies ElementAssemblyValues on every element of every call
 * (AssemblyValsCache.cpp:52-67). */
typedef struct {
    double grad_t_m[64 * 30];
    double jac_it[64 * 9];
    double det[64];
    double da[64];
} elem_vals;

struct orc_baseline {
    const orc_problem *p;
    int n_threads;
    double *vals_cache; /* [e][q][n][3] grad_t_m, precomputed like AssemblyValsCache::init */
    /* SparseMatrixCache state (MatrixCache.cpp:134-212) */
    int64_t nnz;
    int32_t *outer, *inner;
    int64_t *slot_offset; /* per element start into slots */
    int32_t *slots;       /* second_cache_[e][k] */
    double **thread_values;
};

orc_baseline *orc_baseline_create(const orc_problem *p, int n_threads)
{
    orc_baseline *b = (orc_baseline *)calloc(1, sizeof(orc_baseline));
    b->p = p;
    b->n_threads = n_threads > 0 ? n_threads : 1;
    const int n = p->n_local, nq = p->n_quad;
    b->vals_cache = (double *)malloc(sizeof(double) * (size_t)p->n_elements * nq * n * 3);
    for (int e = 0; e < p->n_elements; ++e)
        for (int q = 0; q < nq; ++q) phys_grad(p, e, q, b->vals_cache + ((size_t)e * nq + q) * n * 3);
    return b;
}

void orc_baseline_destroy(orc_baseline *b)
{
    if (!b) return;
    free(b->vals_cache);
    free(b->outer);
    free(b->inner);
    free(b->slot_offset);
    free(b->slots);
    if (b->thread_values)
        for (int t = 0; t < b->n_threads; ++t) free(b->thread_values[t]);
    free(b->thread_values);
    free(b);
}

int orc_baseline_threads(const orc_baseline *b) { return b->n_threads; }

static void load_vals(const orc_baseline *b, int e, elem_vals *v)
{
    const orc_problem *p = b->p;
    const int n = p->n_local, nq = p->n_quad;
    memcpy(v->grad_t_m, b->vals_cache + (size_t)e * nq * n * 3, sizeof(double) * nq * n * 3);
    memcpy(v->jac_it, p->jac_it + (size_t)e * nq * 9, sizeof(double) * nq * 9);
    memcpy(v->det, p->det + (size_t)e * nq, sizeof(double) * nq);
    for (int q = 0; q < nq; ++q) v->da[q] = v->det[q] * p->qw[q];
}

typedef struct {
    orc_baseline *b;
    int t;
    const double *x;
    double scalar;
    double sink; /* keeps the per-element copy observable */
    double *vec;
    int mode; /* 0 energy, 1 gradient, 2 hessian values */
} job_t;

/* static chunks [t*E/T, (t+1)*E/T) of the CPP backend (utils/par_for.cpp:10-31) */
static void *worker(void *arg)
{
    job_t *j = (job_t *)arg;
    orc_baseline *b = j->b;
    const orc_problem *p = b->p;
    const int E = p->n_elements, T = b->n_threads, n = p->n_local, N = 3 * n;
    const int start = (int)((int64_t)j->t * E / T), end = (int)((int64_t)(j->t + 1) * E / T);
    elem_vals v;
    double loc[30];
    double *H = (double *)malloc(sizeof(double) * N * N);
    for (int e = start; e < end; ++e) {
        load_vals(b, e, &v);
        j->sink += v.da[0];
        if (j->mode == 0) {
            j->scalar += orc_element_energy(p, e, j->x);
        } else if (j->mode == 1) {
            orc_element_gradient(p, e, j->x, loc);
            for (int a = 0; a < n; ++a) {
                const size_t node = (size_t)p->elem_nodes[(size_t)e * n + a];
                for (int m = 0; m < 3; ++m) j->vec[node * 3 + m] += loc[a * 3 + m];
            }
        } else {
            orc_element_hessian(p, e, j->x, H);
            const int32_t *s = b->slots + b->slot_offset[e];
            int k = 0;
            for (int i = 0; i < n; ++i)
                for (int jj = 0; jj < n; ++jj)
                    for (int nn = 0; nn < 3; ++nn)
                        for (int m = 0; m < 3; ++m) j->vec[s[k++]] += H[(i * 3 + m) * N + jj * 3 + nn];
        }
    }
    free(H);
    return NULL;
}

static void run_jobs(orc_baseline *b, job_t *jobs)
{
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * b->n_threads);
    for (int t = 0; t < b->n_threads; ++t) pthread_create(&th[t], NULL, worker, &jobs[t]);
    for (int t = 0; t < b->n_threads; ++t) pthread_join(th[t], NULL);
    free(th);
}

/* NLAssembler::assemble_energy with per-thread scalars merged serially (Assembler.cpp:526-530) */
double orc_baseline_energy(orc_baseline *b, const double *x)
{
    job_t *jobs = (job_t *)calloc(b->n_threads, sizeof(job_t));
    for (int t = 0; t < b->n_threads; ++t) {
        jobs[t].b = b; jobs[t].t = t; jobs[t].x = x; jobs[t].mode = 0;
    }
    run_jobs(b, jobs);
    double s = 0;
    for (int t = 0; t < b->n_threads; ++t) s += jobs[t].scalar;
    free(jobs);
    return s;
}

/* per-thread dense vectors merged serially (Assembler.cpp:585-642) */
void orc_baseline_gradient(orc_baseline *b, const double *x, double *g)
{
    const size_t ndof = 3 * (size_t)b->p->n_nodes;
    job_t *jobs = (job_t *)calloc(b->n_threads, sizeof(job_t));
    for (int t = 0; t < b->n_threads; ++t) {
        jobs[t].b = b; jobs[t].t = t; jobs[t].x = x; jobs[t].mode = 1;
        jobs[t].vec = (double *)calloc(ndof, sizeof(double));
    }
    run_jobs(b, jobs);
    memset(g, 0, sizeof(double) * ndof);
    for (int t = 0; t < b->n_threads; ++t) {
        for (size_t i = 0; i < ndof; ++i) g[i] += jobs[t].vec[i];
        free(jobs[t].vec);
    }
    free(jobs);
}

/* Hessian: the first call assembles triplets into CSC and builds the
 * per-element slot map by a linear search of each column's entries
 * (MatrixCache.cpp:134-212); later calls add into per-thread full-length
 * value arrays through the slot map and sum them (MatrixCache.cpp:88-113,
 * 289-322). */
int64_t orc_baseline_hessian(orc_baseline *b, const double *x, double *values)
{
    const orc_problem *p = b->p;
    const int n = p->n_local, N = 3 * n, ndof = 3 * p->n_nodes;
    if (!b->outer) {
        const int64_t nt = (int64_t)p->n_elements * N * N;
        int32_t *r = (int32_t *)malloc(sizeof(int32_t) * nt), *c = (int32_t *)malloc(sizeof(int32_t) * nt);
        double *v = (double *)malloc(sizeof(double) * nt);
        hessian_triplets(p, x, r, c, v);
        b->nnz = orc_triplets_to_csc(ndof, nt, r, c, NULL, NULL, NULL, NULL);
        b->outer = (int32_t *)malloc(sizeof(int32_t) * (ndof + 1));
        b->inner = (int32_t *)malloc(sizeof(int32_t) * b->nnz);
        double *first = (double *)malloc(sizeof(double) * b->nnz);
        orc_triplets_to_csc(ndof, nt, r, c, v, b->outer, b->inner, first);
        b->slot_offset = (int64_t *)malloc(sizeof(int64_t) * (p->n_elements + 1));
        b->slots = (int32_t *)malloc(sizeof(int32_t) * nt);
        for (int e = 0; e <= p->n_elements; ++e) b->slot_offset[e] = (int64_t)e * N * N;
        /* mapping_[row] lists (col, index); the search finds the entry (row, col) */
        for (int64_t k = 0; k < nt; ++k) {
            const int32_t row = r[k], col = c[k];
            int32_t idx = -1;
            for (int32_t ii = b->outer[col]; ii < b->outer[col + 1]; ++ii)
                if (b->inner[ii] == row) { idx = ii; break; }
            b->slots[k] = idx;
        }
        if (values) memcpy(values, first, sizeof(double) * b->nnz);
        free(first);
        free(r);
        free(c);
        free(v);
        b->thread_values = (double **)calloc(b->n_threads, sizeof(double *));
        for (int t = 0; t < b->n_threads; ++t) b->thread_values[t] = (double *)malloc(sizeof(double) * b->nnz);
        return b->nnz;
    }
    job_t *jobs = (job_t *)calloc(b->n_threads, sizeof(job_t));
    for (int t = 0; t < b->n_threads; ++t) {
        memset(b->thread_values[t], 0, sizeof(double) * b->nnz);
        jobs[t].b = b; jobs[t].t = t; jobs[t].x = x; jobs[t].mode = 2;
        jobs[t].vec = b->thread_values[t];
    }
    run_jobs(b, jobs);
    if (values) {
        memcpy(values, b->thread_values[0], sizeof(double) * b->nnz);
        for (int t = 1; t < b->n_threads; ++t)
            for (int64_t i = 0; i < b->nnz; ++i) values[i] += b->thread_values[t][i];
    }
    free(jobs);
    return b->nnz;
}
