/*
 * FP64 CPU oracle for PolyFEM's nonlinear-elasticity assembly path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a restatement, in plain C, of the
 * reference's arithmetic for the NeoHookean and linear-elastic materials on
 * P1/P2 tetrahedra.  It is the checker for the HIP library and the CPU
 * baseline timed beside it; the product never links, loads or calls it.
 *
 * Each function cites the reference file:line it follows.  The reference
 * itself cannot be compiled here (Eigen, oneTBB and polysolve are fetched over
 * the network by its build), so the restatement is pinned by the
 * reference's own known-answer tests and by mathematics (see the tests).
 */
#ifndef PF_ORACLE_H
#define PF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_NEOHOOKEAN = 0, ORC_LINEAR_ELASTICITY = 1, ORC_SAINT_VENANT = 2 };

typedef struct orc_problem orc_problem;

/* Build the per-element geometry cache (AssemblyValsCache::init,
 * AssemblyValsCache.cpp:11-32; ElementAssemblyValues::finalize3d,
 * ElementAssemblyValues.cpp:65-104).  quad_order <= 0 selects the default
 * max(2(p-1),1) (AssemblerUtils.cpp:225-228).  Arrays are borrowed and must
 * outlive the problem. */
orc_problem *orc_create(int order, int quad_order, int material,
                        int n_elements, const int32_t *elem_nodes, int n_nodes,
                        int n_vertices, const double *vertices, const int32_t *elem_vertices,
                        const double *lambda, const double *mu);
void orc_destroy(orc_problem *p);
int orc_n_quad(const orc_problem *p);
int orc_n_local(const orc_problem *p);

/* Reference-element tables. */
int orc_quadrature(int quad_order, double *points /*nq*3*/, double *weights /*nq*/);
void orc_basis_grad(int order, const double xi[3], double *grad /*n*3*/);
void orc_basis_val(int order, const double xi[3], double *val /*n*/);
/* cached J^{-T} (row-major 3x3) and det J at each quadrature point */
void orc_geometry(const orc_problem *p, int e, double *jac_it /*nq*9*/, double *det /*nq*/);

/* Per-element kernels (local ordering a*3+d). */
double orc_element_energy(const orc_problem *p, int e, const double *x);
void orc_element_gradient(const orc_problem *p, int e, const double *x, double *g /*3n*/);
void orc_element_hessian(const orc_problem *p, int e, const double *x, double *H /*3n x 3n row-major*/);

/* Whole-mesh assembly, single thread, element order. */
double orc_energy(const orc_problem *p, const double *x);
void orc_energy_per_element(const orc_problem *p, const double *x, double *out);
void orc_gradient(const orc_problem *p, const double *x, double *g);
int64_t orc_hessian_nnz(const orc_problem *p);
/* CSC (== CSR, the pattern is symmetric): outer[ndof+1], inner[nnz], values[nnz] */
void orc_hessian(const orc_problem *p, const double *x, int32_t *outer, int32_t *inner, double *values);

/* Adjoint terms (ElasticForceDerivative.cpp:35-264). */
void orc_material_vjp(const orc_problem *p, const double *x, const double *adjoint, double *term /*2E*/);
void orc_shape_vjp(const orc_problem *p, const double *x, const double *adjoint, double *term /*3 Nv*/);

/* SparseMatrixCache restatement (MatrixCache.cpp:88-228): scalar triplets ->
 * sorted CSC with duplicate sums.  Returns nnz; pass NULL outputs to count. */
int64_t orc_triplets_to_csc(int n, int64_t n_trip, const int32_t *rows, const int32_t *cols,
                            const double *vals, int32_t *outer, int32_t *inner, double *values);

/* Reference-algorithm CPU baseline (Assembler.cpp:495-771 with the CPP
 * threading backend, utils/par_for.cpp:10-31).  A baseline object keeps the
 * pattern and the per-element slot map between Hessian calls, as
 * SparseMatrixCache does.  Times are seconds of wall clock. */
typedef struct orc_baseline orc_baseline;
orc_baseline *orc_baseline_create(const orc_problem *p, int n_threads);
void orc_baseline_destroy(orc_baseline *b);
int orc_baseline_threads(const orc_baseline *b);
double orc_baseline_energy(orc_baseline *b, const double *x);
void orc_baseline_gradient(orc_baseline *b, const double *x, double *g);
/* first call builds pattern + slot map from triplets; later calls reuse them */
int64_t orc_baseline_hessian(orc_baseline *b, const double *x, double *values /*nnz or NULL*/);

#ifdef __cplusplus
}
#endif
#endif
