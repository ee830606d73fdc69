/*
 * pfgpu — MI355X-native nonlinear-elasticity assembly for PolyFEM.
 *
 * C-ABI of the GPU library.  Plain pointers and sizes only; no exceptions
 * cross this boundary.  Every entry point returns a pfgpu_status and leaves a
 * message for pfgpu_last_error() (thread-local) when it fails.
 *
 * The boundary replaces PolyFEM's in-process assembler API
 * (src/polyfem/assembler/Assembler.hpp:53-376): an ElasticityNLAssembler
 * adapter forwards its overrides to these entry points.  Conventions follow
 * the reference:
 *   - dof layout is node-major and interleaved, dof = node*3 + d
 *     (Assembler.cpp:624,729,734);
 *   - element-local ordering is a*3 + d (NeoHookeanElasticity.cpp:540-544);
 *   - P2 local node order is v0..v3, e01, e12, e20, e03, e13, e23
 *     (autogen/auto_p_bases.cpp:1437-1448);
 *   - the Hessian is returned as sorted CSC with int32 indices, the
 *     StiffnessMatrix type (utils/Types.hpp:21-25); the pattern is symmetric,
 *     so outer/inner are also the CSR row pointers / column indices;
 *   - the handle owns the sparsity pattern and the scatter maps that the
 *     caller-owned, stateful MatrixCache carries in the reference
 *     (utils/MatrixCache.hpp:43-134).
 *
 * Threading: one handle belongs to one HIP stream; calls on one handle are
 * serialised by the caller; distinct handles are independent.  Host-pointer
 * entry points are synchronous.  *_dev entry points take device pointers,
 * enqueue on the handle's stream and return without synchronising.
 */
#ifndef PFGPU_H
#define PFGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    PFGPU_OK = 0,
    PFGPU_ERR_INVALID_ARGUMENT = 1,
    PFGPU_ERR_HIP = 2,
    /* a non-finite energy, gradient or Hessian entry (an inverted element):
     * outputs hold NaN, as the reference's do (ElasticForm.cpp:388-405) */
    PFGPU_ERR_INVALID_STATE = 3,
    PFGPU_ERR_UNSUPPORTED = 4,
    PFGPU_ERR_OUT_OF_MEMORY = 5
} pfgpu_status;

/* SaintVenant: psi = lambda/2 tr(E)^2 + mu E:E with E = (F^T F - I)/2, the
 * isotropic case of SaintVenantElasticity (SaintVenantElasticity.cpp:235-278);
 * the plain path only */
/* FixedCorotational: psi = mu |F - R|^2 + lambda/2 (J - 1)^2 with R the
 * rotation of F (FixedCorotational.cpp:589-706); elements with det F <= 0
 * are reported like inverted NeoHookean elements; the plain path only */
/* MooneyRivlin: psi = c1 (I1~ - 3) + c2 (I2~ - 3) + k/2 ln^2 J, the invariants
 * of C~ = F~ F~^T with F~ = J^{-1/3} F (MooneyRivlinElasticity.hpp:25-44, a
 * GenericElastic material, GenericElastic.hpp:93-348); per-element c1, c2,
 * k; the plain path only; no material (Lame) derivatives, as in the
 * reference */
typedef enum {
    PFGPU_NEOHOOKEAN = 0,
    PFGPU_LINEAR_ELASTICITY = 1,
    PFGPU_SAINT_VENANT = 2,
    PFGPU_FIXED_COROTATIONAL = 3,
    PFGPU_MOONEY_RIVLIN = 4
} pfgpu_material_kind;

/* FP64: every kernel in double.  MIXED: element kernels in float; the energy
 * sum, gradient accumulation, norms and CG dot products in double; Hessian
 * values are float. */
typedef enum { PFGPU_FP64 = 0, PFGPU_MIXED = 1 } pfgpu_precision;

/* Assembly algorithm.  PLAIN: element-local results written to HBM, then
 * gathered into the global vector / CSR in a fixed order (the default for
 * SaintVenant and FixedCorotational).  FACTORED (the default for NeoHookean
 * and linear elasticity): PLAIN, except that the Hessian's element blocks
 * are not stored: per-element factors (per point F^{-T} J^{-T} and two
 * curvature weights, and the reference metric) are written instead and the
 * CSR gather recomputes each block from them.  FUSED: patch-owned rows
 * accumulated on chip (LDS) in a fixed order, one persistent launch.  All
 * three are deterministic and give the same results to rounding.  FUSED and
 * FACTORED cover NeoHookean and linear elasticity. */
typedef enum { PFGPU_PATH_FUSED = 0, PFGPU_PATH_PLAIN = 1, PFGPU_PATH_FACTORED = 2 } pfgpu_path;

typedef struct {
    int32_t order;            /* 1 (P1) or 2 (P2) */
    int32_t quadrature_order; /* <= 0: max(2(p-1),1) (AssemblerUtils.cpp:225-228); 1 to 4 (3 and 4: plain path) */
    int32_t n_elements;
    int32_t n_nodes;             /* discretisation nodes; n_dofs = 3 n_nodes */
    const int32_t *elem_nodes;   /* n_elements x (4|10), PolyFEM local order */
    int32_t n_vertices;          /* geometric vertices */
    const double *vertices;      /* n_vertices x 3 */
    const int32_t *elem_vertices;/* n_elements x 4: the P1 geometry */
} pfgpu_mesh_desc;

typedef struct {
    int32_t kind;          /* pfgpu_material_kind */
    const double *lambda;  /* per-element Lame lambda (n_elements), or NULL */
    const double *mu;      /* per-element Lame mu (n_elements), or NULL */
    double young;          /* uniform E when lambda/mu are NULL */
    double poisson;        /* uniform nu when lambda/mu are NULL */
    /* MooneyRivlin only (the Lame fields are then unused): per-element c1,
     * c2, k (n_elements each, all three or none), else the uniform values */
    const double *c1, *c2, *k;
    double c1_uniform, c2_uniform, k_uniform;
} pfgpu_material_desc;

typedef struct {
    int32_t order, n_local, n_quad, material, precision, path;
    int32_t n_elements, n_nodes, n_vertices, n_dofs;
    int64_t nnz;           /* Hessian non-zeros */
    int64_t device_bytes;  /* device memory held by the handle */
} pfgpu_info_t;

typedef struct pfgpu_handle pfgpu_handle;

/* ---- lifetime --------------------------------------------------------- */

/* Builds the geometry cache (AssemblyValsCache::init, AssemblyValsCache.cpp:11-32),
 * the Lame parameters (LameParameters::lambda_mu, MatParams.cpp:368-401; the
 * per-element override of update_lame_params, NeoHookeanElasticity.hpp:58-62),
 * the Hessian pattern (SparseMatrixCache first assembly, MatrixCache.cpp:134-212)
 * and the scatter maps; uploads everything to the current device. */
pfgpu_status pfgpu_create(const pfgpu_mesh_desc *mesh, const pfgpu_material_desc *material,
                          int32_t precision, pfgpu_handle **out);
pfgpu_status pfgpu_destroy(pfgpu_handle *h);
const char *pfgpu_last_error(void);
pfgpu_status pfgpu_info(const pfgpu_handle *h, pfgpu_info_t *info);
/* hipStream_t passed as void*; NULL selects the default stream */
pfgpu_status pfgpu_set_stream(pfgpu_handle *h, void *stream);
pfgpu_status pfgpu_set_path(pfgpu_handle *h, int32_t path);
/* Launch-shape autotuner (no reference counterpart): times every candidate
 * configuration of the energy + gradient + Hessian assembly on this handle's
 * mesh at state x (device pointer; 3 timed runs each, after one warm-up) and
 * keeps the fastest.  Candidates: the fused path, the plain path with
 * block-gather workgroups of 128/256/512 threads and Hessian-kernel
 * workgroups of 64/128/256 threads, and the factored path with gather
 * workgroups of 64/128/256/512 threads.  All candidates give identical results
 * up to the summation order of the path. */
typedef struct {
    int32_t path, gather_threads, hessian_threads; /* chosen configuration */
    int32_t candidates;                            /* configurations timed */
    double ms;                                     /* best time, E+G+H */
    double ms_plain_default;                       /* plain, 256/256 */
    double ms_fused;                               /* fused (negative: not available) */
    double ms_factored;                            /* best factored (negative: not available) */
} pfgpu_tune_report;
pfgpu_status pfgpu_autotune(pfgpu_handle *h, const double *x, pfgpu_tune_report *report);
/* sets a configuration directly (gather_threads in {64,128,256,512},
 * hessian_threads in {64,128,256}) */
pfgpu_status pfgpu_set_tuning(pfgpu_handle *h, int32_t path, int32_t gather_threads, int32_t hessian_threads);
/* replaces ElasticityAssembler::update_lame_params (Assembler.hpp:147-156) */
pfgpu_status pfgpu_set_lame(pfgpu_handle *h, const double *lambda, const double *mu);
/* MooneyRivlin parameters per element (n_elements each), as the reference's
 * c1 / c2 / k material parameters (MooneyRivlinElasticity.cpp:10-15) */
pfgpu_status pfgpu_set_mooney_rivlin(pfgpu_handle *h, const double *c1, const double *c2, const double *k);
/* new rest positions of the mesh vertices (n_vertices x 3, same topology):
 * replaces Mesh::set_point over the vertices followed by the rebuild of the
 * assembly-values cache (optimization/var2sims/ShapeVariableToSimulation.cpp:84-95,
 * State.cpp:998-1000); the mass matrix, if set, is rebuilt with the same density */
pfgpu_status pfgpu_set_vertices(pfgpu_handle *h, const double *vertices);

/* ---- assembly, host pointers (synchronous) ------------------------------ */

/* NLAssembler::assemble_energy (Assembler.cpp:495-531) */
pfgpu_status pfgpu_value(pfgpu_handle *h, const double *x, double *energy);
/* NLAssembler::assemble_energy_per_element (Assembler.cpp:533-572) */
pfgpu_status pfgpu_value_per_element(pfgpu_handle *h, const double *x, double *out);
/* NLAssembler::assemble_gradient (Assembler.cpp:574-643); g has n_dofs entries */
pfgpu_status pfgpu_gradient(pfgpu_handle *h, const double *x, double *g);
pfgpu_status pfgpu_hessian_nnz(const pfgpu_handle *h, int64_t *nnz);
/* the CSC pattern of NLAssembler::assemble_hessian (Assembler.cpp:645-771) */
pfgpu_status pfgpu_hessian_pattern(const pfgpu_handle *h, int32_t *outer, int32_t *inner);
/* Hessian values in pattern order: double* for FP64, float* for MIXED.
 * project_to_psd != 0: every element's local Hessian is projected onto the
 * positive semidefinite cone before the sum, as NLAssembler::assemble_hessian
 * does with ipc::project_to_psd (eigenvalues <= 0 clamped to 0).  FP64: by
 * eigendecomposition (a positive definite matrix is left as it is); MIXED: by
 * the matrix sign function, which agrees with it to about 1e-6 of the element
 * matrix's largest entry. */
pfgpu_status pfgpu_hessian_values(pfgpu_handle *h, const double *x, int32_t project_to_psd, void *values);
/* energy, gradient and Hessian values from one pass over the mesh; any
 * output pointer may be NULL */
pfgpu_status pfgpu_energy_gradient_hessian(pfgpu_handle *h, const double *x, double *energy,
                                           double *g, void *values);
/* y = H(x) v without forming H (no reference counterpart; the reference
 * multiplies the assembled matrix) */
pfgpu_status pfgpu_hvp(pfgpu_handle *h, const double *x, const double *v, double *y);
/* ElasticForceDerivative::force_material_derivative (ElasticForceDerivative.cpp:35-132):
 * out[e] = -sum_q da <dP/dlambda, grad p>, out[E+e] = -sum_q da <dP/dmu, grad p> */
pfgpu_status pfgpu_material_vjp(pfgpu_handle *h, const double *x, const double *p, double *out);
/* ElasticForceDerivative::force_shape_derivative (ElasticForceDerivative.cpp:135-264),
 * out has 3 n_vertices entries, vertex*3 + d */
pfgpu_status pfgpu_shape_vjp(pfgpu_handle *h, const double *x, const double *p, double *out);
/* explicit derivative of the elastic energy w.r.t. the per-element Lame
 * parameters, out[e] = dE/dlambda_e, out[E+e] = dE/dmu_e (the energy density
 * is linear in (lambda, mu): NeoHookean 1/2 ln^2 J and
 * 1/2 (|F|^2 - 3 - 2 ln J), NeoHookeanElasticity.cpp:384; the term a
 * material objective that contains the elastic energy adds to the adjoint
 * terms) */
pfgpu_status pfgpu_energy_lame_gradient(pfgpu_handle *h, const double *x, double *out);
/* mixed derivative y = (d grad / d lambda) dlambda + (d grad / d mu) dmu,
 * the forward counterpart of pfgpu_material_vjp (out = -(that)^T p) */
pfgpu_status pfgpu_material_jvp(pfgpu_handle *h, const double *x, const double *dlambda,
                                const double *dmu, double *y);

/* ---- assembly, device pointers (asynchronous on the handle's stream) ---- */

pfgpu_status pfgpu_value_dev(pfgpu_handle *h, const double *x, double *energy);
/* The energy of pfgpu_value_dev without any allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  Its
 * scratch is a fixed device array of the library (65,536 + 128 doubles per
 * GPU, held until the process ends): meshes of up to 16,777,216 elements;
 * larger ones return PFGPU_ERR_UNSUPPORTED.  Every call on one GPU shares
 * that array, so calls that overlap on one GPU (other streams, other
 * handles) race.  The handle's device must be current.  Results equal
 * pfgpu_value_dev's bit for bit. */
pfgpu_status pfgpu_value_noalloc(pfgpu_handle *h, const double *x, double *energy);
pfgpu_status pfgpu_gradient_dev(pfgpu_handle *h, const double *x, double *g);
pfgpu_status pfgpu_hessian_values_dev(pfgpu_handle *h, const double *x, void *values);
pfgpu_status pfgpu_energy_gradient_hessian_dev(pfgpu_handle *h, const double *x, double *energy,
                                               double *g, void *values);
pfgpu_status pfgpu_hvp_dev(pfgpu_handle *h, const double *x, const double *v, double *y);
pfgpu_status pfgpu_material_vjp_dev(pfgpu_handle *h, const double *x, const double *p, double *out);
pfgpu_status pfgpu_shape_vjp_dev(pfgpu_handle *h, const double *x, const double *p, double *out);
/* device copies of the pattern held by the handle (valid until destroy) */
pfgpu_status pfgpu_hessian_pattern_dev(const pfgpu_handle *h, const int32_t **outer, const int32_t **inner);
pfgpu_status pfgpu_synchronize(pfgpu_handle *h);

/* ---- solves on the free dofs (callers of the assembly) ------------------ */

/* Newton-CG minimisation of the elastic energy over the free dofs, the
 * reference's nonlinear solve of the elastic form: Dirichlet dofs removed
 * from the unknowns (NLProblem::full_to_reduced / reduced_to_full,
 * NLProblem.cpp:583-629; BCLagrangianForm::project_gradient / hessian,
 * BCLagrangianForm.cpp:149-200), Newton direction H_ff d = -g_f from a
 * matrix-free CG on the Hessian-vector product (polysolve Newton with a CG
 * linear solver), backtracking Armijo line search on the energy, with
 * non-finite energies (inverted elements) rejected as invalid steps
 * (ElasticForm::is_step_valid, ElasticForm.cpp:388-405).  Every vector, dot
 * product and energy difference is FP64 with a fixed reduction order. */
typedef struct {
    int32_t max_iterations;    /* Newton iterations */
    double grad_norm;          /* stop when ||g_free||_2 <= grad_norm */
    int32_t cg_max_iterations; /* per Newton step */
    double cg_tolerance;       /* stop CG when ||r|| <= cg_tolerance ||g_free|| */
    int32_t line_search;       /* 0: full steps (alpha = 1), 1: backtracking Armijo */
    double armijo_c;           /* sufficient-decrease constant */
    int32_t max_line_search;   /* step halvings before giving up */
    int32_t use_graph;         /* 1: CG iterations replayed from a captured hipGraph */
    /* also converged once ||g_f|| <= relative_grad_norm ||g_f(x0)|| (per
     * problem in a batch); 0 disables */
    double relative_grad_norm;
    /* regularized Newton (polysolve's RegularizedNewton, with the
     * block-diagonal scaling of Levenberg-Marquardt): when the CG meets
     * non-positive curvature, the direction is solved again on
     * H + delta D, D the node 3x3 diagonal blocks of H (the block-Jacobi
     * preconditioner's), delta = regularization_min, then x10 per failed
     * solve up to regularization_max; after an accepted step delta /= 10
     * (0 below regularization_min).  regularization_min <= 0 disables it;
     * it needs the block-Jacobi preconditioner.  pfgpu_newton_solve only. */
    double regularization_min, regularization_max;
} pfgpu_newton_params;

typedef struct {
    int32_t iterations, cg_iterations, line_search_halvings, converged;
    double energy, grad_norm, last_alpha;
    /* steps that fell back to steepest descent (Cauchy step) after the
     * Newton direction was rejected by the line search */
    int32_t gradient_steps;
    /* how each Newton iteration's CG ended: relative residual at the
     * tolerance, non-positive curvature p^T H p <= 0 (the direction is the
     * iterate reached so far, the iteration where it happens), or the
     * iteration cap */
    int32_t cg_stops_tolerance, cg_stops_curvature, cg_stops_cap;
    double initial_grad_norm; /* reduced gradient norm at the initial guess */
    /* CG solves repeated on a regularized operator, and the delta of the
     * last accepted step's direction (0: plain Newton) */
    int32_t regularized_solves;
    double last_regularization;
} pfgpu_newton_report;
enum { PFGPU_CG_TOLERANCE = 0, PFGPU_CG_CURVATURE = 1, PFGPU_CG_ITERATION_CAP = 2 };

/* x: n_dofs values, the initial guess with the Dirichlet values already on
 * the fixed dofs, overwritten by the result; fixed: n_dofs bytes, 1 = fixed */
pfgpu_status pfgpu_newton_solve(pfgpu_handle *h, double *x, const uint8_t *fixed, const pfgpu_newton_params *params,
                                pfgpu_newton_report *report);
pfgpu_status pfgpu_newton_solve_dev(pfgpu_handle *h, double *x, const uint8_t *fixed,
                                    const pfgpu_newton_params *params, pfgpu_newton_report *report);
/* Batched Newton-CG over independent problems held by one handle as one
 * mesh with disjoint node ranges (same order and material; build the handle
 * from the concatenated meshes): problem s owns elements
 * [elem_offsets[s], elem_offsets[s+1]) and nodes [node_offsets[s],
 * node_offsets[s+1]).  Every assembly kernel is one launch for the whole
 * batch; one CG on the block-diagonal system gives all Newton directions;
 * line search, step length and convergence are per problem (reports has
 * n_problems entries); each problem's CG stops at its own tolerance and a
 * problem that has converged costs no further element work.  Host pointers; the elastic energy only (no load, no
 * inertia). */
pfgpu_status pfgpu_newton_solve_batch(pfgpu_handle *h, double *x, const uint8_t *fixed, int32_t n_problems,
                                      const int32_t *elem_offsets, const int32_t *node_offsets,
                                      const pfgpu_newton_params *params, pfgpu_newton_report *reports);
/* A batch of independent problems of uneven size and mixed order and
 * material (P1/P2 x NeoHookean/linear elasticity, default quadrature), the
 * reference's independent states solved side by side (the parallel loop of
 * AdjointNLProblem.cpp:439): one handle whose element kernels (energy,
 * gradient, Hv, the preconditioner's diagonal blocks) and node-row gathers run
 * as one launch each over all problems, whatever their order and material;
 * the problems are grouped by (order, material) internally, each group's
 * meshes concatenated.  meshes / materials: n_problems descriptors. */
pfgpu_status pfgpu_batch_create(int32_t n_problems, const pfgpu_mesh_desc *meshes,
                                const pfgpu_material_desc *materials, int32_t precision, pfgpu_handle **out);
/* Newton-CG on every problem of a pfgpu_batch_create handle at once, with
 * the per-problem line search, CG stopping and convergence of
 * pfgpu_newton_solve_batch.  x[i]: problem i's n_dofs initial guess (host;
 * Dirichlet values already on its fixed dofs), overwritten by the result;
 * fixed[i]: its mask (1 = fixed) or NULL; reports: n_problems entries, in
 * the caller's order. */
pfgpu_status pfgpu_batch_newton(pfgpu_handle *h, double *const *x, const uint8_t *const *fixed,
                                const pfgpu_newton_params *params, pfgpu_newton_report *reports);
/* CG solve of H(x)_ff y_f = b_f with y = 0 on fixed dofs, matrix-free; the
 * static adjoint solve (solve_static_adjoint, StateDiff.cpp:230-318; the
 * Hessian is symmetric, so H^T = H).  Stops at ||r|| <= tolerance ||b_f||. */
pfgpu_status pfgpu_cg_solve(pfgpu_handle *h, const double *x, const uint8_t *fixed, const double *b, double *y,
                            int32_t max_iterations, double tolerance, int32_t *iterations, double *rel_residual);
pfgpu_status pfgpu_cg_solve_dev(pfgpu_handle *h, const double *x, const uint8_t *fixed, const double *b, double *y,
                                int32_t max_iterations, double tolerance, int32_t use_graph, int32_t *iterations,
                                double *rel_residual);
/* Linear algebra of the CG solves (the reference picks polysolve's linear
 * solver and preconditioner from the "solver.linear" options passed where the
 * nonlinear solver is built, StateSolveNonlinear.cpp:47-50; here the CG is
 * the solver).  hessian: the objective Hessian as values assembled once per
 * Newton iteration and applied by a block-sparse product, or the matrix-free
 * element product; AUTO assembles for P1 and stays matrix-free for P2.
 * preconditioner: none, or the inverse 3x3 diagonal blocks (block Jacobi). */
enum {
    PFGPU_SOLVER_HESSIAN_AUTO = 0,
    PFGPU_SOLVER_HESSIAN_ASSEMBLED = 1,
    PFGPU_SOLVER_HESSIAN_MATRIX_FREE = 2,
    /* assembled from element Hessians projected onto the PSD cone (the
     * reference's project_to_psd solver option) */
    PFGPU_SOLVER_HESSIAN_ASSEMBLED_PSD = 3
};
enum { PFGPU_PRECOND_NONE = 0, PFGPU_PRECOND_BLOCK_JACOBI = 1 };
pfgpu_status pfgpu_set_solver_options(pfgpu_handle *h, int32_t hessian, int32_t preconditioner);

/* external load f (n_dofs, host pointer; NULL removes it): the Newton solve
 * then minimises E(x) - f^T x (a body force such as gravity, assembled by
 * the caller, e.g. f = M g with pfgpu_mass_matvec) */
pfgpu_status pfgpu_set_load(pfgpu_handle *h, const double *f);

/* ---- transient: mass, implicit Euler, checkpointed adjoint --------------- */

/* consistent mass matrix M_ab = rho int phi_a phi_b (block = M_ab I_3), the
 * reference's Mass assembler (assembler/Mass.cpp:4-20) integrated exactly;
 * rho: per-element density (n_elements) or NULL for rho_uniform */
pfgpu_status pfgpu_set_density(pfgpu_handle *h, const double *rho, double rho_uniform);
/* mass matrix kind:
 *   PFGPU_MASS_CONSISTENT (0, the default) the consistent matrix;
 *   PFGPU_MASS_ROW_SUM (1) row-sum lumping, the reference's
 *     solver.advanced.lump_mass_matrix (State.cpp:1542-1545,
 *     utils::lump_matrix, MatrixUtils.cpp:84-101): every row's sum on its
 *     diagonal block (at P2 the vertex rows sum to negative masses, as in the
 *     reference);
 *   PFGPU_MASS_HRZ (2) the Hinton-Rock-Zienkiewicz diagonal of HRZMass
 *     (assembler/Mass.cpp:48-86): per element, diagonal entry a is
 *     M_e(a,a) sum_bc M_e(b,c) / sum_b M_e(b,b), nothing off the diagonal
 *     (positive at P2; with density 1 it is the reference's pure mass).
 * Applies to the current and later pfgpu_set_density calls. */
enum { PFGPU_MASS_CONSISTENT = 0, PFGPU_MASS_ROW_SUM = 1, PFGPU_MASS_HRZ = 2 };
pfgpu_status pfgpu_set_mass_lumping(pfgpu_handle *h, int32_t lumped);
/* y = M v */
pfgpu_status pfgpu_mass_matvec(pfgpu_handle *h, const double *v, double *y);
pfgpu_status pfgpu_mass_matvec_dev(pfgpu_handle *h, const double *v, double *y);

typedef struct {
    double dt;                /* time step */
    int32_t time_steps;       /* T */
    int32_t checkpoint_every; /* states kept every this many steps (1: all) */
} pfgpu_transient_params;

/* Implicit Euler (ImplicitEuler.cpp:13-31; InertiaForm.cpp:17-35): step i
 * minimises 1/2 (x - xt)^T M (x - xt) + dt^2 (E(x) - f^T x) over the free
 * dofs with xt = x_{i-1} + dt v_{i-1}, by the Newton solve above; fixed dofs
 * keep their x0 values.  Host pointers; f may be NULL; xT, vT receive the
 * final state.  The handle keeps the checkpoints for pfgpu_transient_adjoint. */
pfgpu_status pfgpu_transient_forward(pfgpu_handle *h, const double *x0, const double *v0, const uint8_t *fixed,
                                     const double *f, const pfgpu_transient_params *params,
                                     const pfgpu_newton_params *newton, double *xT, double *vT);
/* Gradient of J(x_T, v_T) through the last forward run, from dJ/dx_T and
 * dJ/dv_T (either may be NULL): dtheta (2 n_elements, lambda block then mu
 * block), dx0 and dv0 (n_dofs, zero on fixed dofs); any output may be NULL.
 * The backward recurrence of solve_transient_adjoint (StateDiff.cpp:320-413)
 * for BDF-1, recomputing the states between checkpoints. */
pfgpu_status pfgpu_transient_adjoint(pfgpu_handle *h, const double *dJdx, const double *dJdv, int32_t cg_max_iterations,
                                     double cg_tolerance, double *dtheta, double *dx0, double *dv0);

/* work counts of the last pfgpu_transient_forward and pfgpu_transient_adjoint
 * (the adjoint's recomputed forward steps are counted in recompute_*), the
 * per-step solver statistics the reference logs */
typedef struct pfgpu_transient_stats {
    int64_t steps;
    int64_t newton_iterations;
    int64_t cg_iterations;
    int64_t line_search_halvings;
    int64_t recompute_newton_iterations;
    int64_t recompute_cg_iterations;
    int64_t adjoint_cg_iterations;
} pfgpu_transient_stats;
pfgpu_status pfgpu_get_transient_stats(pfgpu_handle *h, pfgpu_transient_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
