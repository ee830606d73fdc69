// The handle behind the C-ABI and the launch functions that operate on it.
#pragma once

#include <array>
#include <memory>

#include "common.hpp"

struct pfgpu_handle {
    int order = 1, n_local = 4, n_quad = 1, material = PFGPU_NEOHOOKEAN;
    int precision = PFGPU_FP64, path = PFGPU_PATH_PLAIN;
    int n_elements = 0, n_nodes = 0, n_vertices = 0, n_dofs = 0;
    hipStream_t stream = nullptr;
    int device = 0;

    pfgpu::Pattern pat;
    std::vector<int32_t> h_elem_nodes; // host copy, element-major (E x n_local)

    // element data (SoA), value type = precision
    pfgpu::DevBuf conn, vconn, geom, lame;
    // node / vertex incidence (gathers of element-local vectors)
    pfgpu::DevBuf node_ptr, node_inc, vert_ptr, vert_inc;
    // plain-path Hessian maps: the contributions of every pattern block
    // (host copies, for the gather plans), every block's first CSR value,
    // and one gather plan per Hessian route (built for one workgroup size)
    bool plain_hess_ready = false;
    pfgpu::DevBuf brow_ptr, blk_out;
    std::vector<int32_t> h_contrib_ptr, h_contrib_code;
    pfgpu::GatherPlan blk_plan, fac_plan;
    pfgpu::UnitPlan unit_plan; // the factored route's row-unit gather
    // plan of the segmented node-row gather (chunks, padded slots, row metadata)
    pfgpu::DevBuf row_chunk, row_slot, row_meta;
    int n_row_chunks = 0;
    // factored Hessian (PFGPU_PATH_FACTORED): per-element factor records
    pfgpu::DevBuf fac_rec;
    int n_cu = 0; // compute units of the device (queried on first use)
    // launch shapes of the plain path (set by pfgpu_autotune / pfgpu_set_tuning)
    int tune_gather_threads = 256, tune_hessian_threads = 256;
    // fused-path patch tables (built on first use)
    bool fused_ready = false;
    pfgpu::DevBuf patch_desc, patch_rows, patch_row_acc, patch_elems, patch_cs, patch_cg, patch_hdr, patch_pos, patch_vnodes;
    int n_patches = 0;
    int fused_max_acc = 0, fused_max_rows = 0;
    int64_t fused_visits = 0, fused_chunks = 0;
    std::vector<double> h_vertices;       // host copy, n_vertices x 3
    std::vector<int32_t> h_elem_vertices; // host copy, E x 4
    // lazily built device pattern
    pfgpu::DevBuf d_outer, d_inner;
    pfgpu::DevBuf d_vertices; // vertex positions (FP64) of the device geometry

    // scratch
    pfgpu::DevBuf loc, hloc, partial, red_out, sum_part;
    pfgpu::DevBuf h_x, h_v, h_out, h_out2, h_vals; // staging for host-pointer calls
    // consistent mass matrix: one value per 3x3 block (block = m I_3)
    pfgpu::DevBuf mass, pat_bcol;
    bool has_mass = false;
    int lump_mass = 0; // 0 consistent, 1 row sums, 2 HRZ (PFGPU_MASS_*)
    std::vector<double> h_rho; // per-element density of the mass matrix (empty: rho_uniform)
    double rho_uniform = 0;
    // objective of the solves: w E(x) + 1/2 (x - xt)^T M (x - xt) - w f^T x
    // (xt, f: device vectors or null; the mass term needs xt)
    double obj_w = 1.0;
    const double *obj_xt = nullptr, *obj_f = nullptr;
    int obj_generation = 0;
    // transient run: checkpoints of (x, v) every ckpt_every steps
    std::vector<pfgpu::DevBuf> ckpt;
    pfgpu::DevBuf tr_fixed, tr_f, tr_seg, load;
    bool has_load = false;
    double tr_dt = 0;
    int tr_steps = 0, tr_every = 0;
    bool tr_has_f = false, tr_has_fixed = false;
    pfgpu_newton_params tr_newton{};
    pfgpu_transient_stats tr_stats{};
    bool tr_recompute = false; // forward steps run by the adjoint count as recomputation
    // solver workspace: vectors, scalars, the captured CG iteration
    pfgpu::DevBuf sv, h_mask;
    void *cg_graph = nullptr; // hipGraphExec_t
    std::array<const void *, 9> cg_graph_key{};
    // CG operator: assembled block-sparse values or matrix-free, block-Jacobi preconditioner
    int solver_hessian = PFGPU_SOLVER_HESSIAN_AUTO, solver_precond = PFGPU_PRECOND_BLOCK_JACOBI;
    bool op_valid = false, op_assembled = false, op_precond = false;
    bool op_records = false; // fac_rec holds the factor records at the operator's x
    // regularization weight delta of the operator (w H + M + delta D, D the
    // preconditioner's diagonal blocks, kept in reg_diag as 9 floats per node)
    double op_reg = 0.0;
    pfgpu::DevBuf reg_diag;
    // how prepare_operator built the preconditioner (node-diagonal blocks,
    // mass folded into the values, the fixed-dof mask)
    bool op_ndiag = false, op_fold = false;
    const uint8_t *op_fixed = nullptr;
    pfgpu::SegSkip skip; // set only inside a batched CG
    pfgpu::DevBuf solver_vals, pinv, diag_pos, node_diag;
    hipStream_t own_stream = nullptr; // used when the caller's stream is the legacy default
    // mixed batch (pfgpu_batch_create): one sub-handle per (order, material)
    // group; this handle's node space concatenates theirs, its element kernels
    // run over all groups in one launch (batch.hip).  Empty otherwise.
    std::vector<std::unique_ptr<pfgpu_handle>> groups;
    std::vector<int32_t> batch_perm;     // caller's problem index of each problem, in group order
    std::vector<int32_t> batch_eo, batch_no; // element / node offsets of the problems, group order
    ~pfgpu_handle();

    size_t value_size() const { return precision == PFGPU_MIXED ? sizeof(float) : sizeof(double); }
    int n_blocks_sym() const { return n_local * (n_local + 1) / 2; }
    int64_t device_bytes() const;
    template <typename T> pfgpu::DevMesh<T> mesh() const
    {
        return pfgpu::DevMesh<T>{n_elements, n_nodes, n_vertices, conn.as<int32_t>(), vconn.as<int32_t>(),
                                 geom.as<T>(), lame.as<T>(), material == PFGPU_MOONEY_RIVLIN ? 3 : 2};
    }
};

namespace pfgpu {

using Handle = pfgpu_handle;

// Deterministic FP64 sum of n doubles into out[0] (fixed two-level tree).
void sum_doubles(Handle &h, const double *in, int n, double *out);
// the same with the caller's first-level scratch (SUM_BLOCKS doubles, used
// when n > 32 BLOCK); allocates nothing
constexpr int SUM_BLOCKS = 128;
void sum_doubles_into(Handle &h, const double *in, int n, double *out, double *part);

// ---- plain path (plain.hip) -------------------------------------------
void plain_energy(Handle &h, const double *x, double *energy, double *per_element);
// the energy with the caller's scratch: partial (ceil(E / BLOCK) doubles) and
// part (SUM_BLOCKS doubles); allocates nothing
void energy_into(Handle &h, const double *x, double *energy, double *partial, double *part);
void plain_gradient(Handle &h, const double *x, double *g, double *energy);
void plain_hessian(Handle &h, const double *x, void *values, bool project_to_psd = false);
// every element's local Hessian in the block tiles projected onto the PSD cone (psd.hip)
void psd_project_blocks(Handle &h, void *hloc);
void plain_hvp(Handle &h, const double *x, const double *v, double *y);
void plain_material_jvp(Handle &h, const double *x, const double *dl, const double *dm, double *y);
void element_material_vjp(Handle &h, const double *x, const double *p, double *out);
void energy_lame_gradient(Handle &h, const double *x, double *out);
void plain_shape_vjp(Handle &h, const double *x, const double *p, double *out);
void ensure_plain_hessian_maps(Handle &h);
void factored_energy_gradient_hessian(Handle &h, const double *x, double *energy, double *g, void *values);
void factored_node_diagonal(Handle &h, const double *x, void *diag);
// H v from the factor records in fac_rec (valid only while they belong to x)
void factored_hvp(Handle &h, const double *v, double *y);

// ---- mixed batches (batch.hip) -----------------------------------------
void mixed_energy(Handle &h, const double *x, double *energy, double *per_element);
void mixed_gradient(Handle &h, const double *x, double *g, double *energy);
void mixed_hvp(Handle &h, const double *x, const double *v, double *y);
void mixed_node_diagonal(Handle &h, const double *x, void *diag);
void ensure_row_chunks(Handle &h);

// ---- fused path (fused.hip) -------------------------------------------
void ensure_fused_tables(Handle &h);
void fused_gradient(Handle &h, const double *x, double *g, double *energy);
void fused_energy_gradient_hessian(Handle &h, const double *x, double *energy, double *g, void *values);
void fused_hvp(Handle &h, const double *x, const double *v, double *y);

// ---- mass (mass.cpp) ---------------------------------------------------
std::vector<double> reference_mass(int order);
void build_mass(Handle &h, const double *rho, double rho_uniform);

// ---- geometry (geometry.hip) -------------------------------------------
bool device_geometry(Handle &h, const double *vertices);

// ---- solver (solver.hip) ------------------------------------------------
void mass_apply(Handle &h, const double *v, double *y);
void newton_solve(Handle &h, double *x, const uint8_t *fixed, const pfgpu_newton_params &p, pfgpu_newton_report &r);
void newton_solve_batch(Handle &h, double *x, const uint8_t *fixed, int nseg, const int32_t *eo, const int32_t *no,
                        const pfgpu_newton_params &p, pfgpu_newton_report *reps);
void transient_forward(Handle &h, const double *x0, const double *v0, const uint8_t *fixed, const double *f,
                       double dt, int steps, int every, const pfgpu_newton_params &np, double *xT, double *vT);
void transient_adjoint(Handle &h, const double *dJdx, const double *dJdv, int cg_max, double cg_tol, double *dtheta,
                       double *dx0, double *dv0);
// exit_reason (optional): PFGPU_CG_TOLERANCE, PFGPU_CG_CURVATURE or PFGPU_CG_ITERATION_CAP
int cg_solve_ws(Handle &h, const double *x, const uint8_t *fixed, const double *b, int max_it, double tol, bool use_graph,
                double *rel, int *exit_reason = nullptr);
double *cg_solution(Handle &h);
void solver_release(Handle &h);

// ---- misc kernels (reduce.hip) ----------------------------------------
void build_device_pattern(Handle &h);
// returns true when every value in the device array is finite
bool all_finite(Handle &h, const void *dev, int64_t n, bool is_float);

} // namespace pfgpu
