// The handle behind the C-ABI and the launch functions that operate on it.
#pragma once

#include "common.hpp"

struct pfgpu_handle {
    int order = 1, n_local = 4, n_quad = 1, material = PFGPU_NEOHOOKEAN;
    int precision = PFGPU_FP64, path = PFGPU_PATH_FUSED;
    int n_elements = 0, n_nodes = 0, n_vertices = 0, n_dofs = 0;
    hipStream_t stream = nullptr;
    int device = 0;

    pfgpu::Pattern pat;
    std::vector<int32_t> h_elem_nodes; // host copy, element-major (E x n_local)

    // element data (SoA), value type = precision
    pfgpu::DevBuf conn, vconn, geom, lame;
    // node / vertex incidence (gathers of element-local vectors)
    pfgpu::DevBuf node_ptr, node_inc, vert_ptr, vert_inc;
    // plain-path Hessian maps
    bool plain_hess_ready = false;
    pfgpu::DevBuf brow_ptr, block_row, contrib_ptr, contrib;
    // fused-path patch tables (built on first use)
    bool fused_ready = false;
    pfgpu::DevBuf patch_desc, patch_rows, patch_row_acc, patch_elems, patch_cs, patch_cg, patch_hdr, patch_pos;
    int n_patches = 0;
    int fused_max_acc = 0, fused_max_rows = 0;
    int64_t fused_visits = 0, fused_chunks = 0;
    std::vector<double> h_vertices;       // host copy, n_vertices x 3
    std::vector<int32_t> h_elem_vertices; // host copy, E x 4
    // lazily built device pattern
    pfgpu::DevBuf d_outer, d_inner;

    // scratch
    pfgpu::DevBuf loc, hloc, partial, red_out;
    pfgpu::DevBuf h_x, h_v, h_out, h_out2, h_vals; // staging for host-pointer calls

    size_t value_size() const { return precision == PFGPU_MIXED ? sizeof(float) : sizeof(double); }
    int n_blocks_sym() const { return n_local * (n_local + 1) / 2; }
    int64_t device_bytes() const;
    template <typename T> pfgpu::DevMesh<T> mesh() const
    {
        return pfgpu::DevMesh<T>{n_elements, n_nodes, n_vertices, conn.as<int32_t>(), vconn.as<int32_t>(),
                                 geom.as<T>(), lame.as<T>()};
    }
};

namespace pfgpu {

using Handle = pfgpu_handle;

// Deterministic FP64 sum of n doubles into out[0] (fixed two-level tree).
void sum_doubles(Handle &h, const double *in, int n, double *out);

// ---- plain path (plain.hip) -------------------------------------------
void plain_energy(Handle &h, const double *x, double *energy, double *per_element);
void plain_gradient(Handle &h, const double *x, double *g, double *energy);
void plain_hessian(Handle &h, const double *x, void *values);
void plain_hvp(Handle &h, const double *x, const double *v, double *y);
void plain_material_jvp(Handle &h, const double *x, const double *dl, const double *dm, double *y);
void element_material_vjp(Handle &h, const double *x, const double *p, double *out);
void plain_shape_vjp(Handle &h, const double *x, const double *p, double *out);
void ensure_plain_hessian_maps(Handle &h);

// ---- fused path (fused.hip) -------------------------------------------
void ensure_fused_tables(Handle &h);
void fused_gradient(Handle &h, const double *x, double *g, double *energy);
void fused_energy_gradient_hessian(Handle &h, const double *x, double *energy, double *g, void *values);
void fused_hvp(Handle &h, const double *x, const double *v, double *y);

// ---- misc kernels (reduce.hip) ----------------------------------------
void build_device_pattern(Handle &h);
// returns true when every value in the device array is finite
bool all_finite(Handle &h, const void *dev, int64_t n, bool is_float);

} // namespace pfgpu
