// The C-ABI of libpfgpu (include/pfgpu.h).  Exceptions never cross this
// boundary: every entry point converts them to a status and a message.
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>

#include "dispatch.hpp"
#include "handle.hpp"

using namespace pfgpu;

namespace {

thread_local std::string g_last_error;

template <typename F> pfgpu_status guard(F &&f)
{
    try {
        f();
        return PFGPU_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.status;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return PFGPU_ERR_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return PFGPU_ERR_HIP;
    }
}

void require(bool ok, const std::string &msg)
{
    if (!ok) fail(PFGPU_ERR_INVALID_ARGUMENT, msg);
}

void check_handle(const pfgpu_handle *h) { require(h != nullptr, "null handle"); }

// a host table element-major (E x k) as component-major (k x E)
template <typename S, typename D> std::vector<D> soa(const S *src, int E, int k)
{
    std::vector<D> v(size_t(E) * k);
    for (int e = 0; e < E; ++e)
        for (int c = 0; c < k; ++c) v[size_t(c) * E + e] = D(src[size_t(e) * k + c]);
    return v;
}
template <typename S, typename D> DevBuf upload_soa(const S *src, int E, int k, hipStream_t s)
{
    return upload(soa<S, D>(src, E, k), s);
}
// element geometry (J^{-T}, det J) of the handle's vertices, in place
void upload_geometry(pfgpu_handle &h, const pfgpu_mesh_desc &m);

// ElementAssemblyValues::finalize3d for the P1 geometry (ElementAssemblyValues.cpp:73-103):
// tmp.row(r) = sum_j dphi_j/dxi_r X_j, i.e. rows X1-X0, X2-X0, X3-X0; stores
// J^{-T} row-major and det J.
void geometry_host(const pfgpu_mesh_desc &m, std::vector<double> &g /* E x 10 */)
{
    g.resize(size_t(m.n_elements) * 10);
    for (int e = 0; e < m.n_elements; ++e) {
        const int32_t *v = m.elem_vertices + size_t(e) * 4;
        for (int j = 0; j < 4; ++j)
            require(v[j] >= 0 && v[j] < m.n_vertices, "elem_vertices references a vertex out of range");
        const double *X0 = m.vertices + 3 * size_t(v[0]);
        double t[9];
        for (int r = 0; r < 3; ++r) {
            const double *X = m.vertices + 3 * size_t(v[r + 1]);
            for (int c = 0; c < 3; ++c) t[r * 3 + c] = X[c] - X0[c];
        }
        const double c00 = t[4] * t[8] - t[5] * t[7], c01 = t[5] * t[6] - t[3] * t[8], c02 = t[3] * t[7] - t[4] * t[6];
        const double c10 = t[2] * t[7] - t[1] * t[8], c11 = t[0] * t[8] - t[2] * t[6], c12 = t[1] * t[6] - t[0] * t[7];
        const double c20 = t[1] * t[5] - t[2] * t[4], c21 = t[2] * t[3] - t[0] * t[5], c22 = t[0] * t[4] - t[1] * t[3];
        const double det = t[0] * c00 + t[1] * c01 + t[2] * c02;
        require(std::isfinite(det) && det != 0.0, "element " + std::to_string(e) + " has a degenerate geometry");
        const double id = 1.0 / det;
        // inverse = adj / det with adj[r][c] = cof[c][r]; J^{-T}[r][c] = inverse[c][r] = cof[r][c] / det
        const double jit[9] = {c00 * id, c01 * id, c02 * id, c10 * id, c11 * id, c12 * id, c20 * id, c21 * id, c22 * id};
        double *out = g.data() + size_t(e) * 10;
        for (int k = 0; k < 9; ++k) out[k] = jit[k];
        out[9] = det;
    }
}

// per-element material parameters, k rows (lambda, mu; MooneyRivlin c1, c2, k)
void upload_params(pfgpu_handle &h, const double *const *rows, int k)
{
    std::vector<double> lm(size_t(h.n_elements) * k);
    for (int e = 0; e < h.n_elements; ++e)
        for (int r = 0; r < k; ++r) {
            require(std::isfinite(rows[r][e]), "non-finite material parameter");
            lm[size_t(e) * k + r] = rows[r][e];
        }
    // in place: graphs captured over the handle's kernels keep reading the current values
    if (h.precision == PFGPU_MIXED)
        upload_into(h.lame, soa<double, float>(lm.data(), h.n_elements, k), h.stream);
    else
        upload_into(h.lame, soa<double, double>(lm.data(), h.n_elements, k), h.stream);
}

void upload_lame(pfgpu_handle &h, const double *lambda, const double *mu)
{
    const double *rows[2] = {lambda, mu};
    upload_params(h, rows, 2);
}

void upload_geometry(pfgpu_handle &h, const pfgpu_mesh_desc &m)
{
    // computed on the device; a degenerate element is located by the host
    // computation, which raises with its index
    for (int e = 0; e < m.n_elements; ++e)
        for (int j = 0; j < 4; ++j) {
            const int32_t v = m.elem_vertices[size_t(e) * 4 + j];
            require(v >= 0 && v < m.n_vertices, "elem_vertices references a vertex out of range");
        }
    if (device_geometry(h, m.vertices)) return;
    std::vector<double> g;
    geometry_host(m, g);
    fail(PFGPU_ERR_INVALID_ARGUMENT, "an element has a non-finite geometry");
}

// staging of host arrays for the synchronous entry points
const double *stage(pfgpu_handle &h, DevBuf &buf, const double *host, size_t n)
{
    buf.reserve(sizeof(double) * n);
    PF_HIP(hipMemcpyAsync(buf.get(), host, sizeof(double) * n, hipMemcpyHostToDevice, h.stream));
    return buf.as<double>();
}

void fetch(pfgpu_handle &h, void *host, const void *dev, size_t bytes)
{
    PF_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, h.stream));
    PF_HIP(hipStreamSynchronize(h.stream));
}

void gradient_dev(pfgpu_handle &h, const double *x, double *g, double *energy)
{
    if (h.path == PFGPU_PATH_FUSED)
        fused_gradient(h, x, g, energy);
    else
        plain_gradient(h, x, g, energy);
}

void egh_dev(pfgpu_handle &h, const double *x, double *energy, double *g, void *values)
{
    if (h.path == PFGPU_PATH_FUSED) {
        fused_energy_gradient_hessian(h, x, energy, g, values);
        return;
    }
    if (h.path == PFGPU_PATH_FACTORED && g && values) {
        factored_energy_gradient_hessian(h, x, energy, g, values);
        return;
    }
    if (g)
        plain_gradient(h, x, g, energy);
    else if (energy)
        plain_energy(h, x, energy, nullptr);
    if (values) plain_hessian(h, x, values);
}

void hvp_dev(pfgpu_handle &h, const double *x, const double *v, double *y)
{
    if (h.path == PFGPU_PATH_FUSED)
        fused_hvp(h, x, v, y);
    else
        plain_hvp(h, x, v, y);
}

void check_finite(pfgpu_handle &h, const void *dev, int64_t n, bool is_float, const char *what)
{
    if (!all_finite(h, dev, n, is_float))
        fail(PFGPU_ERR_INVALID_STATE, std::string(what) + " has non-finite entries (inverted element?)");
}

} // namespace

pfgpu_handle::~pfgpu_handle()
{
    pfgpu::solver_release(*this);
    if (own_stream) (void)hipStreamDestroy(own_stream);
}

int64_t pfgpu_handle::device_bytes() const
{
    const DevBuf *bufs[] = {&conn, &vconn, &geom, &lame, &node_ptr, &node_inc, &vert_ptr, &vert_inc,
                            &brow_ptr, &blk_out, &blk_plan.chunk, &blk_plan.slot, &blk_plan.meta, &fac_plan.chunk,
                            &fac_plan.slot, &fac_plan.meta, &fac_plan.dst, &unit_plan.chunk_row, &unit_plan.udst, &unit_plan.bq0, &unit_plan.lanes, &row_chunk, &row_slot, &row_meta, &node_diag, &reg_diag, &patch_desc, &patch_rows, &fac_rec,
                            &patch_row_acc, &patch_elems, &patch_cs, &patch_cg, &patch_hdr, &patch_pos, &patch_vnodes,
                            &d_outer, &d_inner, &loc, &hloc, &partial, &red_out, &sum_part,
                            &h_x, &h_v, &h_out, &h_out2, &h_vals, &sv, &h_mask, &mass, &pat_bcol, &tr_fixed, &tr_f, &tr_seg, &load};
    int64_t s = 0;
    for (const DevBuf *b : bufs) s += int64_t(b->bytes());
    for (const DevBuf &b : ckpt) s += int64_t(b.bytes());
    for (const auto &g : groups) s += g->device_bytes();
    return s;
}

extern "C" {

const char *pfgpu_last_error(void) { return g_last_error.c_str(); }

} // extern "C"

namespace {

// the quadrature rules of 1 and 4 points weigh every point alike; the fused
// and factored paths take no other
bool uniform_rule(const pfgpu_handle &h) { return h.n_quad == 1 || h.n_quad == 4; }

// the fused and factored paths: NeoHookean and linear elasticity, rules of
// 1 or 4 points
void require_path(const pfgpu_handle &h, int32_t path)
{
    if (path == PFGPU_PATH_PLAIN) return;
    if (h.material == PFGPU_SAINT_VENANT || h.material == PFGPU_FIXED_COROTATIONAL || h.material == PFGPU_MOONEY_RIVLIN)
        fail(PFGPU_ERR_UNSUPPORTED, "the fused and factored paths cover NeoHookean and linear elasticity");
    if (!uniform_rule(h))
        fail(PFGPU_ERR_UNSUPPORTED, "the fused and factored paths take quadrature orders 1 and 2 (the plain path takes 3 and 4)");
}

// the handle of one mesh and material (pfgpu_create)
std::unique_ptr<pfgpu_handle> create_handle(const pfgpu_mesh_desc *mesh, const pfgpu_material_desc *material,
                                            int32_t precision)
{
    require(mesh != nullptr && material != nullptr, "null mesh or material descriptor");
    const pfgpu_mesh_desc &m = *mesh;
    require(m.order == 1 || m.order == 2, "order must be 1 or 2");
    require(m.quadrature_order <= 4, "quadrature orders above 4 are not provided");
    require(m.n_elements > 0 && m.n_nodes > 0 && m.n_vertices > 0, "empty mesh");
    require(m.elem_nodes && m.vertices && m.elem_vertices, "null mesh array");
    require(material->kind == PFGPU_NEOHOOKEAN || material->kind == PFGPU_LINEAR_ELASTICITY ||
                material->kind == PFGPU_SAINT_VENANT || material->kind == PFGPU_FIXED_COROTATIONAL ||
                material->kind == PFGPU_MOONEY_RIVLIN,
            "unknown material");
    require(precision == PFGPU_FP64 || precision == PFGPU_MIXED, "unknown precision");
    require((material->lambda == nullptr) == (material->mu == nullptr), "lambda and mu must both be given or both be NULL");

    auto h = std::make_unique<pfgpu_handle>();
    PF_HIP(hipGetDevice(&h->device));
    h->order = m.order;
    h->n_local = m.order == 1 ? 4 : 10;
    // default max(2 (p - 1), 1) (AssemblerUtils.cpp:225-228); orders 1 and 2: 1 and 4 points, 3 and 4: the
    // 6- and 11-point Xiao-Gimbutas rules (TetQuadrature.cpp:14-36)
    const int qo = m.quadrature_order <= 0 ? (m.order == 1 ? 1 : 2) : m.quadrature_order;
    h->n_quad = qo == 1 ? 1 : qo == 2 ? 4 : qo == 3 ? 6 : 11;
    h->material = material->kind;
    h->precision = precision;
    h->path = (material->kind == PFGPU_NEOHOOKEAN || material->kind == PFGPU_LINEAR_ELASTICITY) && uniform_rule(*h)
                  ? PFGPU_PATH_FACTORED
                  : PFGPU_PATH_PLAIN;
    h->n_elements = m.n_elements;
    h->n_nodes = m.n_nodes;
    h->n_vertices = m.n_vertices;
    h->n_dofs = 3 * m.n_nodes;

    h->h_elem_nodes.assign(m.elem_nodes, m.elem_nodes + size_t(m.n_elements) * h->n_local);
    h->h_vertices.assign(m.vertices, m.vertices + size_t(m.n_vertices) * 3);
    h->h_elem_vertices.assign(m.elem_vertices, m.elem_vertices + size_t(m.n_elements) * 4);
    h->pat = build_pattern(m.n_elements, h->n_local, m.elem_nodes, m.n_nodes, m.n_vertices, m.elem_vertices);

    h->vconn = upload_soa<int32_t, int32_t>(m.elem_vertices, m.n_elements, 4, h->stream);
    upload_geometry(*h, m);

    if (material->kind == PFGPU_MOONEY_RIVLIN) {
        const bool given = material->c1 != nullptr;
        require(given == (material->c2 != nullptr) && given == (material->k != nullptr),
                "c1, c2 and k must all be given or all be NULL");
        std::vector<double> u1, u2, u3;
        const double *rows[3] = {material->c1, material->c2, material->k};
        if (!given) {
            u1.assign(size_t(m.n_elements), material->c1_uniform);
            u2.assign(size_t(m.n_elements), material->c2_uniform);
            u3.assign(size_t(m.n_elements), material->k_uniform);
            rows[0] = u1.data(), rows[1] = u2.data(), rows[2] = u3.data();
        }
        upload_params(*h, rows, 3);
    } else if (material->lambda) {
        upload_lame(*h, material->lambda, material->mu);
    } else {
        const double E = material->young, nu = material->poisson;
        require(std::isfinite(E) && E > 0 && nu > -1.0 && nu < 0.5, "invalid Young's modulus / Poisson ratio");
        // LameParameters (MatParams.cpp:11-22), 3D
        const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = E / (2.0 * (1.0 + nu));
        std::vector<double> la(m.n_elements, lam), mv(m.n_elements, mu);
        upload_lame(*h, la.data(), mv.data());
    }
    h->conn = upload_soa<int32_t, int32_t>(m.elem_nodes, m.n_elements, h->n_local, h->stream);
    h->node_ptr = upload(h->pat.node_ptr, h->stream);
    h->node_inc = upload(h->pat.node_inc, h->stream);
    h->vert_ptr = upload(h->pat.vert_ptr, h->stream);
    h->vert_inc = upload(h->pat.vert_inc, h->stream);
    // fixed-size scratch of the reductions, and the element-level scratch
    // that several entry points share, sized here for the largest use
    // (energy partials per 64 elements, 3 n_local local values per element,
    // the factor records) so that a graph captured over one *_dev call stays
    // valid after any other call (a path's gather plans are built by its
    // first call)
    h->sum_part.reserve(sizeof(double) * 128);
    h->red_out.reserve(sizeof(int) * 4);
    h->partial.reserve(sizeof(double) * size_t((int64_t(m.n_elements) + 63) / 64 + 1));
    h->loc.reserve(h->value_size() * size_t(m.n_elements) * size_t(h->n_local) * 3);
    if ((h->material == PFGPU_NEOHOOKEAN || h->material == PFGPU_LINEAR_ELASTICITY) && uniform_rule(*h))
        h->fac_rec.reserve(h->value_size() * size_t(m.n_elements) * size_t((11 * h->n_quad + 6 + 3) & ~3));
    return h;
}

} // namespace

extern "C" {

pfgpu_status pfgpu_create(const pfgpu_mesh_desc *mesh, const pfgpu_material_desc *material, int32_t precision,
                          pfgpu_handle **out)
{
    return guard([&] {
        require(out != nullptr, "null output handle pointer");
        *out = nullptr;
        *out = create_handle(mesh, material, precision).release();
    });
}

pfgpu_status pfgpu_destroy(pfgpu_handle *h)
{
    return guard([&] {
        if (!h) return;
        (void)hipStreamSynchronize(h->stream);
        delete h;
    });
}

pfgpu_status pfgpu_info(const pfgpu_handle *h, pfgpu_info_t *info)
{
    return guard([&] {
        check_handle(h);
        require(info != nullptr, "null info");
        info->order = h->order;
        info->n_local = h->n_local;
        info->n_quad = h->n_quad;
        info->material = h->material;
        info->precision = h->precision;
        info->path = h->path;
        info->n_elements = h->n_elements;
        info->n_nodes = h->n_nodes;
        info->n_vertices = h->n_vertices;
        info->n_dofs = h->n_dofs;
        info->nnz = h->pat.nnz;
        info->device_bytes = h->device_bytes();
    });
}

pfgpu_status pfgpu_set_stream(pfgpu_handle *h, void *stream)
{
    return guard([&] {
        check_handle(h);
        h->stream = static_cast<hipStream_t>(stream);
    });
}

pfgpu_status pfgpu_set_tuning(pfgpu_handle *h, int32_t path, int32_t gather_threads, int32_t hessian_threads)
{
    return guard([&] {
        check_handle(h);
        require(path == PFGPU_PATH_FUSED || path == PFGPU_PATH_PLAIN || path == PFGPU_PATH_FACTORED, "unknown path");
        require(gather_threads == 64 || gather_threads == 128 || gather_threads == 256 || gather_threads == 512,
                "gather_threads must be 64, 128, 256 or 512");
        require(hessian_threads == 64 || hessian_threads == 128 || hessian_threads == 256,
                "hessian_threads must be 64, 128 or 256");
        require_path(*h, path);
        h->path = path;
        h->tune_gather_threads = gather_threads;
        h->tune_hessian_threads = hessian_threads;
    });
}

pfgpu_status pfgpu_autotune(pfgpu_handle *h, const double *x, pfgpu_tune_report *report)
{
    return guard([&] {
        check_handle(h);
        require(x && report, "null argument");
        const size_t nnz = size_t(h->pat.nnz);
        DevBuf out(sizeof(double) * (size_t(h->n_dofs) + 1) + h->value_size() * nnz);
        double *de = out.as<double>(), *dg = de + 1;
        void *dv = dg + h->n_dofs;
        hipEvent_t t0, t1;
        PF_HIP(hipEventCreate(&t0));
        PF_HIP(hipEventCreate(&t1));
        auto time_config = [&]() {
            egh_dev(*h, x, de, dg, dv); // warm-up; builds any table of the configuration
            float best = 1e30f;
            for (int r = 0; r < 3; ++r) {
                PF_HIP(hipEventRecord(t0, h->stream));
                egh_dev(*h, x, de, dg, dv);
                PF_HIP(hipEventRecord(t1, h->stream));
                PF_HIP(hipEventSynchronize(t1));
                float ms = 0;
                PF_HIP(hipEventElapsedTime(&ms, t0, t1));
                best = std::min(best, ms);
            }
            return double(best);
        };
        pfgpu_tune_report r{};
        r.ms = 1e300;
        r.ms_fused = -1;
        r.ms_factored = -1;
        const int G[3] = {128, 256, 512}, HB[3] = {64, 128, 256};
        try {
            for (int gi = 0; gi < 3; ++gi)
                for (int hi = 0; hi < 3; ++hi) {
                    h->path = PFGPU_PATH_PLAIN;
                    h->tune_gather_threads = G[gi];
                    h->tune_hessian_threads = HB[hi];
                    const double ms = time_config();
                    r.candidates++;
                    if (G[gi] == 256 && HB[hi] == 256) r.ms_plain_default = ms;
                    if (ms < r.ms) {
                        r.ms = ms;
                        r.path = PFGPU_PATH_PLAIN;
                        r.gather_threads = G[gi];
                        r.hessian_threads = HB[hi];
                    }
                }
            // the factored path keeps the plain winner's Hessian-kernel size
            // (used by its PSD-projected assemblies)
            h->tune_hessian_threads = r.hessian_threads;
            const int GF[4] = {64, 128, 256, 512};
            if (h->material == PFGPU_NEOHOOKEAN || h->material == PFGPU_LINEAR_ELASTICITY)
                for (int gi = 0; gi < 4; ++gi) {
                    h->path = PFGPU_PATH_FACTORED;
                    h->tune_gather_threads = GF[gi];
                    try {
                        const double ms = time_config();
                        r.candidates++;
                        if (r.ms_factored < 0 || ms < r.ms_factored) r.ms_factored = ms;
                        if (ms < r.ms) {
                            r.ms = ms;
                            r.path = PFGPU_PATH_FACTORED;
                            r.gather_threads = GF[gi];
                        }
                    } catch (const Error &e) {
                        if (e.status != PFGPU_ERR_UNSUPPORTED) throw;
                    }
                }
            h->path = PFGPU_PATH_FUSED;
            try {
                r.ms_fused = time_config();
                r.candidates++;
                if (r.ms_fused < r.ms) {
                    r.ms = r.ms_fused;
                    r.path = PFGPU_PATH_FUSED;
                }
            } catch (const Error &e) {
                if (e.status != PFGPU_ERR_UNSUPPORTED) throw;
            }
        } catch (...) {
            (void)hipEventDestroy(t0);
            (void)hipEventDestroy(t1);
            throw;
        }
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        h->path = r.path;
        h->tune_gather_threads = r.gather_threads ? r.gather_threads : 256;
        h->tune_hessian_threads = r.hessian_threads ? r.hessian_threads : 256;
        // one call in the chosen configuration builds its tables (gather
        // plans are built for one workgroup size), so that the caller's
        // first assembly does not
        egh_dev(*h, x, de, dg, dv);
        PF_HIP(hipStreamSynchronize(h->stream));
        *report = r;
    });
}

pfgpu_status pfgpu_set_path(pfgpu_handle *h, int32_t path)
{
    return guard([&] {
        check_handle(h);
        require(path == PFGPU_PATH_FUSED || path == PFGPU_PATH_PLAIN || path == PFGPU_PATH_FACTORED, "unknown path");
        require_path(*h, path);
        h->path = path;
    });
}

pfgpu_status pfgpu_set_mooney_rivlin(pfgpu_handle *h, const double *c1, const double *c2, const double *k)
{
    return guard([&] {
        check_handle(h);
        require(c1 && c2 && k, "null c1/c2/k");
        require(h->material == PFGPU_MOONEY_RIVLIN, "the handle's material is not MooneyRivlin");
        PF_HIP(hipStreamSynchronize(h->stream));
        const double *rows[3] = {c1, c2, k};
        upload_params(*h, rows, 3);
        solver_release(*h);
    });
}

pfgpu_status pfgpu_set_lame(pfgpu_handle *h, const double *lambda, const double *mu)
{
    return guard([&] {
        check_handle(h);
        require(lambda && mu, "null lambda/mu");
        require(h->material != PFGPU_MOONEY_RIVLIN, "MooneyRivlin takes c1, c2, k (pfgpu_set_mooney_rivlin)");
        PF_HIP(hipStreamSynchronize(h->stream));
        upload_lame(*h, lambda, mu);
        solver_release(*h);
    });
}

pfgpu_status pfgpu_set_vertices(pfgpu_handle *h, const double *vertices)
{
    return guard([&] {
        check_handle(h);
        require(vertices, "null vertices");
        const size_t n = size_t(h->n_vertices) * 3;
        for (size_t i = 0; i < n; ++i) require(std::isfinite(vertices[i]), "non-finite vertex coordinate");
        PF_HIP(hipStreamSynchronize(h->stream));
        pfgpu_mesh_desc m{h->order, 0, h->n_elements, h->n_nodes, h->h_elem_nodes.data(), h->n_vertices, vertices,
                          h->h_elem_vertices.data()};
        try {
            upload_geometry(*h, m);
        } catch (...) {
            (void)device_geometry(*h, h->h_vertices.data()); // the previous geometry stays
            throw;
        }
        h->h_vertices.assign(vertices, vertices + n);
        if (h->has_mass) build_mass(*h, h->h_rho.empty() ? nullptr : h->h_rho.data(), h->rho_uniform);
        solver_release(*h);
        PF_HIP(hipStreamSynchronize(h->stream));
    });
}

pfgpu_status pfgpu_value(pfgpu_handle *h, const double *x, double *energy)
{
    return guard([&] {
        check_handle(h);
        require(x && energy, "null argument");
        const double *dx = stage(*h, h->h_x, x, h->n_dofs);
        h->h_out.reserve(sizeof(double));
        plain_energy(*h, dx, h->h_out.as<double>(), nullptr);
        fetch(*h, energy, h->h_out.get(), sizeof(double));
        if (!std::isfinite(*energy)) fail(PFGPU_ERR_INVALID_STATE, "energy is not finite (inverted element?)");
    });
}

pfgpu_status pfgpu_value_per_element(pfgpu_handle *h, const double *x, double *out)
{
    return guard([&] {
        check_handle(h);
        require(x && out, "null argument");
        const double *dx = stage(*h, h->h_x, x, h->n_dofs);
        h->h_out.reserve(sizeof(double) * size_t(h->n_elements));
        plain_energy(*h, dx, nullptr, h->h_out.as<double>());
        fetch(*h, out, h->h_out.get(), sizeof(double) * size_t(h->n_elements));
        for (int e = 0; e < h->n_elements; ++e)
            if (!std::isfinite(out[e])) fail(PFGPU_ERR_INVALID_STATE, "element energy is not finite (inverted element?)");
    });
}

pfgpu_status pfgpu_gradient(pfgpu_handle *h, const double *x, double *g)
{
    return guard([&] {
        check_handle(h);
        require(x && g, "null argument");
        const double *dx = stage(*h, h->h_x, x, h->n_dofs);
        h->h_out.reserve(sizeof(double) * size_t(h->n_dofs));
        gradient_dev(*h, dx, h->h_out.as<double>(), nullptr);
        fetch(*h, g, h->h_out.get(), sizeof(double) * size_t(h->n_dofs));
        for (int i = 0; i < h->n_dofs; ++i)
            if (!std::isfinite(g[i])) fail(PFGPU_ERR_INVALID_STATE, "gradient is not finite (inverted element?)");
    });
}

pfgpu_status pfgpu_hessian_nnz(const pfgpu_handle *h, int64_t *nnz)
{
    return guard([&] {
        check_handle(h);
        require(nnz != nullptr, "null argument");
        *nnz = h->pat.nnz;
    });
}

pfgpu_status pfgpu_hessian_pattern(const pfgpu_handle *h, int32_t *outer, int32_t *inner)
{
    return guard([&] {
        check_handle(h);
        require(outer && inner, "null argument");
        const Pattern &p = h->pat;
        std::memcpy(outer, p.outer.data(), sizeof(int32_t) * p.outer.size());
        for (int i = 0; i < p.n_nodes; ++i) {
            const int r0 = p.brow_ptr[i], deg = p.brow_ptr[i + 1] - r0;
            for (int m = 0; m < 3; ++m)
                for (int k = 0; k < deg; ++k)
                    for (int n = 0; n < 3; ++n) inner[9 * int64_t(r0) + int64_t(m) * 3 * deg + 3 * k + n] = 3 * p.bcol[r0 + k] + n;
        }
    });
}

pfgpu_status pfgpu_hessian_values(pfgpu_handle *h, const double *x, int32_t project_to_psd, void *values)
{
    return guard([&] {
        check_handle(h);
        require(x && values, "null argument");
        const double *dx = stage(*h, h->h_x, x, h->n_dofs);
        const size_t bytes = h->value_size() * size_t(h->pat.nnz);
        h->h_vals.reserve(bytes);
        if (project_to_psd)
            plain_hessian(*h, dx, h->h_vals.get(), true); // the projection runs on the element tiles of the plain path
        else
            egh_dev(*h, dx, nullptr, nullptr, h->h_vals.get());
        check_finite(*h, h->h_vals.get(), h->pat.nnz, h->precision == PFGPU_MIXED, "Hessian");
        fetch(*h, values, h->h_vals.get(), bytes);
    });
}

pfgpu_status pfgpu_energy_gradient_hessian(pfgpu_handle *h, const double *x, double *energy, double *g, void *values)
{
    return guard([&] {
        check_handle(h);
        require(x != nullptr, "null x");
        const double *dx = stage(*h, h->h_x, x, h->n_dofs);
        double *de = nullptr, *dg = nullptr;
        void *dv = nullptr;
        if (energy) {
            h->h_out2.reserve(sizeof(double));
            de = h->h_out2.as<double>();
        }
        if (g) {
            h->h_out.reserve(sizeof(double) * size_t(h->n_dofs));
            dg = h->h_out.as<double>();
        }
        const size_t bytes = h->value_size() * size_t(h->pat.nnz);
        if (values) {
            h->h_vals.reserve(bytes);
            dv = h->h_vals.get();
        }
        egh_dev(*h, dx, de, dg, dv);
        if (energy) fetch(*h, energy, de, sizeof(double));
        if (g) fetch(*h, g, dg, sizeof(double) * size_t(h->n_dofs));
        if (values) fetch(*h, values, dv, bytes);
        if (energy && !std::isfinite(*energy)) fail(PFGPU_ERR_INVALID_STATE, "energy is not finite (inverted element?)");
        if (g)
            for (int i = 0; i < h->n_dofs; ++i)
                if (!std::isfinite(g[i])) fail(PFGPU_ERR_INVALID_STATE, "gradient is not finite (inverted element?)");
        if (values) check_finite(*h, dv, h->pat.nnz, h->precision == PFGPU_MIXED, "Hessian");
    });
}

pfgpu_status pfgpu_hvp(pfgpu_handle *h, const double *x, const double *v, double *y)
{
    return guard([&] {
        check_handle(h);
        require(x && v && y, "null argument");
        const double *dx = stage(*h, h->h_x, x, h->n_dofs);
        const double *dv = stage(*h, h->h_v, v, h->n_dofs);
        h->h_out.reserve(sizeof(double) * size_t(h->n_dofs));
        hvp_dev(*h, dx, dv, h->h_out.as<double>());
        fetch(*h, y, h->h_out.get(), sizeof(double) * size_t(h->n_dofs));
        for (int i = 0; i < h->n_dofs; ++i)
            if (!std::isfinite(y[i])) fail(PFGPU_ERR_INVALID_STATE, "Hessian-vector product is not finite");
    });
}

pfgpu_status pfgpu_material_vjp(pfgpu_handle *h, const double *x, const double *p, double *out)
{
    return guard([&] {
        check_handle(h);
        require(x && p && out, "null argument");
        const double *dx = stage(*h, h->h_x, x, h->n_dofs);
        const double *dp = stage(*h, h->h_v, p, h->n_dofs);
        h->h_out.reserve(sizeof(double) * 2 * size_t(h->n_elements));
        element_material_vjp(*h, dx, dp, h->h_out.as<double>());
        fetch(*h, out, h->h_out.get(), sizeof(double) * 2 * size_t(h->n_elements));
    });
}

pfgpu_status pfgpu_energy_lame_gradient(pfgpu_handle *h, const double *x, double *out)
{
    return guard([&] {
        check_handle(h);
        require(x && out, "null argument");
        const double *dx = stage(*h, h->h_x, x, h->n_dofs);
        h->h_out.reserve(sizeof(double) * 2 * size_t(h->n_elements));
        energy_lame_gradient(*h, dx, h->h_out.as<double>());
        fetch(*h, out, h->h_out.get(), sizeof(double) * 2 * size_t(h->n_elements));
    });
}

pfgpu_status pfgpu_shape_vjp(pfgpu_handle *h, const double *x, const double *p, double *out)
{
    return guard([&] {
        check_handle(h);
        require(x && p && out, "null argument");
        const double *dx = stage(*h, h->h_x, x, h->n_dofs);
        const double *dp = stage(*h, h->h_v, p, h->n_dofs);
        h->h_out.reserve(sizeof(double) * 3 * size_t(h->n_vertices));
        plain_shape_vjp(*h, dx, dp, h->h_out.as<double>());
        fetch(*h, out, h->h_out.get(), sizeof(double) * 3 * size_t(h->n_vertices));
    });
}

pfgpu_status pfgpu_material_jvp(pfgpu_handle *h, const double *x, const double *dlambda, const double *dmu, double *y)
{
    return guard([&] {
        check_handle(h);
        require(x && dlambda && dmu && y, "null argument");
        const double *dx = stage(*h, h->h_x, x, h->n_dofs);
        h->h_v.reserve(sizeof(double) * 2 * size_t(h->n_elements));
        double *dl = h->h_v.as<double>();
        PF_HIP(hipMemcpyAsync(dl, dlambda, sizeof(double) * size_t(h->n_elements), hipMemcpyHostToDevice, h->stream));
        PF_HIP(hipMemcpyAsync(dl + h->n_elements, dmu, sizeof(double) * size_t(h->n_elements), hipMemcpyHostToDevice, h->stream));
        h->h_out.reserve(sizeof(double) * size_t(h->n_dofs));
        plain_material_jvp(*h, dx, dl, dl + h->n_elements, h->h_out.as<double>());
        fetch(*h, y, h->h_out.get(), sizeof(double) * size_t(h->n_dofs));
    });
}

// ---- device-pointer entry points ----------------------------------------

pfgpu_status pfgpu_value_dev(pfgpu_handle *h, const double *x, double *energy)
{
    return guard([&] {
        check_handle(h);
        require(x && energy, "null argument");
        plain_energy(*h, x, energy, nullptr);
    });
}

} // extern "C"

namespace {

// Scratch of pfgpu_value_noalloc: the energy partials of up to
// NOALLOC_MAX_PARTIALS workgroups of BLOCK elements, then SUM_BLOCKS
// second-level sums.  Each GPU the library runs on holds its own copy until
// the process ends; host code uses only the per-device addresses from
// hipGetSymbolAddress, never the array's name.
constexpr int NOALLOC_MAX_PARTIALS = 65536;
constexpr int64_t NOALLOC_MAX_ELEMENTS = int64_t(NOALLOC_MAX_PARTIALS) * BLOCK;
static __device__ __align__(256) double scratch[NOALLOC_MAX_PARTIALS + SUM_BLOCKS];

std::mutex g_scratch_mutex;
std::vector<double *> g_scratch_ptr; // per device, set on the first call on that device

// the device address of `scratch` on device dev (the current device)
double *scratch_address(int dev)
{
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (size_t(dev) >= g_scratch_ptr.size()) g_scratch_ptr.resize(size_t(dev) + 1, nullptr);
    if (!g_scratch_ptr[size_t(dev)]) {
        void *p = nullptr;
        PF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[size_t(dev)] = static_cast<double *>(p);
    }
    return g_scratch_ptr[size_t(dev)];
}

} // namespace

extern "C" {

pfgpu_status pfgpu_value_noalloc(pfgpu_handle *h, const double *x, double *energy)
{
    return guard([&] {
        check_handle(h);
        require(x && energy, "null argument");
        if (int64_t(h->n_elements) > NOALLOC_MAX_ELEMENTS)
            fail(PFGPU_ERR_UNSUPPORTED, "pfgpu_value_noalloc holds at most " + std::to_string(NOALLOC_MAX_ELEMENTS) +
                                            " elements (its scratch is a fixed device array); use pfgpu_value_dev");
        int cur = 0;
        PF_HIP(hipGetDevice(&cur));
        require(cur == h->device, "the handle's device must be the current device");
        double *s = scratch_address(cur);
        energy_into(*h, x, energy, s, s + NOALLOC_MAX_PARTIALS);
    });
}

pfgpu_status pfgpu_gradient_dev(pfgpu_handle *h, const double *x, double *g)
{
    return guard([&] {
        check_handle(h);
        require(x && g, "null argument");
        gradient_dev(*h, x, g, nullptr);
    });
}

pfgpu_status pfgpu_hessian_values_dev(pfgpu_handle *h, const double *x, void *values)
{
    return guard([&] {
        check_handle(h);
        require(x && values, "null argument");
        egh_dev(*h, x, nullptr, nullptr, values);
    });
}

pfgpu_status pfgpu_energy_gradient_hessian_dev(pfgpu_handle *h, const double *x, double *energy, double *g, void *values)
{
    return guard([&] {
        check_handle(h);
        require(x != nullptr, "null x");
        egh_dev(*h, x, energy, g, values);
    });
}

pfgpu_status pfgpu_hvp_dev(pfgpu_handle *h, const double *x, const double *v, double *y)
{
    return guard([&] {
        check_handle(h);
        require(x && v && y, "null argument");
        hvp_dev(*h, x, v, y);
    });
}

pfgpu_status pfgpu_material_vjp_dev(pfgpu_handle *h, const double *x, const double *p, double *out)
{
    return guard([&] {
        check_handle(h);
        require(x && p && out, "null argument");
        element_material_vjp(*h, x, p, out);
    });
}

pfgpu_status pfgpu_shape_vjp_dev(pfgpu_handle *h, const double *x, const double *p, double *out)
{
    return guard([&] {
        check_handle(h);
        require(x && p && out, "null argument");
        plain_shape_vjp(*h, x, p, out);
    });
}

pfgpu_status pfgpu_hessian_pattern_dev(const pfgpu_handle *h, const int32_t **outer, const int32_t **inner)
{
    return guard([&] {
        check_handle(h);
        require(outer && inner, "null argument");
        pfgpu_handle *mh = const_cast<pfgpu_handle *>(h);
        build_device_pattern(*mh);
        *outer = mh->d_outer.as<int32_t>();
        *inner = mh->d_inner.as<int32_t>();
    });
}

pfgpu_status pfgpu_synchronize(pfgpu_handle *h)
{
    return guard([&] {
        check_handle(h);
        PF_HIP(hipStreamSynchronize(h->stream));
    });
}

} // extern "C"

namespace {

// Graph capture needs a stream other than the legacy default one: a solve on
// a handle bound to the default stream runs on the handle's own (blocking)
// stream, which orders itself after the default stream's earlier work.
struct SolverStream {
    pfgpu_handle &h;
    hipStream_t saved;
    explicit SolverStream(pfgpu_handle &hh) : h(hh), saved(hh.stream)
    {
        if (h.stream == nullptr) {
            if (!h.own_stream) PF_HIP(hipStreamCreate(&h.own_stream));
            h.stream = h.own_stream;
        }
    }
    ~SolverStream()
    {
        if (saved == nullptr) (void)hipStreamSynchronize(h.stream);
        h.stream = saved;
    }
};

void check_newton_params(const pfgpu_newton_params *p)
{
    require(p != nullptr, "null Newton parameters");
    require(p->max_iterations >= 0 && p->cg_max_iterations >= 0 && p->max_line_search >= 0, "negative iteration count");
    require(p->grad_norm >= 0 && p->cg_tolerance >= 0 && p->armijo_c >= 0 && p->armijo_c < 1 &&
                p->relative_grad_norm >= 0,
            "invalid tolerance");
    require(p->line_search == 0 || p->line_search == 1, "line_search must be 0 or 1");
    require(!(p->regularization_min > 0) || p->regularization_max >= p->regularization_min,
            "regularization_max must be at least regularization_min");
}

const uint8_t *stage_mask(pfgpu_handle &h, const uint8_t *fixed)
{
    if (!fixed) return nullptr;
    h.h_mask.reserve(size_t(h.n_dofs));
    PF_HIP(hipMemcpyAsync(h.h_mask.get(), fixed, size_t(h.n_dofs), hipMemcpyHostToDevice, h.stream));
    return h.h_mask.as<uint8_t>();
}

} // namespace

extern "C" {

pfgpu_status pfgpu_newton_solve_dev(pfgpu_handle *h, double *x, const uint8_t *fixed, const pfgpu_newton_params *params,
                                    pfgpu_newton_report *report)
{
    return guard([&] {
        check_handle(h);
        require(x != nullptr && report != nullptr, "null argument");
        check_newton_params(params);
        SolverStream ss(*h);
        newton_solve(*h, x, fixed, *params, *report);
    });
}

pfgpu_status pfgpu_newton_solve(pfgpu_handle *h, double *x, const uint8_t *fixed, const pfgpu_newton_params *params,
                                pfgpu_newton_report *report)
{
    return guard([&] {
        check_handle(h);
        require(x != nullptr && report != nullptr, "null argument");
        check_newton_params(params);
        SolverStream ss(*h);
        h->h_out2.reserve(sizeof(double) * size_t(h->n_dofs));
        double *dx = h->h_out2.as<double>();
        PF_HIP(hipMemcpyAsync(dx, x, sizeof(double) * size_t(h->n_dofs), hipMemcpyHostToDevice, h->stream));
        const uint8_t *dm = stage_mask(*h, fixed);
        newton_solve(*h, dx, dm, *params, *report);
        fetch(*h, x, dx, sizeof(double) * size_t(h->n_dofs));
    });
}

pfgpu_status pfgpu_newton_solve_batch(pfgpu_handle *h, double *x, const uint8_t *fixed, int32_t n_problems,
                                      const int32_t *elem_offsets, const int32_t *node_offsets,
                                      const pfgpu_newton_params *params, pfgpu_newton_report *reports)
{
    return guard([&] {
        check_handle(h);
        require(x && elem_offsets && node_offsets && reports, "null argument");
        check_newton_params(params);
        require(n_problems >= 1, "empty batch");
        require(elem_offsets[0] == 0 && elem_offsets[n_problems] == h->n_elements && node_offsets[0] == 0 &&
                    node_offsets[n_problems] == h->n_nodes,
                "batch offsets must cover the handle's elements and nodes");
        for (int s = 0; s < n_problems; ++s) {
            require(elem_offsets[s + 1] > elem_offsets[s] && node_offsets[s + 1] > node_offsets[s], "empty problem in the batch");
            // every element of problem s must use nodes of problem s only
            for (int e = elem_offsets[s]; e < elem_offsets[s + 1]; ++e)
                for (int a = 0; a < h->n_local; ++a) {
                    const int32_t nd = h->h_elem_nodes[size_t(e) * h->n_local + a];
                    require(nd >= node_offsets[s] && nd < node_offsets[s + 1], "a batch problem's element uses another problem's node");
                }
        }
        SolverStream ss(*h);
        h->h_out2.reserve(sizeof(double) * size_t(h->n_dofs));
        double *dx = h->h_out2.as<double>();
        PF_HIP(hipMemcpyAsync(dx, x, sizeof(double) * size_t(h->n_dofs), hipMemcpyHostToDevice, h->stream));
        const uint8_t *dm = stage_mask(*h, fixed);
        newton_solve_batch(*h, dx, dm, n_problems, elem_offsets, node_offsets, *params, reports);
        fetch(*h, x, dx, sizeof(double) * size_t(h->n_dofs));
    });
}

pfgpu_status pfgpu_batch_create(int32_t n_problems, const pfgpu_mesh_desc *meshes, const pfgpu_material_desc *materials,
                                int32_t precision, pfgpu_handle **out)
{
    return guard([&] {
        require(out != nullptr, "null output handle pointer");
        *out = nullptr;
        require(n_problems >= 1 && meshes && materials, "empty batch or null descriptors");
        require(precision == PFGPU_FP64 || precision == PFGPU_MIXED, "unknown precision");
        // group of each problem: (order, material) with the default quadrature
        std::vector<int> var(static_cast<size_t>(n_problems) + 0);
        for (int i = 0; i < n_problems; ++i) {
            const pfgpu_mesh_desc &m = meshes[i];
            const pfgpu_material_desc &mt = materials[i];
            require(m.order == 1 || m.order == 2, "order must be 1 or 2");
            require(m.quadrature_order <= 0 || m.quadrature_order == (m.order == 1 ? 1 : 2),
                    "batched problems use the default quadrature");
            require(mt.kind == PFGPU_NEOHOOKEAN || mt.kind == PFGPU_LINEAR_ELASTICITY,
                    "batched problems are NeoHookean or linear elasticity");
            require(m.n_elements > 0 && m.n_nodes > 0 && m.elem_nodes, "empty problem in the batch");
            var[size_t(i)] = (m.order == 2 ? 2 : 0) + (mt.kind == PFGPU_LINEAR_ELASTICITY ? 1 : 0);
        }
        auto H = std::make_unique<pfgpu_handle>();
        PF_HIP(hipGetDevice(&H->device));
        H->precision = precision;
        H->path = PFGPU_PATH_PLAIN;
        H->solver_hessian = PFGPU_SOLVER_HESSIAN_MATRIX_FREE;
        int64_t N = 0, E = 0, NV = 0;
        H->batch_eo.push_back(0);
        H->batch_no.push_back(0);
        for (int v = 0; v < 4; ++v) {
            std::vector<int> idx;
            for (int i = 0; i < n_problems; ++i)
                if (var[size_t(i)] == v) idx.push_back(i);
            if (idx.empty()) continue;
            const int order = v >= 2 ? 2 : 1, NB = order == 1 ? 4 : 10;
            std::vector<int32_t> en, ev;
            std::vector<double> vx, lam, mu;
            int64_t nn = 0, nv = 0, ne = 0;
            for (int i : idx) {
                const pfgpu_mesh_desc &m = meshes[i];
                const pfgpu_material_desc &mt = materials[i];
                require(m.vertices && m.elem_vertices && m.n_vertices > 0, "null mesh array");
                for (int64_t k = 0; k < int64_t(m.n_elements) * NB; ++k) {
                    const int32_t nd = m.elem_nodes[k];
                    require(nd >= 0 && nd < m.n_nodes, "elem_nodes references a node out of range");
                    en.push_back(int32_t(nd + nn));
                }
                for (int64_t k = 0; k < int64_t(m.n_elements) * 4; ++k) {
                    const int32_t vv = m.elem_vertices[k];
                    require(vv >= 0 && vv < m.n_vertices, "elem_vertices references a vertex out of range");
                    ev.push_back(int32_t(vv + nv));
                }
                vx.insert(vx.end(), m.vertices, m.vertices + size_t(m.n_vertices) * 3);
                require((mt.lambda == nullptr) == (mt.mu == nullptr), "lambda and mu must both be given or both be NULL");
                for (int e = 0; e < m.n_elements; ++e) {
                    if (mt.lambda) {
                        lam.push_back(mt.lambda[e]);
                        mu.push_back(mt.mu[e]);
                    } else {
                        const double Y = mt.young, nu = mt.poisson;
                        require(std::isfinite(Y) && Y > 0 && nu > -1.0 && nu < 0.5, "invalid Young's modulus / Poisson ratio");
                        lam.push_back(Y * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)));
                        mu.push_back(Y / (2.0 * (1.0 + nu)));
                    }
                }
                H->batch_perm.push_back(i);
                H->batch_eo.push_back(int32_t(E + ne + m.n_elements));
                H->batch_no.push_back(int32_t(N + nn + m.n_nodes));
                nn += m.n_nodes;
                nv += m.n_vertices;
                ne += m.n_elements;
            }
            require(N + nn < (int64_t(1) << 31) / 3 && E + ne < (int64_t(1) << 31), "batch too large for 32-bit indices");
            pfgpu_mesh_desc gm{order, 0, int32_t(ne), int32_t(nn), en.data(), int32_t(nv), vx.data(), ev.data()};
            pfgpu_material_desc gmat{};
            gmat.kind = v % 2 == 1 ? PFGPU_LINEAR_ELASTICITY : PFGPU_NEOHOOKEAN;
            gmat.lambda = lam.data();
            gmat.mu = mu.data();
            std::unique_ptr<pfgpu_handle> sub = create_handle(&gm, &gmat, precision);
            sub->path = PFGPU_PATH_PLAIN;
            // the group's block pattern, shifted into the batch's node space
            Pattern &P = H->pat;
            if (P.brow_ptr.empty()) P.brow_ptr.push_back(0);
            const int32_t b0 = P.brow_ptr.back();
            for (int i = 0; i < sub->n_nodes; ++i) P.brow_ptr.push_back(b0 + sub->pat.brow_ptr[size_t(i) + 1]);
            for (int32_t c : sub->pat.bcol) P.bcol.push_back(int32_t(c + N));
            P.n_blocks += sub->pat.n_blocks;
            P.nnz += sub->pat.nnz;
            N += nn;
            E += ne;
            NV += nv;
            H->groups.push_back(std::move(sub));
        }
        H->pat.n_nodes = int(N);
        H->n_nodes = int(N);
        H->n_dofs = int(3 * N);
        H->n_elements = int(E);
        H->n_vertices = int(NV);
        H->order = H->groups.front()->order;
        H->n_local = H->groups.front()->n_local;
        H->n_quad = H->groups.front()->n_quad;
        H->material = H->groups.front()->material;
        H->sum_part.reserve(sizeof(double) * 128);
        H->red_out.reserve(sizeof(int) * 4);
        *out = H.release();
    });
}

pfgpu_status pfgpu_batch_newton(pfgpu_handle *h, double *const *x, const uint8_t *const *fixed,
                                const pfgpu_newton_params *params, pfgpu_newton_report *reports)
{
    return guard([&] {
        check_handle(h);
        require(!h->groups.empty(), "not a batch handle (pfgpu_batch_create)");
        require(x && reports, "null argument");
        check_newton_params(params);
        const int n = int(h->batch_perm.size());
        // the problems' vectors in the batch's order (groups back to back)
        std::vector<double> xg(size_t(h->n_dofs));
        std::vector<uint8_t> mg(size_t(h->n_dofs), 0);
        for (int k = 0; k < n; ++k) {
            const int i = h->batch_perm[size_t(k)];
            const size_t o = 3 * size_t(h->batch_no[size_t(k)]), len = 3 * size_t(h->batch_no[size_t(k) + 1] - h->batch_no[size_t(k)]);
            require(x[i] != nullptr, "null problem vector");
            std::copy(x[i], x[i] + len, xg.begin() + long(o));
            if (fixed && fixed[i]) std::copy(fixed[i], fixed[i] + len, mg.begin() + long(o));
        }
        SolverStream ss(*h);
        h->h_out2.reserve(sizeof(double) * size_t(h->n_dofs));
        double *dx = h->h_out2.as<double>();
        PF_HIP(hipMemcpyAsync(dx, xg.data(), sizeof(double) * size_t(h->n_dofs), hipMemcpyHostToDevice, h->stream));
        const uint8_t *dm = stage_mask(*h, mg.data());
        std::vector<pfgpu_newton_report> reps(static_cast<size_t>(n) + 0);
        newton_solve_batch(*h, dx, dm, n, h->batch_eo.data(), h->batch_no.data(), *params, reps.data());
        fetch(*h, xg.data(), dx, sizeof(double) * size_t(h->n_dofs));
        for (int k = 0; k < n; ++k) {
            const int i = h->batch_perm[size_t(k)];
            const size_t o = 3 * size_t(h->batch_no[size_t(k)]), len = 3 * size_t(h->batch_no[size_t(k) + 1] - h->batch_no[size_t(k)]);
            std::copy(xg.begin() + long(o), xg.begin() + long(o + len), x[i]);
            reports[i] = reps[size_t(k)];
        }
    });
}

pfgpu_status pfgpu_cg_solve_dev(pfgpu_handle *h, const double *x, const uint8_t *fixed, const double *b, double *y,
                                int32_t max_iterations, double tolerance, int32_t use_graph, int32_t *iterations,
                                double *rel_residual)
{
    return guard([&] {
        check_handle(h);
        require(x && b && y, "null argument");
        require(max_iterations >= 0 && tolerance >= 0, "invalid CG limits");
        SolverStream ss(*h);
        double rel = 0;
        h->op_valid = false; // the operator is built at this x
        const int it = cg_solve_ws(*h, x, fixed, b, max_iterations, tolerance, use_graph != 0, &rel);
        PF_HIP(hipMemcpyAsync(y, cg_solution(*h), sizeof(double) * size_t(h->n_dofs), hipMemcpyDeviceToDevice, h->stream));
        if (iterations) *iterations = it;
        if (rel_residual) *rel_residual = rel;
    });
}

pfgpu_status pfgpu_set_solver_options(pfgpu_handle *h, int32_t hessian, int32_t preconditioner)
{
    return guard([&] {
        check_handle(h);
        require(hessian >= PFGPU_SOLVER_HESSIAN_AUTO && hessian <= PFGPU_SOLVER_HESSIAN_ASSEMBLED_PSD, "unknown hessian mode");
        require(preconditioner == PFGPU_PRECOND_NONE || preconditioner == PFGPU_PRECOND_BLOCK_JACOBI,
                "unknown preconditioner");
        h->solver_hessian = hessian;
        h->solver_precond = preconditioner;
        solver_release(*h);
    });
}

pfgpu_status pfgpu_get_transient_stats(pfgpu_handle *h, pfgpu_transient_stats *stats)
{
    return guard([&] {
        check_handle(h);
        require(stats, "null argument");
        *stats = h->tr_stats;
    });
}

pfgpu_status pfgpu_set_load(pfgpu_handle *h, const double *f)
{
    return guard([&] {
        check_handle(h);
        if (!f) {
            h->has_load = false;
            h->obj_f = nullptr;
            h->obj_generation++;
            return;
        }
        for (int i = 0; i < h->n_dofs; ++i) require(std::isfinite(f[i]), "non-finite load");
        h->load.reserve(sizeof(double) * size_t(h->n_dofs));
        PF_HIP(hipMemcpyAsync(h->load.get(), f, sizeof(double) * size_t(h->n_dofs), hipMemcpyHostToDevice, h->stream));
        PF_HIP(hipStreamSynchronize(h->stream));
        h->has_load = true;
        h->obj_f = h->load.as<double>();
        h->obj_generation++;
    });
}

pfgpu_status pfgpu_set_mass_lumping(pfgpu_handle *h, int32_t lumped)
{
    return guard([&] {
        check_handle(h);
        require(lumped == PFGPU_MASS_CONSISTENT || lumped == PFGPU_MASS_ROW_SUM || lumped == PFGPU_MASS_HRZ,
                "lumped must be 0 (consistent), 1 (row sums) or 2 (HRZ)");
        if (h->lump_mass == lumped) return;
        h->lump_mass = lumped;
        if (h->has_mass) {
            build_mass(*h, h->h_rho.empty() ? nullptr : h->h_rho.data(), h->rho_uniform);
            solver_release(*h);
            PF_HIP(hipStreamSynchronize(h->stream));
        }
    });
}

pfgpu_status pfgpu_set_density(pfgpu_handle *h, const double *rho, double rho_uniform)
{
    return guard([&] {
        check_handle(h);
        if (rho)
            for (int e = 0; e < h->n_elements; ++e) require(std::isfinite(rho[e]) && rho[e] > 0, "density must be positive");
        else
            require(std::isfinite(rho_uniform) && rho_uniform > 0, "density must be positive");
        build_mass(*h, rho, rho_uniform);
        if (rho)
            h->h_rho.assign(rho, rho + h->n_elements);
        else
            h->h_rho.clear();
        h->rho_uniform = rho_uniform;
        solver_release(*h);
        PF_HIP(hipStreamSynchronize(h->stream));
    });
}

pfgpu_status pfgpu_mass_matvec_dev(pfgpu_handle *h, const double *v, double *y)
{
    return guard([&] {
        check_handle(h);
        require(v && y, "null argument");
        mass_apply(*h, v, y);
    });
}

pfgpu_status pfgpu_mass_matvec(pfgpu_handle *h, const double *v, double *y)
{
    return guard([&] {
        check_handle(h);
        require(v && y, "null argument");
        const double *dv = stage(*h, h->h_v, v, h->n_dofs);
        h->h_out.reserve(sizeof(double) * size_t(h->n_dofs));
        mass_apply(*h, dv, h->h_out.as<double>());
        fetch(*h, y, h->h_out.get(), sizeof(double) * size_t(h->n_dofs));
    });
}

pfgpu_status pfgpu_transient_forward(pfgpu_handle *h, const double *x0, const double *v0, const uint8_t *fixed,
                                     const double *f, const pfgpu_transient_params *params,
                                     const pfgpu_newton_params *newton, double *xT, double *vT)
{
    return guard([&] {
        check_handle(h);
        require(x0 && v0 && params, "null argument");
        check_newton_params(newton);
        require(params->dt > 0 && std::isfinite(params->dt), "dt must be positive");
        require(params->time_steps >= 1 && params->checkpoint_every >= 1, "time_steps and checkpoint_every must be >= 1");
        SolverStream ss(*h);
        const size_t n = size_t(h->n_dofs), bytes = sizeof(double) * n;
        DevBuf buf(4 * bytes + n);
        double *dx0 = buf.as<double>(), *dv0 = dx0 + n, *dxT = dv0 + n, *dvT = dxT + n, *df = nullptr;
        uint8_t *dm = reinterpret_cast<uint8_t *>(dvT + n);
        PF_HIP(hipMemcpyAsync(dx0, x0, bytes, hipMemcpyHostToDevice, h->stream));
        PF_HIP(hipMemcpyAsync(dv0, v0, bytes, hipMemcpyHostToDevice, h->stream));
        if (fixed) PF_HIP(hipMemcpyAsync(dm, fixed, n, hipMemcpyHostToDevice, h->stream));
        DevBuf fb;
        if (f) {
            fb.alloc(bytes);
            df = fb.as<double>();
            PF_HIP(hipMemcpyAsync(df, f, bytes, hipMemcpyHostToDevice, h->stream));
        }
        transient_forward(*h, dx0, dv0, fixed ? dm : nullptr, df, params->dt, params->time_steps,
                          params->checkpoint_every, *newton, dxT, dvT);
        if (xT) PF_HIP(hipMemcpyAsync(xT, dxT, bytes, hipMemcpyDeviceToHost, h->stream));
        if (vT) PF_HIP(hipMemcpyAsync(vT, dvT, bytes, hipMemcpyDeviceToHost, h->stream));
        PF_HIP(hipStreamSynchronize(h->stream));
    });
}

pfgpu_status pfgpu_transient_adjoint(pfgpu_handle *h, const double *dJdx, const double *dJdv, int32_t cg_max_iterations,
                                     double cg_tolerance, double *dtheta, double *dx0, double *dv0)
{
    return guard([&] {
        check_handle(h);
        require(cg_max_iterations > 0 && cg_tolerance >= 0, "invalid CG limits");
        SolverStream ss(*h);
        const size_t n = size_t(h->n_dofs), bytes = sizeof(double) * n, E = size_t(h->n_elements);
        DevBuf buf(4 * bytes + 2 * E * sizeof(double));
        double *djx = buf.as<double>(), *djv = djx + n, *ox = djv + n, *ov = ox + n, *ot = ov + n;
        if (dJdx) PF_HIP(hipMemcpyAsync(djx, dJdx, bytes, hipMemcpyHostToDevice, h->stream));
        if (dJdv) PF_HIP(hipMemcpyAsync(djv, dJdv, bytes, hipMemcpyHostToDevice, h->stream));
        transient_adjoint(*h, dJdx ? djx : nullptr, dJdv ? djv : nullptr, cg_max_iterations, cg_tolerance, ot, ox, ov);
        if (dtheta) PF_HIP(hipMemcpyAsync(dtheta, ot, 2 * E * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        if (dx0) PF_HIP(hipMemcpyAsync(dx0, ox, bytes, hipMemcpyDeviceToHost, h->stream));
        if (dv0) PF_HIP(hipMemcpyAsync(dv0, ov, bytes, hipMemcpyDeviceToHost, h->stream));
        PF_HIP(hipStreamSynchronize(h->stream));
    });
}

pfgpu_status pfgpu_cg_solve(pfgpu_handle *h, const double *x, const uint8_t *fixed, const double *b, double *y,
                            int32_t max_iterations, double tolerance, int32_t *iterations, double *rel_residual)
{
    return guard([&] {
        check_handle(h);
        require(x && b && y, "null argument");
        require(max_iterations >= 0 && tolerance >= 0, "invalid CG limits");
        SolverStream ss(*h);
        const double *dx = stage(*h, h->h_x, x, h->n_dofs);
        const double *db = stage(*h, h->h_v, b, h->n_dofs);
        const uint8_t *dm = stage_mask(*h, fixed);
        double rel = 0;
        h->op_valid = false; // the operator is built at this x
        const int it = cg_solve_ws(*h, dx, dm, db, max_iterations, tolerance, true, &rel);
        fetch(*h, y, cg_solution(*h), sizeof(double) * size_t(h->n_dofs));
        if (iterations) *iterations = it;
        if (rel_residual) *rel_residual = rel;
    });
}

} // extern "C"
