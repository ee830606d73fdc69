// Element geometry on the device: the P1 Jacobian of every tet from the
// vertex positions, as ElementAssemblyValues::finalize3d computes it
// (ElementAssemblyValues.cpp:73-103): the rows of t are X1-X0, X2-X0, X3-X0;
// the stored values are J^{-T} = cof(t) / det t (row-major) and det t, in
// FP64 and rounded once to the handle's value type.
#include "dispatch.hpp"

namespace pfgpu {

namespace {

template <typename T>
__global__ __launch_bounds__(BLOCK) void k_geometry(int n_elements, const double *__restrict__ X, const int32_t *__restrict__ vconn,
                                                    T *__restrict__ geom)
{
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= n_elements) return;
    const int64_t v0 = vconn[e];
    double t[9];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int64_t v = vconn[int64_t(r + 1) * n_elements + e];
#pragma unroll
        for (int c = 0; c < 3; ++c) t[r * 3 + c] = X[3 * v + c] - X[3 * v0 + c];
    }
    const double c00 = t[4] * t[8] - t[5] * t[7], c01 = t[5] * t[6] - t[3] * t[8], c02 = t[3] * t[7] - t[4] * t[6];
    const double c10 = t[2] * t[7] - t[1] * t[8], c11 = t[0] * t[8] - t[2] * t[6], c12 = t[1] * t[6] - t[0] * t[7];
    const double c20 = t[1] * t[5] - t[2] * t[4], c21 = t[2] * t[3] - t[0] * t[5], c22 = t[0] * t[4] - t[1] * t[3];
    const double det = t[0] * c00 + t[1] * c01 + t[2] * c02;
    const double id = 1.0 / det; // a degenerate tet gives non-finite entries, which the host rejects
    const double g[10] = {c00 * id, c01 * id, c02 * id, c10 * id, c11 * id, c12 * id, c20 * id, c21 * id, c22 * id, det};
#pragma unroll
    for (int k = 0; k < 10; ++k) geom[int64_t(k) * n_elements + e] = T(g[k]);
}

} // namespace

// vertices: host n_vertices x 3; the handle's vconn must be on the device.
// Returns false when some element is degenerate (non-finite geometry).
bool device_geometry(Handle &h, const double *vertices)
{
    const size_t nv = size_t(h.n_vertices) * 3;
    h.d_vertices.reserve(sizeof(double) * nv);
    PF_HIP(hipMemcpyAsync(h.d_vertices.get(), vertices, sizeof(double) * nv, hipMemcpyHostToDevice, h.stream));
    h.geom.reserve(h.value_size() * 10 * size_t(h.n_elements));
    const unsigned g = grid_for(h.n_elements);
    if (h.precision == PFGPU_MIXED)
        k_geometry<float><<<g, BLOCK, 0, h.stream>>>(h.n_elements, h.d_vertices.as<double>(), h.vconn.as<int32_t>(),
                                                     h.geom.as<float>());
    else
        k_geometry<double><<<g, BLOCK, 0, h.stream>>>(h.n_elements, h.d_vertices.as<double>(), h.vconn.as<int32_t>(),
                                                      h.geom.as<double>());
    PF_LAUNCH_CHECK();
    return all_finite(h, h.geom.get(), 10 * int64_t(h.n_elements), h.precision == PFGPU_MIXED);
}

} // namespace pfgpu
