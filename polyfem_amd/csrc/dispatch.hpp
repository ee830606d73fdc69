// Compile-time dispatch over (value type, element order, quadrature size, material).
#pragma once

#include "elem.hpp"
#include "handle.hpp"

namespace pfgpu {

constexpr int BLOCK = 256;

inline unsigned grid_for(int64_t n, int block = BLOCK) { return unsigned((n + block - 1) / block); }

// f.template operator()<T, ORDER, NQ, MAT>() for the handle's value type,
// order, quadrature (1, 4, 6 or 11 points) and material; UNIFORM: only the
// rules that weigh every point alike (1 and 4 points), the kernels of the
// factored and fused paths
template <bool UNIFORM = false, typename F> void dispatch(const Handle &h, F &&f)
{
    auto mat = [&]<typename T, int O, int Q>() {
        if (h.material == MAT_NH)
            f.template operator()<T, O, Q, MAT_NH>();
        else if (h.material == MAT_SV)
            f.template operator()<T, O, Q, MAT_SV>();
        else if (h.material == MAT_FC)
            f.template operator()<T, O, Q, MAT_FC>();
        else if (h.material == MAT_MR)
            f.template operator()<T, O, Q, MAT_MR>();
        else
            f.template operator()<T, O, Q, MAT_LE>();
    };
    auto quad = [&]<typename T, int O>() {
        if (h.n_quad == 1) {
            mat.template operator()<T, O, 1>();
        } else if (h.n_quad == 4) {
            mat.template operator()<T, O, 4>();
        } else if constexpr (UNIFORM) {
            fail(PFGPU_ERR_UNSUPPORTED, "this path takes quadrature orders 1 and 2 only (the plain path takes 3 and 4)");
        } else if (h.n_quad == 6) {
            mat.template operator()<T, O, 6>();
        } else {
            mat.template operator()<T, O, 11>();
        }
    };
    auto order = [&]<typename T>() {
        if (h.order == 1)
            quad.template operator()<T, 1>();
        else
            quad.template operator()<T, 2>();
    };
    if (h.precision == PFGPU_MIXED)
        order.template operator()<float>();
    else
        order.template operator()<double>();
}

} // namespace pfgpu
