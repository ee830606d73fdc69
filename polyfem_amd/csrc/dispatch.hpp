// Compile-time dispatch over (value type, element order, quadrature size, material).
#pragma once

#include "elem.hpp"
#include "handle.hpp"

namespace pfgpu {

constexpr int BLOCK = 256;

inline unsigned grid_for(int64_t n, int block = BLOCK) { return unsigned((n + block - 1) / block); }

// f.template operator()<T, ORDER, NQ, MAT>()
template <typename F> void dispatch(const Handle &h, F &&f)
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
        if (h.n_quad == 1)
            mat.template operator()<T, O, 1>();
        else
            mat.template operator()<T, O, 4>();
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
