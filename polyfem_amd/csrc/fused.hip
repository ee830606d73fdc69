// Fused assembly path (patch-owned rows, on-chip reduction).
#include "dispatch.hpp"

namespace pfgpu {

void ensure_fused_tables(Handle &) { fail(PFGPU_ERR_UNSUPPORTED, "fused path not built yet"); }
void fused_gradient(Handle &h, const double *, double *, double *) { ensure_fused_tables(h); }
void fused_energy_gradient_hessian(Handle &h, const double *, double *, double *, void *) { ensure_fused_tables(h); }
void fused_hvp(Handle &h, const double *, const double *, double *) { ensure_fused_tables(h); }

} // namespace pfgpu
