// ge_sph.hip -- SPH uniform grid and PCISPH passes (placeholder until the
// SPH kernels land).
#include "ge_internal.hpp"

namespace ge {
int sph_new_grid(ge_ctx *c, const ge_step_params *) { c->err = "SPH not built yet"; return GE_E_UNSUPPORTED; }
int sph_pcisph(ge_ctx *c, const ge_step_params *, float) { c->err = "SPH not built yet"; return GE_E_UNSUPPORTED; }
int sph_neighbors(ge_ctx *c, int, int64_t *, uint32_t *, int64_t, int64_t *) { c->err = "SPH not built yet"; return GE_E_UNSUPPORTED; }
}  // namespace ge
