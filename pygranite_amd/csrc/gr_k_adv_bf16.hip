// Single-job step kernels over BF16 slices.
#include "gr_advance.h"

namespace gr
{

AdvanceFn pick_advance_bf16(uint32_t dim, uint32_t integrator, uint32_t mode, int spec, bool wide, bool narrow,
                           bool quad)
{
    return pick_advance_t<bf16_t>(dim, integrator, mode, spec, wide, narrow, quad);
}

} // namespace gr
