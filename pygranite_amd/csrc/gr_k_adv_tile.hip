// Single-job 3D step kernels with a per-workgroup field tile in LDS (tile_step): every step the
// workgroup stages the box of time-paired nodes around its live particles' cells and samples the
// stencils inside it from LDS. Results are those of the untiled time-paired kernels.
#include "gr_advance.h"

namespace gr
{

template <int DIM, int INTEG, class T, int MODE, int SPEC>
__global__ void __launch_bounds__(256) k_advance_tile(const StepParams P, uint32_t nxcd)
{
    advance_block<DIM, INTEG, T, MODE, true, SPEC, 1, true>(P, nxcd);
}

template <int DIM, int INTEG, class T, int MODE, int SPEC> struct KTileFn
{
    static AdvanceFn fn() { return k_advance_tile<DIM, INTEG, T, MODE, SPEC>; }
};

template <class T> static AdvanceFn pick_tile_t(uint32_t integrator, int spec)
{
    return integrator == GR_RK4 ? SpecPick<KTileFn, 3, GR_RK4, T, M_PAIRED>::get(spec)
                                : SpecPick<KTileFn, 3, GR_EULER, T, M_PAIRED>::get(spec);
}

AdvanceFn pick_advance_tile(uint32_t dtype, uint32_t integrator, int spec)
{
    if (dtype == GR_F16) return pick_tile_t<__half>(integrator, spec);
    if (dtype == GR_BF16) return pick_tile_t<bf16_t>(integrator, spec);
    return pick_tile_t<float>(integrator, spec);
}

} // namespace gr
