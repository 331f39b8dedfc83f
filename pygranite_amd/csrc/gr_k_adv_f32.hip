// Single-job step kernels over FP32 slices, and the launcher of every single-job step kernel.
#include "gr_advance.h"

namespace gr
{

AdvanceFn pick_advance_f32(uint32_t dim, uint32_t integrator, uint32_t mode, int spec, bool wide, bool narrow,
                           bool quad)
{
    return pick_advance_t<float>(dim, integrator, mode, spec, wide, narrow, quad);
}

int launch_advance(const StepParams & P, uint32_t dim, uint32_t integrator, uint32_t mode, int spec,
                   uint32_t block, uint32_t nxcd, uint64_t nrun, bool wide, bool narrow, bool quad,
                   uint32_t lds_bytes, void * stream, bool tile)
{
    AdvanceFn fn = tile ? (dim == 3 && mode == M_PAIRED && !quad ? pick_advance_tile(P.wind.dtype, integrator, spec)
                                                                 : nullptr)
                   : P.wind.dtype == GR_F16  ? pick_advance_f16(dim, integrator, mode, spec, wide, narrow, quad)
                   : P.wind.dtype == GR_BF16 ? pick_advance_bf16(dim, integrator, mode, spec, wide, narrow, quad)
                                             : pick_advance_f32(dim, integrator, mode, spec, wide, narrow, quad);
    if (!fn) return -1;
    if (block == 0 || block > 256) block = 256;
    const uint64_t threads = std::min<uint64_t>(nrun, P.n) * (quad ? 4u : 1u);
    if (threads >= (1ull << 31)) return -1;
    const uint32_t grid = uint32_t((threads + block - 1) / block);
    if (grid == 0) return 0;
    hipLaunchKernelGGL(fn, dim3(grid), dim3(block), lds_bytes, hipStream_t(stream), P, nxcd);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

} // namespace gr
