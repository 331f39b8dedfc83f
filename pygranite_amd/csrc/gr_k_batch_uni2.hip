// The per-variant batched kernels (k_batch_uni) of the 2D item variants: one kernel per
// integrator, storage type and specialised feature set.
#include "gr_batch.h"

namespace gr
{
namespace
{

typedef void (*UniKernel)(const StepParams *, const BatchItem *, uint32_t *, uint32_t, const uint32_t *, uint32_t,
                          unsigned long long *);

// specialisation index s of batch_variant (1 .. 4) -> the kernel of that feature set
template <int INTEG, class T>
UniKernel pick_spec(uint32_t s)
{
    switch (s)
    {
        case 1: return k_batch_uni<2, INTEG, T, 0>;
        case 2: return k_batch_uni<2, INTEG, T, 3>;
        case 3: return k_batch_uni<2, INTEG, T, 16>;
        case 4: return k_batch_uni<2, INTEG, T, 19>;
        default: return nullptr;
    }
}

// variant = 12 * s + 6 * (dim == 3) + 3 * (RK4) + storage type
UniKernel pick(uint32_t var)
{
    if (var < 12 || var >= BATCH_VARIANTS || ((var % 12) >= 6) != (2 == 3)) return nullptr;
    const uint32_t s = var / 12;
    switch (var % 6)
    {
        case 0: return pick_spec<GR_EULER, float>(s);
        case 1: return pick_spec<GR_EULER, __half>(s);
        case 2: return pick_spec<GR_EULER, bf16_t>(s);
        case 3: return pick_spec<GR_RK4, float>(s);
        case 4: return pick_spec<GR_RK4, __half>(s);
        default: return pick_spec<GR_RK4, bf16_t>(s);
    }
}

} // namespace

int launch_batch_uni2(uint32_t var, const StepParams * jobs, const BatchItem * items, uint32_t * qdata, uint32_t nq,
                      const uint32_t * culut, uint32_t group, unsigned long long * work, uint32_t grid,
                      uint32_t block, uint32_t lds_bytes, void * stream)
{
    const UniKernel k = pick(var);
    if (!k) return -2;
    hipLaunchKernelGGL(k, dim3(grid), dim3(block), lds_bytes, hipStream_t(stream), jobs, items, qdata, nq, culut,
                       group, work);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int occupancy_batch_uni2(uint32_t var, uint32_t block, uint32_t lds_bytes, int * blocks_per_cu)
{
    const UniKernel k = pick(var);
    if (!k) return -2;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k, int(block), lds_bytes) == hipSuccess ? 0 : -1;
}

} // namespace gr
