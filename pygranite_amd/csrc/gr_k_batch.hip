// The batched kernel's default build (the compiler's register allocation), the launch
// dispatcher and the compute-unit probe.
#include "gr_batch.h"

namespace gr
{

__global__ void k_cu_probe(uint32_t * out, uint32_t n)
{
    if (threadIdx.x == 0 && blockIdx.x < n) out[blockIdx.x] = cu_key();
}

int launch_cu_probe(uint32_t * out, uint32_t n, void * stream)
{
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_cu_probe, dim3(n), dim3(64), 0, hipStream_t(stream), out, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

GR_BATCH_LAUNCHER(launch_batch_default, 1, false)
GR_BATCH_OCCUPANCY(occupancy_batch_default, 1)

int batch_variant(uint32_t dim, uint32_t integrator, uint32_t dtype, int spec)
{
    const int s = spec == 0 ? 1 : spec == 3 ? 2 : spec == 16 ? 3 : spec == 19 ? 4 : 0;
    return int(12 * s + (dim == 3 ? 6 : 0) + (integrator == GR_RK4 ? 3 : 0) + (dtype > 2 ? 0 : dtype));
}

int launch_batch(const StepParams * jobs, const BatchItem * items, uint32_t * qdata, uint32_t nq,
                 const uint32_t * culut, uint32_t group, unsigned long long * work, uint32_t grid,
                 uint32_t block, uint32_t waves, bool narrow, uint32_t lds_bytes, void * stream)
{
    if (nq == 0 || grid == 0 || group == 0) return 0;
    const BatchLaunchFn fn = narrow ? launch_batch_narrow : (waves >= 4 ? launch_batch_w4 : launch_batch_default);
    return fn(jobs, items, qdata, nq, culut, group, work, grid, block, lds_bytes, stream);
}

int launch_batch_uni(uint32_t var, const StepParams * jobs, const BatchItem * items, uint32_t * qdata, uint32_t nq,
                     const uint32_t * culut, uint32_t group, unsigned long long * work, uint32_t grid,
                     uint32_t block, uint32_t lds_bytes, void * stream)
{
    if (nq == 0 || grid == 0 || group == 0) return 0;
    return (var % 12) >= 6 ? launch_batch_uni3(var, jobs, items, qdata, nq, culut, group, work, grid, block,
                                               lds_bytes, stream)
                           : launch_batch_uni2(var, jobs, items, qdata, nq, culut, group, work, grid, block,
                                               lds_bytes, stream);
}

int batch_uni_occupancy(uint32_t var, uint32_t block, uint32_t lds_bytes, int * blocks_per_cu)
{
    return (var % 12) >= 6 ? occupancy_batch_uni3(var, block, lds_bytes, blocks_per_cu)
                           : occupancy_batch_uni2(var, block, lds_bytes, blocks_per_cu);
}

int batch_occupancy(uint32_t block, uint32_t waves, uint32_t lds_bytes, int * blocks_per_cu)
{
    return waves >= 4 ? occupancy_batch_w4(block, lds_bytes, blocks_per_cu)
                      : occupancy_batch_default(block, lds_bytes, blocks_per_cu);
}

} // namespace gr
