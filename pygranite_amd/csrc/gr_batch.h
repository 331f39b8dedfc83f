// The persistent batched kernel (gr_batch_run): many jobs of any dimension, integrator and
// storage type in one launch, wavefronts pulling one-wavefront items from per-CU-group queues.
// Its three builds (compiler register allocation, at most 128 VGPRs, stencil reuse for narrow
// launches) are instantiated in gr_k_batch*.hip, compiled in parallel.
#pragma once

#include "gr_step.h"

namespace gr
{

// Persistent batched kernel: workgroups pull (job, particle block) items from a global counter
// until the list is exhausted; each item runs its job's whole span of steps. Jobs of different
// dimension, integrator and storage type share the launch; the variant is uniform per item.
template <int DIM, int INTEG, class T, bool CACHE, int SPEC>
__device__ __forceinline__ uint32_t run_item(const StepParams & P, uint32_t pid)
{
    // the host builds a paired copy only for jobs that use it (dynamic wind, paired option)
    if (P.wind.pair) return advance_one<DIM, INTEG, T, M_PAIRED, CACHE, SPEC>(P, pid);
    return advance_one<DIM, INTEG, T, M_SEPARATE, false, SPEC>(P, pid);
}

// Quad items (variant BATCH_QUAD + v): a quad of lanes per particle (velocity's Q = 4), in the
// time-paired layout with the stencil reuse; the host plans them only for launches of few live
// particles whose jobs have a specialised feature set and a paired copy smaller than 4 GiB.
template <int DIM, int INTEG, class T, int SPEC>
__device__ __forceinline__ uint32_t run_quad_item(const StepParams & P, uint32_t pid)
{
    return advance_one<DIM, INTEG, T, M_PAIRED, true, SPEC, 4>(P, pid);
}

// the five feature sets of item variant V (V + 12 * specialisation, see batch_variant)
#define GR_BATCH_CASES(V, DIM, INTEG, T)                                                            \
    case V: w = run_item<DIM, INTEG, T, CACHE, -1>(P, pid); break;                                 \
    case V + 12: w = run_item<DIM, INTEG, T, CACHE, 0>(P, pid); break;                             \
    case V + 24: w = run_item<DIM, INTEG, T, CACHE, 3>(P, pid); break;                             \
    case V + 36: w = run_item<DIM, INTEG, T, CACHE, 16>(P, pid); break;                            \
    case V + 48: w = run_item<DIM, INTEG, T, CACHE, 19>(P, pid); break;
// the four specialised feature sets of quad variant BATCH_QUAD + V + 12 * specialisation
#define GR_BATCH_QUAD_CASES(V, DIM, INTEG, T)                                                       \
    case V + 12: w = run_quad_item<DIM, INTEG, T, 0>(P, pid); break;                               \
    case V + 24: w = run_quad_item<DIM, INTEG, T, 3>(P, pid); break;                               \
    case V + 36: w = run_quad_item<DIM, INTEG, T, 16>(P, pid); break;                              \
    case V + 48: w = run_quad_item<DIM, INTEG, T, 19>(P, pid); break;

// Identity of the compute unit running the calling wavefront: XCC, shader engine, shader array
// and CU fields of the HW_ID / XCC_ID hardware registers packed into CU_KEY_BITS bits.
__device__ __forceinline__ uint32_t cu_key()
{
    const uint32_t hw = __builtin_amdgcn_s_getreg((31 << 11) | 4); // HW_REG_HW_ID, bits 31:0
    const uint32_t xcc = __builtin_amdgcn_s_getreg((3 << 11) | 20); // HW_REG_XCC_ID, bits 3:0
    return ((xcc & 7u) << 8) | (((hw >> 13) & 7u) << 5) | (((hw >> 12) & 1u) << 4) | ((hw >> 8) & 15u);
}

// Items are one wavefront of particles (a quarter of one for quad items, which only the
// stencil-reuse builds run) and every wavefront takes its own: a wavefront whose
// particles are all dead moves on at once instead of waiting at a workgroup barrier. There is
// one queue per group of `group` compute units (queue q = items [qbegin[q], qbegin[q + 1]), a
// compact region of the Morton-ordered particle sets, the same region for every job); the
// wavefronts of a group drain its queue first, so the group's caches hold the field tiles of
// one region (one group per XCD: its L2; one per CU: its L1). A drained queue
// clears its bit in masks[q / 32]; wavefronts then move to the nearest queue whose bit is still
// set, and exit when no bit is left. Every wavefront accumulates its integrated particle-steps
// (Body::run(job, item, lane), the item run by the calling lane) and adds them once, into one
// of WORK_SLOTS counters.
template <class Body>
__device__ __forceinline__ void batch_loop(const StepParams * __restrict__ jobs, const BatchItem * __restrict__ items,
                                           uint32_t * qdata, uint32_t nq, const uint32_t * __restrict__ culut,
                                           uint32_t group, unsigned long long * work)
{
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t * counters = qdata;                                  // nq x QUEUE_STRIDE
    const uint32_t * qbegin = qdata + size_t(nq) * QUEUE_STRIDE;  // nq + 1
    uint32_t * masks = qdata + size_t(nq) * QUEUE_STRIDE + nq + 1; // ceil(nq / 32)
    const uint32_t ng = (nq + 31) / 32;
    uint32_t home = culut ? culut[cu_key()] : 0xffffffffu;
    home = home != 0xffffffffu ? home / group : nq;
    if (home >= nq) home = blockIdx.x % nq;
    uint32_t q = home;
    uint32_t acc = 0;
    for (;;)
    {
        uint32_t v = 0;
        if (lane == 0) v = atomicAdd(&counters[size_t(q) * QUEUE_STRIDE], 1u);
        // every lane is active here, so the first active lane is lane 0; the result is
        // wavefront-uniform (a scalar register), which keeps the item and job loads scalar
        const uint32_t it = qbegin[q] + __builtin_amdgcn_readfirstlane(v);
        if (it < qbegin[q + 1])
        {
            const BatchItem b = items[it];
            acc += Body::run(jobs[b.job], b, lane);
            continue;
        }
        // queue q is drained (its counter only grows): clear its bit, find the nearest queue
        // that still has one, starting in q's own mask word
        if (lane == 0) atomicAnd(&masks[q >> 5], ~(1u << (q & 31u)));
        uint32_t next = 0xffffffffu;
        for (uint32_t k = 0; k < ng; ++k)
        {
            const uint32_t g = (q >> 5) + k < ng ? (q >> 5) + k : (q >> 5) + k - ng;
            uint32_t m = 0;
            if (lane == 0) m = __hip_atomic_load(&masks[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            m = __builtin_amdgcn_readfirstlane(m);
            if (!m) continue;
            uint32_t bit;
            if (k == 0)
            {
                const uint32_t sh = q & 31u; // first set bit at or after q's position, cyclically
                const uint32_t rot = sh ? ((m >> sh) | (m << (32u - sh))) : m;
                bit = (uint32_t(__builtin_ctz(rot)) + sh) & 31u;
            }
            else
                bit = uint32_t(__builtin_ctz(m));
            next = g * 32u + bit;
            break;
        }
        if (next >= nq) break; // every queue is drained
        q = next;
    }
    if (work) wave_add_work(work + size_t(home % WORK_SLOTS) * WORK_STRIDE, acc);
}

// the item body of the kernel of every variant
template <bool CACHE> struct BatchEvery
{
    static __device__ __forceinline__ uint32_t run(const StepParams & P, const BatchItem b, uint32_t lane)
    {
        uint32_t w = 0;
        if constexpr (CACHE)
            if (b.variant >= BATCH_QUAD)
            {
                // a quad item: lanes 4k .. 4k + 3 hold its particle k (whole quads active)
                const uint32_t pid = b.first + (lane >> 2);
                if ((lane >> 2) < b.count)
                {
                    switch (b.variant - BATCH_QUAD)
                    {
                        GR_BATCH_QUAD_CASES(0, 2, GR_EULER, float)
                        GR_BATCH_QUAD_CASES(1, 2, GR_EULER, __half)
                        GR_BATCH_QUAD_CASES(2, 2, GR_EULER, bf16_t)
                        GR_BATCH_QUAD_CASES(3, 2, GR_RK4, float)
                        GR_BATCH_QUAD_CASES(4, 2, GR_RK4, __half)
                        GR_BATCH_QUAD_CASES(5, 2, GR_RK4, bf16_t)
                        GR_BATCH_QUAD_CASES(6, 3, GR_EULER, float)
                        GR_BATCH_QUAD_CASES(7, 3, GR_EULER, __half)
                        GR_BATCH_QUAD_CASES(8, 3, GR_EULER, bf16_t)
                        GR_BATCH_QUAD_CASES(9, 3, GR_RK4, float)
                        GR_BATCH_QUAD_CASES(10, 3, GR_RK4, __half)
                        GR_BATCH_QUAD_CASES(11, 3, GR_RK4, bf16_t)
                        default: w = 0; break;
                    }
                }
                return w;
            }
        const uint32_t pid = b.first + lane;
        if (lane < b.count)
        {
            switch (b.variant)
            {
                GR_BATCH_CASES(0, 2, GR_EULER, float)
                GR_BATCH_CASES(1, 2, GR_EULER, __half)
                GR_BATCH_CASES(2, 2, GR_EULER, bf16_t)
                GR_BATCH_CASES(3, 2, GR_RK4, float)
                GR_BATCH_CASES(4, 2, GR_RK4, __half)
                GR_BATCH_CASES(5, 2, GR_RK4, bf16_t)
                GR_BATCH_CASES(6, 3, GR_EULER, float)
                GR_BATCH_CASES(7, 3, GR_EULER, __half)
                GR_BATCH_CASES(8, 3, GR_EULER, bf16_t)
                GR_BATCH_CASES(9, 3, GR_RK4, float)
                GR_BATCH_CASES(10, 3, GR_RK4, __half)
                GR_BATCH_CASES(11, 3, GR_RK4, bf16_t)
                default: w = 0; break;
            }
        }
        return w;
    }
};

// The batched kernel of every item variant. WAVES is the minimum number of wavefronts per SIMD
// the register allocation must allow (1: no limit; 4: at most 128 VGPRs); CACHE: the
// stencil-reuse build, which also runs quad items.
template <int WAVES, bool CACHE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WAVES)))
k_batch(const StepParams * __restrict__ jobs, const BatchItem * __restrict__ items,
        uint32_t * qdata, uint32_t nq, const uint32_t * __restrict__ culut, uint32_t group,
        unsigned long long * work)
{
    batch_loop<BatchEvery<CACHE>>(jobs, items, qdata, nq, culut, group, work);
}

// the item body of the per-variant kernel
template <int DIM, int INTEG, class T, int SPEC> struct BatchOne
{
    static __device__ __forceinline__ uint32_t run(const StepParams & P, const BatchItem b, uint32_t lane)
    {
        if (b.variant >= BATCH_QUAD)
            return (lane >> 2) < b.count ? run_quad_item<DIM, INTEG, T, SPEC>(P, b.first + (lane >> 2)) : 0u;
        return lane < b.count ? run_item<DIM, INTEG, T, true, SPEC>(P, b.first + lane) : 0u;
    }
};

// The batched kernel of one specialised item variant (dimension, integrator, storage type,
// feature set SPEC >= 0; its quad items too), with the stencil reuse, held like the single-job
// stencil-reuse kernels to four wavefronts per SIMD: for launches whose spans all share that
// variant. The kernel of every variant carries the register demand of the largest one (the
// FP32 Kahan build), which leaves three wavefronts per SIMD.
template <int DIM, int INTEG, class T, int SPEC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
k_batch_uni(const StepParams * __restrict__ jobs, const BatchItem * __restrict__ items,
            uint32_t * qdata, uint32_t nq, const uint32_t * __restrict__ culut, uint32_t group,
            unsigned long long * work)
{
    batch_loop<BatchOne<DIM, INTEG, T, SPEC>>(jobs, items, qdata, nq, culut, group, work);
}

// launchers of the three builds (one translation unit each)
#define GR_BATCH_LAUNCHER(NAME, WAVES, CACHE)                                                       \
    int NAME(const StepParams * jobs, const BatchItem * items, uint32_t * qdata, uint32_t nq,       \
             const uint32_t * culut, uint32_t group, unsigned long long * work, uint32_t grid,      \
             uint32_t block, uint32_t lds_bytes, void * stream)                                     \
    {                                                                                               \
        hipLaunchKernelGGL((k_batch<WAVES, CACHE>), dim3(grid), dim3(block), lds_bytes,             \
                           hipStream_t(stream), jobs, items, qdata, nq, culut, group, work);        \
        return hipGetLastError() == hipSuccess ? 0 : -1;                                            \
    }
#define GR_BATCH_OCCUPANCY(NAME, WAVES)                                                             \
    int NAME(uint32_t block, uint32_t lds_bytes, int * blocks_per_cu)                               \
    {                                                                                               \
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_batch<WAVES, false>,   \
                                                            int(block), lds_bytes) == hipSuccess    \
                   ? 0                                                                              \
                   : -1;                                                                            \
    }
// the per-variant kernels of one dimension (gr_k_batch_uni{2,3}.hip): the launcher and the
// occupancy of specialised variant `var` (12 .. 59, see batch_variant) when it has that
// dimension, -2 otherwise
int launch_batch_uni2(uint32_t var, const StepParams * jobs, const BatchItem * items, uint32_t * qdata, uint32_t nq,
                      const uint32_t * culut, uint32_t group, unsigned long long * work, uint32_t grid,
                      uint32_t block, uint32_t lds_bytes, void * stream);
int launch_batch_uni3(uint32_t var, const StepParams * jobs, const BatchItem * items, uint32_t * qdata, uint32_t nq,
                      const uint32_t * culut, uint32_t group, unsigned long long * work, uint32_t grid,
                      uint32_t block, uint32_t lds_bytes, void * stream);
int occupancy_batch_uni2(uint32_t var, uint32_t block, uint32_t lds_bytes, int * blocks_per_cu);
int occupancy_batch_uni3(uint32_t var, uint32_t block, uint32_t lds_bytes, int * blocks_per_cu);
typedef int (*BatchLaunchFn)(const StepParams *, const BatchItem *, uint32_t *, uint32_t, const uint32_t *,
                             uint32_t, unsigned long long *, uint32_t, uint32_t, uint32_t, void *);
int launch_batch_default(const StepParams * jobs, const BatchItem * items, uint32_t * qdata, uint32_t nq,
                         const uint32_t * culut, uint32_t group, unsigned long long * work, uint32_t grid,
                         uint32_t block, uint32_t lds_bytes, void * stream);
int launch_batch_w4(const StepParams * jobs, const BatchItem * items, uint32_t * qdata, uint32_t nq,
                    const uint32_t * culut, uint32_t group, unsigned long long * work, uint32_t grid,
                    uint32_t block, uint32_t lds_bytes, void * stream);
int launch_batch_narrow(const StepParams * jobs, const BatchItem * items, uint32_t * qdata, uint32_t nq,
                        const uint32_t * culut, uint32_t group, unsigned long long * work, uint32_t grid,
                        uint32_t block, uint32_t lds_bytes, void * stream);
int occupancy_batch_default(uint32_t block, uint32_t lds_bytes, int * blocks_per_cu);
int occupancy_batch_w4(uint32_t block, uint32_t lds_bytes, int * blocks_per_cu);

} // namespace gr
