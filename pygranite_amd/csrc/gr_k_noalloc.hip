// Scratch memory of the entry points that allocate nothing (gr_job_advance_noalloc): one
// file-scope device array per GPU, its address taken once per device with hipGetSymbolAddress
// when a context is created there, and a kernel that forms a launch's time table in it.
#include <mutex>

#include "gr_step.h"

namespace gr
{

// Time table of a no-allocation launch: four floats per step (t, t + dt/2, t + dt in the
// window, and "the step starts after time 0"), at most NOALLOC_MAX_STEPS steps. Every launch on
// a GPU shares its copy; calls queued on one context's stream are ordered, calls that overlap
// on one GPU (several contexts or threads) race.
static __device__ __align__(256) float scratch[4 * NOALLOC_MAX_STEPS];

namespace
{
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void * g_ptr[MAX_DEVICES] = {};
} // namespace

int noalloc_setup(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return -1;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_ptr[device]) return 0;
    // the caller made `device` current: its copy of the array is the one whose address we get
    void * p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
    {
        (void)hipGetLastError();
        return -1;
    }
    g_ptr[device] = p;
    return 0;
}

void * noalloc_scratch(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_mu);
    return g_ptr[device];
}

// The host's time table (gr_capi.cpp, time_table) for a launch inside one window, formed on the
// device: thread 0 accumulates the step start times in FP32 and FP64 sequentially (Time += dt,
// as the reference's host loop does, tpp:990), then every thread maps its steps' times into the
// window with the same operations (unit_interval, tpp:127-144). Correctly rounded division on
// both sides makes the table equal to the host's bit for bit.
__global__ void __launch_bounds__(256) k_time_table(float * __restrict__ out, uint32_t nsteps, uint32_t flags,
                                                    float dt, float tb, float tf, float t_start,
                                                    double tb64, double tf64, double t_start64)
{
    // thread 0 leaves step s's start times in step s's own 16 bytes: the FP64 time in floats
    // 0-1, the FP32 time in float 2; the thread that finishes step s reads them first
    double * t64s = reinterpret_cast<double *>(out);
    if (threadIdx.x == 0)
    {
        float t32 = t_start;
        double t64 = t_start64;
        for (uint32_t s = 0; s < nsteps; ++s)
        {
            t64s[2 * size_t(s)] = t64;
            out[4 * size_t(s) + 2] = t32;
            t32 += dt;
            t64 += double(dt);
        }
    }
    __threadfence_block();
    __syncthreads();
    const bool use64 = (flags & F_TIME64) != 0;
    const float half_dt = 0.5f * dt;
    for (uint32_t s = threadIdx.x; s < nsteps; s += blockDim.x)
    {
        const float time32 = out[4 * size_t(s) + 2];
        const double time64 = t64s[2 * size_t(s)];
        auto unit = [&](float add) -> float {
            if (use64) return float((time64 + double(add) - tb64) / (tf64 - tb64));
            const float x = time32 + add;
            return (x - tb) / (tf - tb);
        };
        float t = 0.f;
        if (flags & F_DYNAMIC_ANY) t = unit(0.f);
        float t1 = t, t2 = t;
        if (flags & F_WIND_DYNAMIC)
        {
            t1 = unit(half_dt);
            t2 = unit(dt);
        }
        const float pos = (use64 ? time64 > 0.0 : time32 > 0.f) ? 1.f : 0.f;
        out[4 * size_t(s) + 0] = t;
        out[4 * size_t(s) + 1] = t1;
        out[4 * size_t(s) + 2] = t2;
        out[4 * size_t(s) + 3] = pos;
    }
}

int launch_time_table(const StepParams & P, float * out, void * stream)
{
    if (P.nsteps == 0) return 0;
    if (P.nsteps > NOALLOC_MAX_STEPS || !out) return -1;
    hipLaunchKernelGGL(k_time_table, dim3(1), dim3(256), 0, hipStream_t(stream), out, P.nsteps, P.flags, P.dt,
                       P.tb, P.tf, P.t_start, P.tb64, P.tf64, P.t_start64);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

} // namespace gr
