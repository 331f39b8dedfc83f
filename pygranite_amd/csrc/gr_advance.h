// Step kernels of single-job launches (gr_job_advance): one thread per particle, the
// workgroups of a launch renumbered per XCD. Instantiated per storage type in
// gr_k_adv_{f32,f16,bf16}.hip (compiled in parallel).
#pragma once

#include <type_traits>

#include "gr_step.h"

namespace gr
{

template <int DIM, int INTEG, class T, int MODE, int SPEC>
__global__ void __launch_bounds__(256) k_advance(const StepParams P, uint32_t nxcd)
{
    advance_block<DIM, INTEG, T, MODE, false, SPEC>(P, nxcd);
}

// The kernel without stencil reuse held to at most 128 VGPRs, i.e. four wavefronts per SIMD
// instead of three, for wide FP16 time-paired launches of jobs that turn the reuse off.
template <int DIM, int INTEG, class T, int MODE, int SPEC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
k_advance_w4(const StepParams P, uint32_t nxcd)
{
    advance_block<DIM, INTEG, T, MODE, false, SPEC>(P, nxcd);
}

// The time-paired kernel with the stencil reuse (CACHE): a lane reloads its stencil only when
// its cell changes. The default for every time-paired launch (GR_OPT_REUSE). Held to four
// wavefronts per SIMD (128 VGPRs): the FP32 builds would take 134 and run three; a few spilled
// values cost less than the lost occupancy (ABC workload, FP32: 12.3 -> 11.6 ms per job).
template <int DIM, int INTEG, class T, int MODE, int SPEC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_advance_narrow(const StepParams P, uint32_t nxcd)
{
    advance_block<DIM, INTEG, T, MODE, true, SPEC>(P, nxcd);
}

// The time-paired kernel with a quad of lanes per particle (velocity's Q = 4) and the stencil
// reuse, for launches of few particles: the plain feature sets only.
template <int DIM, int INTEG, class T, int MODE, int SPEC>
__global__ void __launch_bounds__(256) k_advance_quad(const StepParams P, uint32_t nxcd)
{
    advance_block<DIM, INTEG, T, MODE, true, SPEC, 4>(P, nxcd);
}

typedef void (*AdvanceFn)(const StepParams, uint32_t);

// lean kernels exist generic (spec < 0) and for the plain feature set with curvature Off (0) or
// IndividualAndTotal (3), each with and without Kahan-compensated positions (+16); the full
// kernel is generic only
template <template <int, int, class, int, int> class K, int DIM, int INTEG, class T, int MODE>
struct SpecPick
{
    static AdvanceFn get(int spec)
    {
        if (spec == 0) return K<DIM, INTEG, T, MODE, 0>::fn();
        if (spec == 3) return K<DIM, INTEG, T, MODE, 3>::fn();
        if (spec == 16) return K<DIM, INTEG, T, MODE, 16>::fn();
        if (spec == 19) return K<DIM, INTEG, T, MODE, 19>::fn();
        return K<DIM, INTEG, T, MODE, -1>::fn();
    }
};
template <int DIM, int INTEG, class T, int MODE, int SPEC> struct KPlainFn
{
    static AdvanceFn fn() { return k_advance<DIM, INTEG, T, MODE, SPEC>; }
};
template <int DIM, int INTEG, class T, int MODE, int SPEC> struct KW4Fn
{
    static AdvanceFn fn() { return k_advance_w4<DIM, INTEG, T, MODE, SPEC>; }
};
template <int DIM, int INTEG, class T, int MODE, int SPEC> struct KNarrowFn
{
    static AdvanceFn fn() { return k_advance_narrow<DIM, INTEG, T, MODE, SPEC>; }
};

// quad kernels exist for the specialised feature sets only (null otherwise)
template <int DIM, int INTEG, class T>
AdvanceFn pick_quad(int spec)
{
    if (spec == 0) return k_advance_quad<DIM, INTEG, T, M_PAIRED, 0>;
    if (spec == 3) return k_advance_quad<DIM, INTEG, T, M_PAIRED, 3>;
    if (spec == 16) return k_advance_quad<DIM, INTEG, T, M_PAIRED, 16>;
    if (spec == 19) return k_advance_quad<DIM, INTEG, T, M_PAIRED, 19>;
    return nullptr;
}

template <int DIM, int INTEG, class T>
AdvanceFn pick_mode(uint32_t mode, int spec, bool wide, bool narrow, bool quad)
{
    if (mode == M_PAIRED && quad) return pick_quad<DIM, INTEG, T>(spec);
    if (mode == M_PAIRED && narrow) return SpecPick<KNarrowFn, DIM, INTEG, T, M_PAIRED>::get(spec);
    if constexpr (std::is_same<T, __half>::value)
        if (mode == M_PAIRED && wide) return SpecPick<KW4Fn, DIM, INTEG, T, M_PAIRED>::get(spec);
    if (mode == M_FULL) return k_advance<DIM, INTEG, T, M_FULL, -1>;
    if (mode == M_PAIRED) return SpecPick<KPlainFn, DIM, INTEG, T, M_PAIRED>::get(spec);
    return SpecPick<KPlainFn, DIM, INTEG, T, M_SEPARATE>::get(spec);
}

// the kernel of storage type T for a launch (wide / narrow: see launch_advance)
template <class T>
AdvanceFn pick_advance_t(uint32_t dim, uint32_t integrator, uint32_t mode, int spec, bool wide, bool narrow,
                         bool quad)
{
    if (dim == 3)
        return integrator == GR_RK4 ? pick_mode<3, GR_RK4, T>(mode, spec, wide, narrow, quad)
                                    : pick_mode<3, GR_EULER, T>(mode, spec, wide, narrow, quad);
    return integrator == GR_RK4 ? pick_mode<2, GR_RK4, T>(mode, spec, wide, narrow, quad)
                                : pick_mode<2, GR_EULER, T>(mode, spec, wide, narrow, quad);
}

// one per storage type, each in its own translation unit
AdvanceFn pick_advance_f32(uint32_t dim, uint32_t integrator, uint32_t mode, int spec, bool wide, bool narrow,
                           bool quad);
AdvanceFn pick_advance_f16(uint32_t dim, uint32_t integrator, uint32_t mode, int spec, bool wide, bool narrow,
                           bool quad);
AdvanceFn pick_advance_bf16(uint32_t dim, uint32_t integrator, uint32_t mode, int spec, bool wide, bool narrow,
                            bool quad);
// the 3D time-paired kernels with the field tile (gr_k_adv_tile.hip)
AdvanceFn pick_advance_tile(uint32_t dtype, uint32_t integrator, int spec);

} // namespace gr
