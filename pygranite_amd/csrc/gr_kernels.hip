// gfx950 kernels of the trajectory engine.
//
// k_advance runs `nsteps` integration steps of one window for every particle, keeping the
// particle state in registers across steps and writing saved columns to a device ring. It
// replaces the one-launch-per-step `granite::integrate` kernel of the reference
// (source/granite/TrajectoryIntegrator.tpp:108-457) together with the host step loop that
// launched it (tpp:973-994). Sampling is done in software from node-interleaved FP32 / FP16 /
// BF16 slices (one vector load per stencil corner and slice) instead of texture fetches
// (TrajectoryIntegrator.icc:22-129).
//
// Arithmetic follows the reference's operation order and is compiled without FP contraction,
// so the FP32 form performs exactly the IEEE operations of the step (reference semantics);
// only the sampler weights differ from the texture unit (exact here, optionally quantised to
// 1/256 for emulation).
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hipcub/device/device_scan.hpp>
#include <stdint.h>

#include <algorithm>
#include <type_traits>

#include "gr_internal.h"

namespace gr
{

// --------------------------------------------------------------------------------------------
// storage types
// --------------------------------------------------------------------------------------------
struct bf16_t
{
    uint16_t bits;
};

__device__ __forceinline__ float to_f(float v) { return v; }
__device__ __forceinline__ float to_f(__half v) { return __half2float(v); }
__device__ __forceinline__ float to_f(bf16_t v) { return __uint_as_float(uint32_t(v.bits) << 16); }

__device__ __forceinline__ float h_lo(uint32_t w) { return __half2float(__ushort_as_half(uint16_t(w & 0xffffu))); }
__device__ __forceinline__ float h_hi(uint32_t w) { return __half2float(__ushort_as_half(uint16_t(w >> 16))); }
__device__ __forceinline__ float b_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float b_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }

// Global-address-space views of field memory: the slot pointers travel through structs, so
// without the cast the compiler emits flat loads.
#if defined(__HIP_DEVICE_COMPILE__)
#define GR_LOAD(T, b, i) (((const __attribute__((address_space(1))) T *)(b))[i])
#else
#define GR_LOAD(T, b, i) (reinterpret_cast<const T *>(b)[i])
#endif

// Read-only launch data (time tables, window tables) through the constant address space: the
// compiler may then use scalar loads, which it must not do for global memory the kernel could
// also write through another pointer.
#if defined(__HIP_DEVICE_COMPILE__)
#define GR_CONST(T, b, i) (((const __attribute__((address_space(4))) T *)(b))[i])
#else
#define GR_CONST(T, b, i) (reinterpret_cast<const T *>(b)[i])
#endif

// the same for state and output arrays reached through structs in global memory (the batched
// kernel's job table): G(p)[i] is a global-address-space access
#if defined(__HIP_DEVICE_COMPILE__)
template <class T> __device__ __forceinline__ __attribute__((address_space(1))) T * G(T * p)
{
    return (__attribute__((address_space(1))) T *)p;
}
#else
template <class T> __host__ __device__ T * G(T * p) { return p; }
#endif

// Loads the NC components of node `node` (cstride = 4 for NC = 3, else NC).
template <class T, int NC> struct NodeLoad;

template <> struct NodeLoad<float, 3>
{
    static __device__ __forceinline__ void get(const void * b, uint32_t node, float (&v)[3])
    {
        const float4 x = GR_LOAD(float4, b, node);
        v[0] = x.x; v[1] = x.y; v[2] = x.z;
    }
};
template <> struct NodeLoad<float, 2>
{
    static __device__ __forceinline__ void get(const void * b, uint32_t node, float (&v)[2])
    {
        const float2 x = GR_LOAD(float2, b, node);
        v[0] = x.x; v[1] = x.y;
    }
};
template <> struct NodeLoad<__half, 3>
{
    static __device__ __forceinline__ void get(const void * b, uint32_t node, float (&v)[3])
    {
        const uint2 x = GR_LOAD(uint2, b, node);
        v[0] = h_lo(x.x); v[1] = h_hi(x.x); v[2] = h_lo(x.y);
    }
};
template <> struct NodeLoad<__half, 2>
{
    static __device__ __forceinline__ void get(const void * b, uint32_t node, float (&v)[2])
    {
        const uint32_t x = GR_LOAD(uint32_t, b, node);
        v[0] = h_lo(x); v[1] = h_hi(x);
    }
};
template <> struct NodeLoad<bf16_t, 3>
{
    static __device__ __forceinline__ void get(const void * b, uint32_t node, float (&v)[3])
    {
        const uint2 x = GR_LOAD(uint2, b, node);
        v[0] = b_lo(x.x); v[1] = b_hi(x.x); v[2] = b_lo(x.y);
    }
};
template <> struct NodeLoad<bf16_t, 2>
{
    static __device__ __forceinline__ void get(const void * b, uint32_t node, float (&v)[2])
    {
        const uint32_t x = GR_LOAD(uint32_t, b, node);
        v[0] = b_lo(x); v[1] = b_hi(x);
    }
};

// generic scalar component load (volumes, topography)
__device__ __forceinline__ float load_comp(const void * b, uint32_t dtype, uint64_t idx)
{
    if (dtype == GR_F16) return __half2float(GR_LOAD(__half, b, idx));
    if (dtype == GR_BF16) return __uint_as_float(uint32_t(GR_LOAD(uint16_t, b, idx)) << 16);
    return GR_LOAD(float, b, idx);
}

// --------------------------------------------------------------------------------------------
// sampling helpers
// --------------------------------------------------------------------------------------------
// a + t * (b - a) as one fused multiply-add (the contraction nvcc applies to the reference's
// lerp, LinearAlgebra.hpp:617-621; the oracle evaluates the same form)
__device__ __forceinline__ float lerp(float a, float b, float t) { return __builtin_fmaf(t, b - a, a); }

// texel coordinate -> (index, weight). Clamping to [-1, n] leaves every tap outside the grid
// and keeps the conversion defined (NaN -> -1).
__device__ __forceinline__ void split(float a, uint32_t n, uint32_t wmode, int & i, float & w)
{
    a = fminf(fmaxf(a, -1.f), float(n));
    const float f = floorf(a);
    i = int(f);
    w = a - f;
    if (wmode == GR_WEIGHTS_Q8_ROUND) w = floorf(w * 256.f + 0.5f) / 256.f;
    else if (wmode == GR_WEIGHTS_Q8_TRUNC) w = floorf(w * 256.f) / 256.f;
}

struct Taps
{
    uint32_t node[8]; // corner nodes (x fastest: bit0 = +x, bit1 = +y, bit2 = +z)
    uint32_t valid;   // bit per corner (border addressing)
    float wx, wy, wz;
};

// Stencil for 3D (DIM = 3) or 2D (DIM = 2, 4 corners). Border addressing zeroes taps outside
// [0, n-1]; clamp addressing clamps the index.
template <int DIM>
__device__ __forceinline__ void taps(const float * arg, uint32_t nx, uint32_t ny, uint32_t nz,
                                     uint32_t address, uint32_t wmode, Taps & T)
{
    int i, j, k = 0;
    split(arg[0], nx, wmode, i, T.wx);
    split(arg[1], ny, wmode, j, T.wy);
    T.wz = 0.f;
    if (DIM == 3) split(arg[2], nz, wmode, k, T.wz);
    const int xi[2] = {i, i + 1}, yj[2] = {j, j + 1}, zk[2] = {k, k + 1};
    T.valid = 0;
#pragma unroll
    for (int c = 0; c < (DIM == 3 ? 8 : 4); ++c)
    {
        const int a = xi[c & 1], b = yj[(c >> 1) & 1], d = zk[(c >> 2) & 1];
        const bool ok = a >= 0 && a < int(nx) && b >= 0 && b < int(ny) && (DIM == 2 || (d >= 0 && d < int(nz)));
        const int ca = min(max(a, 0), int(nx) - 1);
        const int cb = min(max(b, 0), int(ny) - 1);
        const int cd = DIM == 3 ? min(max(d, 0), int(nz) - 1) : 0;
        T.node[c] = (uint32_t(cd) * ny + uint32_t(cb)) * nx + uint32_t(ca);
        if (ok || address == 1) T.valid |= 1u << c;
    }
}

// Nested lerps x -> y -> z of the corner values (same order as the oracle's restatement).
template <int DIM, int NC>
__device__ __forceinline__ void combine(const float (&v)[8][NC], const Taps & T, float (&out)[NC])
{
#pragma unroll
    for (int c = 0; c < NC; ++c)
    {
        const float c00 = lerp(v[0][c], v[1][c], T.wx);
        const float c10 = lerp(v[2][c], v[3][c], T.wx);
        if (DIM == 3)
        {
            const float c01 = lerp(v[4][c], v[5][c], T.wx);
            const float c11 = lerp(v[6][c], v[7][c], T.wx);
            const float c0 = lerp(c00, c10, T.wy);
            const float c1 = lerp(c01, c11, T.wy);
            out[c] = lerp(c0, c1, T.wz);
        }
        else
            out[c] = lerp(c00, c10, T.wy);
    }
}

// FP16 corner values enter the first (x) lerp without separate conversions: with f16 operands
// fma(b, 1, -a) is b - a and fma(t, d, a) is lerp's fused form, each rounded once after the
// exact f16 -> f32 conversion, i.e. the same IEEE results as lerp((float)a, (float)b, t); the
// compiler forms v_fma_mix_f32 from them. `one` is 1.0 hidden from constant folding (which would
// turn the first fma back into a subtraction of converted values).
typedef _Float16 f16_t;
__device__ __forceinline__ f16_t f16_lo(uint32_t w) { return __builtin_bit_cast(f16_t, uint16_t(w & 0xffffu)); }
__device__ __forceinline__ f16_t f16_hi(uint32_t w) { return __builtin_bit_cast(f16_t, uint16_t(w >> 16)); }
__device__ __forceinline__ float lerp_h(f16_t a, f16_t b, float t, f16_t one)
{
    const float d = __builtin_fmaf(float(b), float(one), -float(a));
    return __builtin_fmaf(t, d, float(a));
}

// Component c of slice s (0 back, 1 front) of a time-paired FP16 node: 3D words
// [u0 v0 | w0 u1 | v1 w1 | 0 0], 2D words [u0 v0 | u1 v1].
template <int DIM>
__device__ __forceinline__ f16_t pair_h(const uint4 & a, int s, int c)
{
    if (DIM == 3)
    {
        const int e = 3 * s + c; // element 0..5
        const uint32_t w = e < 2 ? a.x : (e < 4 ? a.y : a.z);
        return (e & 1) ? f16_hi(w) : f16_lo(w);
    }
    const uint32_t w = s ? a.y : a.x;
    return c ? f16_hi(w) : f16_lo(w);
}

// Trilinear (bilinear) sample of both slices of a paired FP16 stencil, same operation order as
// combine(): x lerps, then y, then z.
template <int DIM>
__device__ __forceinline__ void combine_pair_h(const uint4 (&a)[8], const Taps & T, f16_t one,
                                               float (&b)[DIM], float (&f)[DIM])
{
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int c = 0; c < DIM; ++c)
        {
            const float c00 = lerp_h(pair_h<DIM>(a[0], s, c), pair_h<DIM>(a[1], s, c), T.wx, one);
            const float c10 = lerp_h(pair_h<DIM>(a[2], s, c), pair_h<DIM>(a[3], s, c), T.wx, one);
            float r;
            if (DIM == 3)
            {
                const float c01 = lerp_h(pair_h<DIM>(a[4], s, c), pair_h<DIM>(a[5], s, c), T.wx, one);
                const float c11 = lerp_h(pair_h<DIM>(a[6], s, c), pair_h<DIM>(a[7], s, c), T.wx, one);
                const float c0 = lerp(c00, c10, T.wy);
                const float c1 = lerp(c01, c11, T.wy);
                r = lerp(c0, c1, T.wz);
            }
            else
                r = lerp(c00, c10, T.wy);
            if (s == 0) b[c] = r;
            else f[c] = r;
        }
}

template <int DIM, class T, int NC>
__device__ __forceinline__ void sample_slot(const void * base, const Taps & S, float (&out)[NC])
{
    constexpr int NCOR = DIM == 3 ? 8 : 4;
    float v[8][NC];
#pragma unroll
    for (int c = 0; c < NCOR; ++c)
    {
        float x[NC];
        NodeLoad<T, NC>::get(base, S.node[c], x);
        const bool ok = (S.valid >> c) & 1u;
#pragma unroll
        for (int m = 0; m < NC; ++m) v[c][m] = ok ? x[m] : 0.f;
    }
    combine<DIM, NC>(v, S, out);
}

// generic scalar field sample (component `comp` of a field view), used for volumes and
// topography; bilinear/trilinear with the view's addressing
template <int DIM>
__device__ float sample_scalar(const void * base, const FieldView & f, uint32_t comp,
                               const float * arg, uint32_t wmode)
{
    Taps S;
    taps<DIM>(arg, f.nx, f.ny, f.nz, f.address, wmode, S);
    float v[8][1];
#pragma unroll
    for (int c = 0; c < (DIM == 3 ? 8 : 4); ++c)
    {
        const float x = load_comp(base, f.dtype, uint64_t(S.node[c]) * f.cstride + comp);
        v[c][0] = ((S.valid >> c) & 1u) ? x : 0.f;
    }
    float o[1];
    combine<DIM, 1>(v, S, o);
    return o[0];
}

// --------------------------------------------------------------------------------------------
// expression interpreter (wavefront-uniform: the program lives in the kernel arguments, every
// lane executes the same instruction stream)
// --------------------------------------------------------------------------------------------
// Programs run in their device form (lower_program): binary operations with a leaf operand read
// the leaf directly, and the two top stack values live in registers; deeper entries live in
// LDS, one column per lane (entry d of lane l at stk[d * blockDim.x + l]), so the wave-uniform
// stack depth turns into a fixed LDS offset instead of dynamically indexed registers. Reference
// slots come from a uniform switch over the step's registers (`ref`).
typedef __attribute__((address_space(3))) float lds_float;

// One LDS area per kernel (a non-template function, so every inlined step variant of the
// batched kernel shares it): 19 reference slots (volumes and constants of the full kernel) +
// (stack depth - 1) stack entries per lane.
__device__ __forceinline__ lds_float * expr_lds()
{
    __shared__ float buf[(19 + GR_MAX_STACK - 1) * 256];
    return (lds_float *)buf;
}

// a op b for k = 0 add, 1 sub, 2 mul (branch-free: a - b is a + (-b) exactly), 3 div
__device__ __forceinline__ float apply_op(uint32_t k, float a, float b)
{
    if (k == 3) return a / b; // uniform branch
    const float s = a + (k == 1 ? -b : b);
    return k == 2 ? a * b : s;
}

// Reference slots x y z cv u v w of a step as one vector value: element d for a uniform d is an
// indexed register read (no branches, no memory). Nine elements (the last two zero): a vector
// of more than eight is read with GPR index mode (s_set_gpr_idx_on, one move), where one of
// eight would be expanded into a chain of eight compares and selects per read.
typedef float refvec_t __attribute__((ext_vector_type(9)));

// rv: reference slots x y z cv u v w of the step; R: this lane's LDS reference column (volumes
// and constants, full kernel only)
template <bool FULL>
__device__ __forceinline__ float run_program(const Program & pr, const refvec_t rv,
                                             const lds_float * R, lds_float * stk, uint32_t lane,
                                             uint32_t bs)
{
    auto refv = [&](uint32_t d) -> float {
        if (!FULL) return rv[d & 7u];
        return d < 7 ? rv[d] : R[(d - 7) * bs];
    };
    float acc = 0.f, acc2 = 0.f;
    uint32_t depth = 0; // values on the stack (uniform)
    for (uint32_t i = 0; i < pr.n; ++i)
    {
        const uint32_t w = pr.code[i];
        const uint32_t cls = (w >> 4) & 0xfu, sub = w & 0xfu, arg = w >> 8;
        if (cls == DC_LEAF) // the most common instruction first
        {
            const float x = (sub & 8) ? pr.imm[arg] : refv(arg);
            const bool rev = sub & 4;
            acc = apply_op(sub & 3, rev ? x : acc, rev ? acc : x);
        }
        else if (cls == DC_PUSH || cls == DC_DOT)
        {
            float x;
            if (cls == DC_DOT) // (r0 * r1) + (r2 * r3) [+ (r4 * r5)], register slots [, sqrt]
            {
                x = (rv[arg & 7u] * rv[(arg >> 3) & 7u]) + (rv[(arg >> 6) & 7u] * rv[(arg >> 9) & 7u]);
                if ((sub & 3) == 3) x = x + (rv[(arg >> 12) & 7u] * rv[(arg >> 15) & 7u]);
                if (sub & 4) x = sqrtf(x);
            }
            else
                x = sub ? pr.imm[arg] : refv(arg);
            if (depth >= 2) stk[(depth - 2) * bs + lane] = acc2;
            acc2 = acc;
            acc = x;
            ++depth;
        }
        else if (cls == DC_BINARY)
        {
            const bool rev = sub & 8;
            const float a = rev ? acc : acc2, b = rev ? acc2 : acc;
            if ((sub & 7) == 4) acc = atan2f(a, b);
            else acc = apply_op(sub & 3, a, b);
            --depth;
            if (depth >= 2) acc2 = stk[(depth - 2) * bs + lane];
        }
        else
        {
            switch (sub)
            {
                case 0: acc = sqrtf(acc); break;
                case 1: acc = expf(acc); break;
                case 2: acc = logf(acc); break;
                case 3: acc = sinf(acc); break;
                case 4: acc = cosf(acc); break;
                case 5: acc = tanf(acc); break;
                case 6: acc = asinf(acc); break;
                case 7: acc = acosf(acc); break;
                case 8: acc = atanf(acc); break;
                default: acc = fabsf(acc); break;
            }
        }
    }
    return acc;
}

// --------------------------------------------------------------------------------------------
// the step kernel
// --------------------------------------------------------------------------------------------
struct Ctx
{
    const StepParams & P;
};

// Time-paired node layout: both slices of a node in one vector, 3D as
// [u0 v0 w0 u1 v1 w1 0 0], 2D as [u0 v0 u1 v1] (slice 0 = back, 1 = front).
template <class T, int NC> struct PairLoad;

template <> struct PairLoad<float, 3>
{
    static __device__ __forceinline__ void get(const void * b, uint32_t node, float (&x)[3], float (&y)[3])
    {
        const float4 a = GR_LOAD(float4, b, 2 * uint64_t(node));
        const float4 c = GR_LOAD(float4, b, 2 * uint64_t(node) + 1);
        x[0] = a.x; x[1] = a.y; x[2] = a.z;
        y[0] = a.w; y[1] = c.x; y[2] = c.y;
    }
};
template <> struct PairLoad<float, 2>
{
    static __device__ __forceinline__ void get(const void * b, uint32_t node, float (&x)[2], float (&y)[2])
    {
        const float4 a = GR_LOAD(float4, b, node);
        x[0] = a.x; x[1] = a.y;
        y[0] = a.z; y[1] = a.w;
    }
};
template <> struct PairLoad<__half, 3>
{
    static __device__ __forceinline__ void get(const void * b, uint32_t node, float (&x)[3], float (&y)[3])
    {
        const uint4 a = GR_LOAD(uint4, b, node);
        x[0] = h_lo(a.x); x[1] = h_hi(a.x); x[2] = h_lo(a.y);
        y[0] = h_hi(a.y); y[1] = h_lo(a.z); y[2] = h_hi(a.z);
    }
};
template <> struct PairLoad<__half, 2>
{
    static __device__ __forceinline__ void get(const void * b, uint32_t node, float (&x)[2], float (&y)[2])
    {
        const uint2 a = GR_LOAD(uint2, b, node);
        x[0] = h_lo(a.x); x[1] = h_hi(a.x);
        y[0] = h_lo(a.y); y[1] = h_hi(a.y);
    }
};
template <> struct PairLoad<bf16_t, 3>
{
    static __device__ __forceinline__ void get(const void * b, uint32_t node, float (&x)[3], float (&y)[3])
    {
        const uint4 a = GR_LOAD(uint4, b, node);
        x[0] = b_lo(a.x); x[1] = b_hi(a.x); x[2] = b_lo(a.y);
        y[0] = b_hi(a.y); y[1] = b_lo(a.z); y[2] = b_hi(a.z);
    }
};
template <> struct PairLoad<bf16_t, 2>
{
    static __device__ __forceinline__ void get(const void * b, uint32_t node, float (&x)[2], float (&y)[2])
    {
        const uint2 a = GR_LOAD(uint2, b, node);
        x[0] = b_lo(a.x); x[1] = b_hi(a.x);
        y[0] = b_lo(a.y); y[1] = b_hi(a.y);
    }
};

// kernel modes: lean kernels on the time-paired layout or on separate slots, and the full
// kernel (every reference feature, separate slots)
enum : int { M_PAIRED = 0, M_SEPARATE = 1, M_FULL = 2 };

// slices of the window being integrated (change at window edges)
struct Win
{
    const void * back;
    const void * front;
    const void * pair;
};

// Texel coordinate of height zr (relative to the origin) on a stretched z axis:
// k + (zr - z_k) / (z_{k+1} - z_k) with z_k <= zr < z_{k+1}; linear continuation of the first /
// last cell outside the levels. The lookup table gives a k at or below the answer, the loop
// finishes the search.
__device__ __forceinline__ float zcoord(const FieldView & f, float zr)
{
    const int nz = int(f.nz);
    const float u = (zr - f.zlut_z0) * f.zlut_scale;
    int k;
    if (!(u >= 0.f)) // below the first level (or NaN)
        k = 0;
    else if (u >= float(f.nlut))
        k = nz - 2;
    else
    {
        k = GR_LOAD(int32_t, f.zlut, int(u));
        while (k > 0 && zr < GR_LOAD(float, f.zlev, k)) --k; // guards the float rounding of u at bin edges
        while (k + 2 < nz && zr >= GR_LOAD(float, f.zlev, k + 1)) ++k;
    }
    const float z0 = GR_LOAD(float, f.zlev, k), z1 = GR_LOAD(float, f.zlev, k + 1);
    return float(k) + (zr - z0) / (z1 - z0);
}

template <int DIM>
__device__ __forceinline__ void texel_arg(const StepParams & P, const float * x, float * arg)
{
#pragma unroll
    for (int d = 0; d < DIM; ++d) arg[d] = (x[d] - P.off[d]) * P.inv_h[d];
    if (DIM == 3 && P.wind.zlev) arg[2] = zcoord(P.wind, x[2] - P.off[2]);
}

// The last FP16 time-paired stencil a lane loaded: its base node and the raw words of its
// corners (3D: words x, y, z of the eight 16-byte nodes, the fourth being padding; 2D: both words
// of the four 8-byte nodes). The RK4 stages of a step mostly sample the same cell (a particle
// crosses about a fifth of a cell per step), so with CACHE a lane whose base node is unchanged
// reuses the words instead of loading the stencil again: the same data, the same result. Holding
// the words costs about 50 VGPRs (two wavefronts per SIMD instead of three), so the kernels
// built with CACHE serve narrow launches, whose wavefronts all fit on the device at once and
// whose steps are bound by one wavefront's latency.
template <int DIM>
struct Stencil16
{
    uint32_t base;
    uint32_t w[DIM == 3 ? 8 : 4][DIM == 3 ? 3 : 2];
};

template <int DIM, class T, int MODE, bool CACHE>
__device__ __forceinline__ void velocity(const StepParams & P, const Win & W, const float * x,
                                         float t, float * v, f16_t one, Stencil16<DIM> & C)
{
    float arg[3] = {0.f, 0.f, 0.f};
    texel_arg<DIM>(P, x, arg);
    if (MODE == M_PAIRED)
    {
        // Padded paired copy: node (i, j, k), i in [-1, n + 1], at
        // ((k + 1) * PY + (j + 1)) * PX + (i + 1) with PX = nx + 3, PY = ny + 3; the halo holds
        // zeros (border addressing) or edge copies (clamp addressing), so the 8 corners are
        // fixed offsets from the base node and need no per-corner clamp or select.
        Taps S;
        int i, j, k = -1;
        split(arg[0], P.wind.nx, P.weights, i, S.wx);
        split(arg[1], P.wind.ny, P.weights, j, S.wy);
        S.wz = 0.f;
        if (DIM == 3) split(arg[2], P.wind.nz, P.weights, k, S.wz);
        const uint32_t PX = P.wind.nx + 3, PXY = PX * (P.wind.ny + 3);
        const uint32_t base = uint32_t(k + 1) * PXY + uint32_t(j + 1) * PX + uint32_t(i + 1);
        float b[DIM], f[DIM];
        if constexpr (std::is_same<T, __half>::value)
        {
            if (!CACHE || base != C.base) // divergent: lanes whose cell is unchanged load nothing
            {
#pragma unroll
                for (int c = 0; c < (DIM == 3 ? 8 : 4); ++c)
                {
                    const uint32_t node = base + ((c & 1) ? 1u : 0u) + ((c & 2) ? PX : 0u) + ((c & 4) ? PXY : 0u);
                    if (DIM == 3)
                    {
                        const uint4 h = GR_LOAD(uint4, W.pair, node);
                        C.w[c][0] = h.x;
                        C.w[c][1] = h.y;
                        C.w[c][2] = h.z;
                    }
                    else
                    {
                        const uint2 h = GR_LOAD(uint2, W.pair, node);
                        C.w[c][0] = h.x;
                        C.w[c][1] = h.y;
                    }
                }
                C.base = base;
            }
            uint4 raw[8];
#pragma unroll
            for (int c = 0; c < (DIM == 3 ? 8 : 4); ++c)
                raw[c] = DIM == 3 ? make_uint4(C.w[c][0], C.w[c][1], C.w[c][2], 0u)
                                  : make_uint4(C.w[c][0], C.w[c][1], 0u, 0u);
            combine_pair_h<DIM>(raw, S, one, b, f);
#pragma unroll
            for (int d = 0; d < DIM; ++d) v[d] = lerp(b[d], f[d], t);
            return;
        }
        float vb[8][DIM], vf[8][DIM];
#pragma unroll
        for (int c = 0; c < (DIM == 3 ? 8 : 4); ++c)
        {
            const uint32_t node = base + ((c & 1) ? 1u : 0u) + ((c & 2) ? PX : 0u) + ((c & 4) ? PXY : 0u);
            PairLoad<T, DIM>::get(W.pair, node, vb[c], vf[c]);
        }
        combine<DIM, DIM>(vb, S, b);
        combine<DIM, DIM>(vf, S, f);
#pragma unroll
        for (int d = 0; d < DIM; ++d) v[d] = lerp(b[d], f[d], t);
        return;
    }
    Taps S;
    taps<DIM>(arg, P.wind.nx, P.wind.ny, P.wind.nz, P.wind.address, P.weights, S);
    float b[DIM];
    sample_slot<DIM, T, DIM>(W.back, S, b);
    if (P.flags & F_WIND_DYNAMIC)
    {
        float f[DIM];
        sample_slot<DIM, T, DIM>(W.front, S, f);
#pragma unroll
        for (int d = 0; d < DIM; ++d) v[d] = lerp(b[d], f[d], t);
    }
    else
    {
#pragma unroll
        for (int d = 0; d < DIM; ++d) v[d] = b[d];
    }
}

// sphere_coordinates, TrajectoryIntegrator.tpp:98-104 (2D only)
__device__ __forceinline__ void sphere(const StepParams & P, float py, const float * v, float * o)
{
    const float PI = 3.14159265358979323846f;
    const float PI_2 = PI * 2.f;
    o[0] = (v[0] * 180.f) / (P.sphere_radius * cosf(py * PI_2 / 360.f) * PI);
    o[1] = (v[1] * 180.f) / (PI * P.sphere_radius);
}

// 2D loop borders (TrajectoryIntegrator.tpp:20-96); returns true on Domain abort.
__device__ bool loop_border(const StepParams & P, const float * o, float * n)
{
    const float ox = P.off[0], oy = P.off[1], DX = P.dom[0], DY = P.dom[1];
    const bool sph = (P.flags & F_SPHERE) != 0;
    if (P.border == GR_LOOP_X)
    {
        n[0] = fmodf((n[0] - ox + DX), DX) + ox;
        return n[1] - oy > DY || n[1] - oy < 0.f;
    }
    if (P.border == GR_LOOP_Y)
    {
        if (sph)
            return n[0] - ox > DX || n[0] - ox < 0.f || n[1] - oy > DY || n[1] - oy < 0.f;
        n[1] = fmodf((n[1] - oy + DY), DY) + oy;
        return n[0] - ox > DX || n[0] - ox < 0.f;
    }
    if (sph)
    {
        const float nx_ = n[0] - ox, ny_ = n[1] - oy;
        if (ny_ < 0.f)
        {
            const float ox_ = o[0] - ox, oy_ = o[1] - oy;
            const float tx = (-oy_) / (ny_ - oy_) * (nx_ - ox_) + ox_;
            n[1] = -(ny_) + oy;
            n[0] = fmodf(DX + 180.0f - nx_ + 2.f * tx, DX) + ox;
        }
        else if (ny_ >= DY)
        {
            const float ox_ = o[0] - ox, oy_ = o[1] - oy;
            const float tx = (DY - oy_) / (ny_ - oy_) * (nx_ - ox_) + ox_;
            n[1] = (ny_ - DY) + oy;
            n[0] = fmodf(DX + 180.0f - nx_ + 2.f * tx, DX) + ox;
        }
    }
    else
        n[1] = fmodf((n[1] - oy + DY), DY) + oy;
    n[0] = fmodf((n[0] - ox + DX), DX) + ox;
    return false;
}

__device__ __forceinline__ float global_coord(const StepParams & P, float x, int d)
{
    return (P.flags & F_LOCAL) ? float(P.off64[d] + double(x)) : x;
}

// Advances one particle by the launch's steps; returns the number of steps it integrated
// (steps begun alive, the aborting step included).
template <int DIM, int INTEG, class T, int MODE, bool CACHE = false>
__device__ __forceinline__ uint32_t advance_one(const StepParams & P, const uint32_t pid)
{
    constexpr bool FULL = MODE == M_FULL;
    const uint32_t oid = (FULL && P.orig) ? GR_LOAD(uint32_t, P.orig, pid) : pid; // caller-order index

    // window state; with nwin > 1 the launch crosses window edges (tpp:933-969 per window)
    Win W{P.wind.back, P.wind.front, P.wind.pair};
    uint32_t wi = 0;
    // the window tables are uniform and read-only: scalar loads keep their values in SGPRs
    uint32_t edge = P.nwin > 1 ? GR_CONST(uint32_t, P.win_edge, 0) : 0xffffffffu;

    // A stopped particle's state never changes again (tpp:452-455 copies it unchanged) and its
    // saved columns after the stop are left to the reader, which fills them from the final
    // state: nothing to do.
    int32_t status = G(P.status)[pid];
    if (status != GR_STATUS_TIME) return 0;
    f16_t one = f16_t(1.0f);
    asm volatile("" : "+v"(one)); // opaque 1.0 for the mixed-precision lerps (lerp_h)

    float p[3] = {0.f, 0.f, 0.f}, q[3] = {0.f, 0.f, 0.f}, cmp[3] = {0.f, 0.f, 0.f};
    float ld[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < DIM; ++d) p[d] = G(P.p[d])[pid];
    const bool curv_on = P.curvature != GR_CURV_OFF;
    const bool curv_pos = P.curvature == GR_CURV_TOTAL || P.curvature == GR_CURV_INDIVIDUAL_AND_TOTAL;
    if (curv_pos)
#pragma unroll
        for (int d = 0; d < DIM; ++d) q[d] = G(P.q[d])[pid];
    if (P.curvature == GR_CURV_FAST_TOTAL)
#pragma unroll
        for (int d = 0; d < DIM; ++d) ld[d] = G(P.ld[d])[pid];
    if (P.flags & F_KAHAN)
#pragma unroll
        for (int d = 0; d < DIM; ++d) cmp[d] = G(P.c[d])[pid];
    int32_t abort_step = G(P.abort_step)[pid];
    float total = curv_on ? G(P.total)[pid] : 0.f;
    float curv = P.curvature == GR_CURV_INDIVIDUAL_AND_TOTAL ? G(P.curv)[pid] : 0.f;
    float len = (FULL && P.abort_mode != GR_ABORT_TIME) ? G(P.len)[pid] : 0.f;
    float ex[GR_MAX_PROGRAMS];
#pragma unroll
    for (int e = 0; e < GR_MAX_PROGRAMS; ++e) ex[e] = e < int(P.nprog) ? G(P.expr[e])[pid] : 0.f;
    float vo[GR_MAX_VOLUMES];
#pragma unroll
    for (int k = 0; k < GR_MAX_VOLUMES; ++k) vo[k] = (FULL && k < int(P.nvol)) ? G(P.vol[k])[pid] : 0.f;

    const float dt = P.dt;
    const float half_dt = 0.5f * dt;
    uint64_t next_save = P.save_first;
    // ring column of the next save (global save index modulo the ring), advanced per save
    uint32_t col = P.ring_cap ? uint32_t(P.save_g0 % P.ring_cap) : 0u;

    uint32_t integrated = 0;
    // |p - q| of the next step is |np - p| of this one (same operands, same order): carried
    // over while consecutive steps both take the curvature branch
    bool have_sb = false;
    float sb_prev = 0.f;
    Stencil16<DIM> sc;
    sc.base = 0xffffffffu; // no stencil loaded (also after every window change)
    for (uint32_t s = 0; s < P.nsteps; ++s)
    {
        if (s == edge) // next window: slices c+1, c+2 (wave-uniform branch)
        {
            ++wi;
            typedef const void * cvp_t;
            W.back = GR_CONST(cvp_t, P.win_ptr, 3 * wi + 0);
            W.front = GR_CONST(cvp_t, P.win_ptr, 3 * wi + 1);
            W.pair = GR_CONST(cvp_t, P.win_ptr, 3 * wi + 2);
            edge = wi + 1 < P.nwin ? GR_CONST(uint32_t, P.win_edge, wi) : 0xffffffffu;
            sc.base = 0xffffffffu;
        }
        const bool live = status == GR_STATUS_TIME; // alive at the start of step s
        if (live)
        {
            ++integrated;
            // time weights of this step from the launch's table, which the host forms for every
            // launch with the reference's operations: unit_interval of Time, Time + dt/2 and
            // Time + dt in the step's window (tpp:132, 141-144), FP32 or FP64 time, and whether
            // Time > 0 (curvature, tpp:347)
            const float4 tt = GR_CONST(float4, P.ttab, s);
            const float t = tt.x;
            const bool sph = FULL && DIM == 2 && (P.flags & F_SPHERE);
            float v[3] = {0.f, 0.f, 0.f}, vc[3] = {0.f, 0.f, 0.f};
            if (INTEG == GR_RK4)
            {
                const float t1 = tt.y, t2 = tt.z;
                float k1[3] = {0.f, 0.f, 0.f}, k2[3] = {0.f, 0.f, 0.f}, k3[3] = {0.f, 0.f, 0.f}, k4[3] = {0.f, 0.f, 0.f};
                float x[3] = {0.f, 0.f, 0.f};
                velocity<DIM, T, MODE, CACHE>(P, W, p, t, k1, one, sc);
                float k1c[3] = {k1[0], k1[1], k1[2]};
                if (sph) sphere(P, p[1], k1c, k1);
#pragma unroll
                for (int d = 0; d < DIM; ++d) x[d] = p[d] + k1[d] * half_dt;
                velocity<DIM, T, MODE, CACHE>(P, W, x, t1, k2, one, sc);
                float k2c[3] = {k2[0], k2[1], k2[2]};
                if (sph) sphere(P, x[1], k2c, k2);
#pragma unroll
                for (int d = 0; d < DIM; ++d) x[d] = p[d] + k2[d] * half_dt;
                velocity<DIM, T, MODE, CACHE>(P, W, x, t1, k3, one, sc);
                float k3c[3] = {k3[0], k3[1], k3[2]};
                if (sph) sphere(P, x[1], k3c, k3);
#pragma unroll
                for (int d = 0; d < DIM; ++d) x[d] = p[d] + k3[d] * dt;
                velocity<DIM, T, MODE, CACHE>(P, W, x, t2, k4, one, sc);
                float k4c[3] = {k4[0], k4[1], k4[2]};
                if (sph) sphere(P, x[1], k4c, k4);
#pragma unroll
                for (int d = 0; d < DIM; ++d)
                {
                    v[d] = (((k1[d] + k2[d] * 2.f) + k3[d] * 2.f) + k4[d]) * (1.0f / 6);
                    vc[d] = sph ? (((k1c[d] + k2c[d] * 2.f) + k3c[d] * 2.f) + k4c[d]) * (1.0f / 6) : v[d];
                }
            }
            else
            {
                velocity<DIM, T, MODE, CACHE>(P, W, p, t, vc, one, sc);
#pragma unroll
                for (int d = 0; d < DIM; ++d) v[d] = vc[d];
                if (sph) sphere(P, p[1], vc, v);
            }

            // dp (tpp:233-250)
            float dp[3] = {0.f, 0.f, 0.f};
            const float sgn = (P.flags & F_REVERSE) ? -1.f : 1.f;
#pragma unroll
            for (int d = 0; d < DIM; ++d) dp[d] = v[d] * dt * sgn;
            if (FULL && DIM == 3 && (P.flags & (F_UPLIFT_CONST | F_UPLIFT_DYN)))
            {
#pragma unroll
                for (int d = 0; d < DIM; ++d)
                {
                    const float lb = G(P.lift_b)[uint64_t(oid) * DIM + d];
                    const float l = (P.flags & F_UPLIFT_DYN) ? lerp(lb, G(P.lift_f)[uint64_t(oid) * DIM + d], t) : lb;
                    dp[d] = dp[d] + l;
                }
            }

            int32_t abort = GR_STATUS_TIME;
            if (FULL && P.abort_mode != GR_ABORT_TIME) // tpp:253-280
            {
                float l2 = 0.f;
#pragma unroll
                for (int d = 0; d < DIM; ++d) l2 += dp[d] * dp[d];
                const float dp_l = sqrtf(l2);
                const float c_l = len;
                const float n_l = c_l + dp_l;
                if (P.abort_mode == GR_ABORT_FIT_LENGTH)
                {
                    if (n_l > P.max_length)
                    {
                        const float sc = (P.max_length - c_l) / dp_l;
#pragma unroll
                        for (int d = 0; d < DIM; ++d) dp[d] = sc * dp[d];
                        len = P.max_length;
                        abort = GR_STATUS_LENGTH;
                    }
                    else
                        len = n_l;
                }
                else
                {
                    if (n_l > P.max_length) abort = GR_STATUS_LENGTH;
                    len = n_l;
                }
            }

            float np[3] = {p[0], p[1], p[2]};
            float y[3] = {0.f, 0.f, 0.f};
            if (P.flags & F_KAHAN)
            {
#pragma unroll
                for (int d = 0; d < DIM; ++d)
                {
                    y[d] = ((P.flags & F_REVERSE) ? -dp[d] : dp[d]) - cmp[d];
                    np[d] = p[d] + y[d];
                }
            }
            else
            {
#pragma unroll
                for (int d = 0; d < DIM; ++d) np[d] = (P.flags & F_REVERSE) ? p[d] - dp[d] : p[d] + dp[d];
            }

            float dp2 = 0.f;
#pragma unroll
            for (int d = 0; d < DIM; ++d) dp2 += dp[d] * dp[d];
            if (abort == GR_STATUS_TIME) // tpp:284-323
            {
                if (FULL && DIM == 2 && P.border != GR_BLOCK)
                {
                    if (dp2 < GR_ABORT_VALUE)
                        abort = GR_STATUS_WIND;
                    else if (loop_border(P, p, np))
                        abort = GR_STATUS_DOMAIN;
                }
                else
                {
                    bool topo = false;
                    if (FULL && DIM == 3 && (P.flags & F_TOPO))
                    {
                        float ta[3] = {global_coord(P, np[0], 0) / P.h[0],
                                       global_coord(P, np[1], 1) / P.h[1], 0.f};
                        topo = sample_scalar<2>(P.topo.back, P.topo, 0, ta, P.weights) >
                               global_coord(P, np[2], 2);
                    }
                    if (topo)
                        abort = GR_STATUS_TOPOGRAPHY;
                    else if (dp2 < GR_ABORT_VALUE)
                        abort = GR_STATUS_WIND;
                    else
                    {
                        bool out = false;
#pragma unroll
                        for (int d = 0; d < DIM; ++d)
                            out = out || (np[d] - P.off[d] > P.dom[d] || np[d] - P.off[d] < 0.f);
                        if (out) abort = GR_STATUS_DOMAIN;
                    }
                }
            }
            status = abort;
            if (abort != GR_STATUS_TIME) abort_step = int32_t(P.counter_start + s);

            if (abort == GR_STATUS_WIND || abort == GR_STATUS_TOPOGRAPHY || abort == GR_STATUS_DOMAIN)
            {
                if (curv_on) total = -1.f; // tpp:327-332
            }
            else
            {
                const bool tpos = tt.w != 0.f;
                if (P.curvature == GR_CURV_FAST_TOTAL) // tpp:337-346
                {
                    float l2 = 0.f;
#pragma unroll
                    for (int d = 0; d < DIM; ++d) l2 += v[d] * v[d];
                    const float inv = 1.f / sqrtf(l2);
                    float vn[3] = {0.f, 0.f, 0.f};
#pragma unroll
                    for (int d = 0; d < DIM; ++d) vn[d] = v[d] * inv;
                    if (tpos)
                    {
                        float dd = 0.f;
#pragma unroll
                        for (int d = 0; d < DIM; ++d) dd += ld[d] * vn[d];
                        const float k = acosf(dd);
                        if (!isnan(k)) total += k;
                    }
#pragma unroll
                    for (int d = 0; d < 3; ++d) ld[d] = vn[d];
                }
                else if (curv_pos && tpos) // tpp:347-360, kappa icc:145-159
                {
                    float k = 0.f;
                    if (DIM == 3)
                    {
                        float d1[3], d2[3];
#pragma unroll
                        for (int d = 0; d < 3; ++d)
                        {
                            d1[d] = (np[d] - q[d]) / (2.f * dt);
                            d2[d] = ((q[d] + np[d]) - 2.f * p[d]) / (dt * dt);
                        }
                        const float ldp = sqrtf((d1[0] * d1[0] + d1[1] * d1[1]) + d1[2] * d1[2]);
                        const float ldp3 = (ldp * ldp) * ldp;
                        const float cx = d1[1] * d2[2] - d1[2] * d2[1];
                        const float cy = d1[2] * d2[0] - d1[0] * d2[2];
                        const float cz = d1[0] * d2[1] - d1[1] * d2[0];
                        k = fabsf(sqrtf((cx * cx + cy * cy) + cz * cz) / ldp3);
                    }
                    float sa = sb_prev, b2 = 0.f;
                    if (!have_sb)
                    {
                        float a2 = 0.f;
#pragma unroll
                        for (int d = 0; d < DIM; ++d) a2 += (p[d] - q[d]) * (p[d] - q[d]);
                        sa = sqrtf(a2);
                    }
#pragma unroll
                    for (int d = 0; d < DIM; ++d) b2 += (np[d] - p[d]) * (np[d] - p[d]);
                    const float sb = sqrtf(b2);
                    const float ds = 0.5f * (sa + sb);
                    sb_prev = sb;
                    have_sb = true;
                    if (P.curvature == GR_CURV_INDIVIDUAL_AND_TOTAL) curv = k;
                    if (!isnan(k)) total += k * ds;
                }

                if (FULL && (P.flags & (F_VOL_CONST | F_VOL_DYN))) // tpp:363-377
                {
                    float arg[3] = {0.f, 0.f, 0.f};
                    texel_arg<DIM>(P, np, arg);
                    for (uint32_t k = 0; k < P.nvol; ++k)
                    {
                        float val = sample_scalar<DIM>(P.vols.back, P.vols, k, arg, P.weights);
                        if (P.flags & F_VOL_DYN)
                            val = lerp(val, sample_scalar<DIM>(P.vols.front, P.vols, k, arg, P.weights), t);
                        vo[k] = val;
                    }
                }

                if (P.nprog > 0) // tpp:379-447
                {
                    // reference vector slots (tpp:381-439): position, curvature and velocity from
                    // registers; volumes and constants (full kernel) staged in LDS, only those read
                    lds_float * s_refs = expr_lds();
                    lds_float * s_stk = s_refs + 19 * 256;
                    const uint32_t lane = threadIdx.x, bs = blockDim.x, m = P.refmask;
                    lds_float * R = s_refs + lane;
                    if (FULL)
                    {
#pragma unroll
                        for (int k = 0; k < GR_MAX_VOLUMES; ++k)
                            if (m & (1u << (4 + k))) R[(4 + k) * bs] = k < int(P.nvol) ? vo[k] : 0.f;
#pragma unroll
                        for (int k = 0; k < GR_MAX_CONSTANTS; ++k)
                        {
                            if (!(m & (1u << (12 + k)))) continue;
                            float cv = 0.f;
                            if (k < int(P.nconst))
                            {
                                cv = G(P.cst_b[k])[oid];
                                if (P.flags & F_CONST_DYN) cv = lerp(cv, G(P.cst_f[k])[oid], t);
                            }
                            R[(12 + k) * bs] = cv;
                        }
                    }
                    float gx = 0.f, gy = 0.f, gz = 0.f;
                    if (m & 7u)
                    {
                        gx = global_coord(P, np[0], 0);
                        gy = global_coord(P, np[1], 1);
                        if (DIM == 3) gz = global_coord(P, np[2], 2);
                    }
                    const refvec_t rv = {gx, gy, gz, P.curvature == GR_CURV_INDIVIDUAL_AND_TOTAL ? curv : 0.f,
                                         vc[0], vc[1], DIM == 3 ? vc[2] : 0.f, 0.f, 0.f};
                    for (uint32_t e = 0; e < P.nprog; ++e)
                        ex[e] = run_program<FULL>(P.prog[e], rv, R, s_stk, lane, bs);
                }

                if (P.flags & F_KAHAN)
                {
#pragma unroll
                    for (int d = 0; d < DIM; ++d) cmp[d] = (np[d] - p[d]) - y[d];
                }
#pragma unroll
                for (int d = 0; d < DIM; ++d)
                {
                    q[d] = p[d];
                    p[d] = np[d];
                }
            }
        }

        // copyParticleInfo, tpp:997-1049. Columns after a particle's stop step hold its final
        // state and are filled when read (k_fill_frozen).
        if (s == next_save && live)
        {
            auto rp = G(P.ring_pos) + (uint64_t(col) * P.n + pid) * DIM;
#pragma unroll
            for (int d = 0; d < DIM; ++d) rp[d] = global_coord(P, p[d], d);
            if (P.curvature == GR_CURV_INDIVIDUAL_AND_TOTAL) G(P.ring_curv)[uint64_t(col) * P.n + pid] = curv;
            for (uint32_t e = 0; e < P.nprog; ++e)
                G(P.ring_expr)[(uint64_t(e) * P.ring_cap + col) * P.n + pid] = ex[e];
            if (FULL)
                for (uint32_t k = 0; k < P.nvol; ++k)
                    G(P.ring_vol)[(uint64_t(k) * P.ring_cap + col) * P.n + pid] = vo[k];
        }
        if (s == next_save)
        {
            col = col + 1 == P.ring_cap ? 0u : col + 1;
            next_save += P.save_interval;
        }
        if (status != GR_STATUS_TIME) break; // stopped: nothing more to compute or save
    }

#pragma unroll
    for (int d = 0; d < DIM; ++d) G(P.p[d])[pid] = p[d];
    if (curv_pos)
#pragma unroll
        for (int d = 0; d < DIM; ++d) G(P.q[d])[pid] = q[d];
    if (P.curvature == GR_CURV_FAST_TOTAL)
#pragma unroll
        for (int d = 0; d < DIM; ++d) G(P.ld[d])[pid] = ld[d];
    if (P.flags & F_KAHAN)
#pragma unroll
        for (int d = 0; d < DIM; ++d) G(P.c[d])[pid] = cmp[d];
    G(P.status)[pid] = status;
    G(P.abort_step)[pid] = abort_step;
    if (curv_on) G(P.total)[pid] = total;
    if (P.curvature == GR_CURV_INDIVIDUAL_AND_TOTAL) G(P.curv)[pid] = curv;
    if (FULL && P.abort_mode != GR_ABORT_TIME) G(P.len)[pid] = len;
    for (uint32_t e = 0; e < P.nprog; ++e) G(P.expr[e])[pid] = ex[e];
    if (FULL)
        for (uint32_t k = 0; k < P.nvol; ++k) G(P.vol[k])[pid] = vo[k];
    return integrated;
}

// Adds the wavefront's integrated particle-steps to *work (all 64 lanes must be active). The
// sum is formed from 16-bit halves so that no 32-bit partial sum can overflow.
__device__ __forceinline__ void wave_add_work(unsigned long long * work, uint32_t v)
{
    uint32_t lo = v & 0xffffu, hi = v >> 16;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
    {
        lo += __shfl_xor(lo, o, 64);
        hi += __shfl_xor(hi, o, 64);
    }
    if ((threadIdx.x & 63u) == 0 && (lo | hi))
        atomicAdd(work, (unsigned long long)lo + ((unsigned long long)hi << 16));
}

// Workgroups are handed to the XCDs round-robin (workgroup b runs on XCD b % nxcd). Renumbering
// them so that each XCD gets one contiguous run of workgroups keeps the particles in flight on
// an XCD (Morton-ordered) in one compact region, whose field tiles then stay in that XCD's L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nblocks, uint32_t nxcd)
{
    if (nxcd <= 1) return b;
    const uint32_t x = b % nxcd, local = b / nxcd;
    const uint32_t per = nblocks / nxcd, rem = nblocks % nxcd;
    // XCD x owns per + (x < rem) consecutive logical blocks
    return x < rem ? x * (per + 1) + local : rem * (per + 1) + (x - rem) * per + local;
}

template <int DIM, int INTEG, class T, int MODE, bool CACHE = false>
__device__ __forceinline__ void advance_block(const StepParams & P, uint32_t nxcd)
{
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x, nxcd);
    const uint32_t pid = b * blockDim.x + threadIdx.x;
    uint32_t w = 0;
    if (pid < P.n) w = advance_one<DIM, INTEG, T, MODE, CACHE>(P, pid);
    if (P.work) // block sizes are multiples of 64
        wave_add_work(P.work + size_t(blockIdx.x % WORK_SLOTS) * WORK_STRIDE, w);
}

template <int DIM, int INTEG, class T, int MODE>
__global__ void __launch_bounds__(256) k_advance(const StepParams P, uint32_t nxcd)
{
    advance_block<DIM, INTEG, T, MODE>(P, nxcd);
}

// The same kernel held to at most 128 VGPRs, i.e. four wavefronts per SIMD instead of three.
// Each wavefront's step chain runs a little slower, so this pays only when the launch has many
// more wavefronts than the compute units hold at once. Instantiated for the FP16 time-paired
// form only: it is the variant that fits this budget without spilling registers.
template <int DIM, int INTEG, class T, int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
k_advance_w4(const StepParams P, uint32_t nxcd)
{
    advance_block<DIM, INTEG, T, MODE>(P, nxcd);
}

// The FP16 time-paired kernel with the stencil reuse (CACHE) for narrow launches.
template <int DIM, int INTEG, class T, int MODE>
__global__ void __launch_bounds__(256) k_advance_narrow(const StepParams P, uint32_t nxcd)
{
    advance_block<DIM, INTEG, T, MODE, true>(P, nxcd);
}

// Persistent batched kernel: workgroups pull (job, particle block) items from a global counter
// until the list is exhausted; each item runs its job's whole span of steps. Jobs of different
// dimension, integrator and storage type share the launch; the variant is uniform per item.
template <int DIM, int INTEG, class T, bool CACHE>
__device__ __forceinline__ uint32_t run_item(const StepParams & P, uint32_t pid)
{
    // the host builds a paired copy only for jobs that use it (dynamic wind, paired option)
    if (P.wind.pair) return advance_one<DIM, INTEG, T, M_PAIRED, CACHE>(P, pid);
    return advance_one<DIM, INTEG, T, M_SEPARATE>(P, pid);
}

// Identity of the compute unit running the calling wavefront: XCC, shader engine, shader array
// and CU fields of the HW_ID / XCC_ID hardware registers packed into CU_KEY_BITS bits.
__device__ __forceinline__ uint32_t cu_key()
{
    const uint32_t hw = __builtin_amdgcn_s_getreg((31 << 11) | 4); // HW_REG_HW_ID, bits 31:0
    const uint32_t xcc = __builtin_amdgcn_s_getreg((3 << 11) | 20); // HW_REG_XCC_ID, bits 3:0
    return ((xcc & 7u) << 8) | (((hw >> 13) & 7u) << 5) | (((hw >> 12) & 1u) << 4) | ((hw >> 8) & 15u);
}

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

// Items are one wavefront of particles and every wavefront takes its own: a wavefront whose
// particles are all dead moves on at once instead of waiting at a workgroup barrier. There is
// one queue per group of `group` compute units (queue q = items [qbegin[q], qbegin[q + 1]), a
// compact region of the Morton-ordered particle sets, the same region for every job); the
// wavefronts of a group drain its queue first, so the group's caches hold the field tiles of
// one region (one group per XCD: its L2; one per CU: its L1). A drained queue
// clears its bit in masks[q / 32]; wavefronts then move to the nearest queue whose bit is still
// set, and exit when no bit is left. Every wavefront accumulates its integrated particle-steps
// and adds them once, into one of WORK_SLOTS counters. WAVES is the minimum number of wavefronts
// per SIMD the register allocation must allow (1: no limit; 4: at most 128 VGPRs).
template <int WAVES, bool CACHE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WAVES)))
k_batch(const StepParams * __restrict__ jobs, const BatchItem * __restrict__ items,
        uint32_t * qdata, uint32_t nq, const uint32_t * __restrict__ culut, uint32_t group,
        unsigned long long * work)
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
            const uint32_t pid = b.first + lane;
            const StepParams & P = jobs[b.job];
            if (lane < b.count)
            {
                uint32_t w;
                switch (b.variant)
                {
                    case 0: w = run_item<2, GR_EULER, float, CACHE>(P, pid); break;
                    case 1: w = run_item<2, GR_EULER, __half, CACHE>(P, pid); break;
                    case 2: w = run_item<2, GR_EULER, bf16_t, CACHE>(P, pid); break;
                    case 3: w = run_item<2, GR_RK4, float, CACHE>(P, pid); break;
                    case 4: w = run_item<2, GR_RK4, __half, CACHE>(P, pid); break;
                    case 5: w = run_item<2, GR_RK4, bf16_t, CACHE>(P, pid); break;
                    case 6: w = run_item<3, GR_EULER, float, CACHE>(P, pid); break;
                    case 7: w = run_item<3, GR_EULER, __half, CACHE>(P, pid); break;
                    case 8: w = run_item<3, GR_EULER, bf16_t, CACHE>(P, pid); break;
                    case 9: w = run_item<3, GR_RK4, float, CACHE>(P, pid); break;
                    case 10: w = run_item<3, GR_RK4, __half, CACHE>(P, pid); break;
                    default: w = run_item<3, GR_RK4, bf16_t, CACHE>(P, pid); break;
                }
                acc += w;
            }
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

int batch_variant(uint32_t dim, uint32_t integrator, uint32_t dtype)
{
    return int((dim == 3 ? 6 : 0) + (integrator == GR_RK4 ? 3 : 0) + (dtype > 2 ? 0 : dtype));
}

int launch_batch(const StepParams * jobs, const BatchItem * items, uint32_t * qdata, uint32_t nq,
                 const uint32_t * culut, uint32_t group, unsigned long long * work, uint32_t grid,
                 uint32_t block, uint32_t waves, bool narrow, void * stream)
{
    if (nq == 0 || grid == 0 || group == 0) return 0;
    if (narrow)
        hipLaunchKernelGGL((k_batch<1, true>), dim3(grid), dim3(block), 0, hipStream_t(stream), jobs, items,
                           qdata, nq, culut, group, work);
    else if (waves >= 4)
        hipLaunchKernelGGL((k_batch<4, false>), dim3(grid), dim3(block), 0, hipStream_t(stream), jobs, items,
                           qdata, nq, culut, group, work);
    else
        hipLaunchKernelGGL((k_batch<1, false>), dim3(grid), dim3(block), 0, hipStream_t(stream), jobs, items,
                           qdata, nq, culut, group, work);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int batch_occupancy(uint32_t block, uint32_t waves, int * blocks_per_cu)
{
    const hipError_t e =
        waves >= 4 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_batch<4, false>, int(block), 0)
                   : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_batch<1, false>, int(block), 0);
    return e == hipSuccess ? 0 : -1;
}

// --------------------------------------------------------------------------------------------
// support kernels
// --------------------------------------------------------------------------------------------

// FP32 component-major staging -> node-interleaved storage of the slot
template <class T>
__global__ void k_pack(const float * __restrict__ src, T * __restrict__ dst, uint64_t nodes,
                       uint32_t ncomp, uint32_t cstride)
{
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nodes) return;
    for (uint32_t c = 0; c < cstride; ++c)
    {
        const float v = c < ncomp ? src[uint64_t(c) * nodes + i] : 0.f;
        if constexpr (sizeof(T) == 4)
            reinterpret_cast<float *>(dst)[i * cstride + c] = v;
        else if constexpr (std::is_same<T, __half>::value)
            dst[i * cstride + c] = __float2half_rn(v);
        else
        {
            // round to nearest even to bfloat16
            uint32_t u = __float_as_uint(v);
            uint16_t b;
            if ((u & 0x7fffffffu) > 0x7f800000u)
                b = uint16_t((u >> 16) | 0x0040u);
            else
                b = uint16_t((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
            reinterpret_cast<uint16_t *>(dst)[i * cstride + c] = b;
        }
    }
}

__global__ void k_seed(const float * __restrict__ aos, float * x, float * y, float * z, uint64_t n,
                       uint32_t dim, double o0, double o1, double o2, uint32_t local)
{
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float a = aos[i * dim + 0], b = aos[i * dim + 1];
    x[i] = local ? float(double(a) - o0) : a;
    y[i] = local ? float(double(b) - o1) : b;
    if (dim == 3)
    {
        const float c = aos[i * dim + 2];
        z[i] = local ? float(double(c) - o2) : c;
    }
}

__global__ void k_gather(const float * x, const float * y, const float * z, float * aos, uint64_t n,
                         uint32_t dim, double o0, double o1, double o2, uint32_t local)
{
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    aos[i * dim + 0] = local ? float(o0 + double(x[i])) : x[i];
    aos[i * dim + 1] = local ? float(o1 + double(y[i])) : y[i];
    if (dim == 3) aos[i * dim + 2] = local ? float(o2 + double(z[i])) : z[i];
}

__global__ void k_count_alive(const int32_t * __restrict__ status, uint64_t n,
                              unsigned long long * out)
{
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool alive = i < n && status[i] == GR_STATUS_TIME;
    const unsigned long long m = __ballot(alive);
    if ((threadIdx.x & (warpSize - 1)) == 0 && m) atomicAdd(out, (unsigned long long)__popcll(m));
}

// Completes saved columns: a particle that stopped before column c's step keeps its final
// state, which the step kernels do not write into the ring. ring: `count` columns starting at
// global column `first` of a ring of `cap` columns of n x w floats; src[d]: final values (w
// arrays); column c was saved at step c * save_interval.
__global__ void k_fill_frozen(float * __restrict__ ring, uint32_t cap, uint64_t first, uint32_t count,
                              uint64_t n, uint32_t w, const int32_t * __restrict__ status,
                              const int32_t * __restrict__ abort_step, const float * src0,
                              const float * src1, const float * src2, uint64_t save_interval,
                              uint32_t local, double o0, double o1, double o2)
{
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || status[i] == GR_STATUS_TIME) return;
    const int64_t stop = abort_step[i];
    const float * src[3] = {src0, src1, src2};
    const double off[3] = {o0, o1, o2};
    float v[3];
    for (uint32_t d = 0; d < w; ++d)
        v[d] = local ? float(off[d] + double(src[d][i])) : src[d][i];
    for (uint32_t c = 0; c < count; ++c)
    {
        const uint64_t g = first + c;
        if (int64_t(g * save_interval) <= stop) continue; // written by the step kernel
        float * col = ring + (g % cap) * n * w;
        for (uint32_t d = 0; d < w; ++d) col[i * w + d] = v[d];
    }
}

// --------------------------------------------------------------------------------------------
// launchers
// --------------------------------------------------------------------------------------------
typedef void (*AdvanceFn)(const StepParams, uint32_t);

template <int DIM, int INTEG, class T>
static AdvanceFn pick_mode(uint32_t mode)
{
    if (mode == M_FULL) return k_advance<DIM, INTEG, T, M_FULL>;
    if (mode == M_PAIRED) return k_advance<DIM, INTEG, T, M_PAIRED>;
    return k_advance<DIM, INTEG, T, M_SEPARATE>;
}

template <int DIM, int INTEG>
static AdvanceFn pick_type(uint32_t dtype, uint32_t mode)
{
    if (dtype == GR_F16) return pick_mode<DIM, INTEG, __half>(mode);
    if (dtype == GR_BF16) return pick_mode<DIM, INTEG, bf16_t>(mode);
    return pick_mode<DIM, INTEG, float>(mode);
}

int launch_advance(const StepParams & P, uint32_t dim, uint32_t integrator, uint32_t mode,
                   uint32_t block, uint32_t nxcd, uint64_t nrun, bool wide, bool narrow, void * stream)
{
    AdvanceFn fn;
    if (wide && mode == M_PAIRED && P.wind.dtype == GR_F16)
    {
        if (dim == 3)
            fn = integrator == GR_RK4 ? k_advance_w4<3, GR_RK4, __half, M_PAIRED>
                                      : k_advance_w4<3, GR_EULER, __half, M_PAIRED>;
        else
            fn = integrator == GR_RK4 ? k_advance_w4<2, GR_RK4, __half, M_PAIRED>
                                      : k_advance_w4<2, GR_EULER, __half, M_PAIRED>;
    }
    else if (narrow && mode == M_PAIRED && P.wind.dtype == GR_F16)
    {
        if (dim == 3)
            fn = integrator == GR_RK4 ? k_advance_narrow<3, GR_RK4, __half, M_PAIRED>
                                      : k_advance_narrow<3, GR_EULER, __half, M_PAIRED>;
        else
            fn = integrator == GR_RK4 ? k_advance_narrow<2, GR_RK4, __half, M_PAIRED>
                                      : k_advance_narrow<2, GR_EULER, __half, M_PAIRED>;
    }
    else if (dim == 3)
        fn = integrator == GR_RK4 ? pick_type<3, GR_RK4>(P.wind.dtype, mode)
                                  : pick_type<3, GR_EULER>(P.wind.dtype, mode);
    else
        fn = integrator == GR_RK4 ? pick_type<2, GR_RK4>(P.wind.dtype, mode)
                                  : pick_type<2, GR_EULER>(P.wind.dtype, mode);
    if (block == 0 || block > 256) block = 256;
    const uint32_t grid = uint32_t((std::min<uint64_t>(nrun, P.n) + block - 1) / block);
    if (grid == 0) return 0;
    hipLaunchKernelGGL(fn, dim3(grid), dim3(block), 0, hipStream_t(stream), P, nxcd);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_fill_frozen(float * ring, uint32_t cap, uint64_t first, uint32_t count, uint64_t n,
                       uint32_t w, const int32_t * status, const int32_t * abort_step,
                       const float * const * src, uint64_t save_interval, uint32_t local,
                       const double * off64, void * stream)
{
    if (n == 0 || count == 0) return 0;
    const uint32_t block = 256;
    const uint32_t grid = uint32_t((n + block - 1) / block);
    hipLaunchKernelGGL(k_fill_frozen, dim3(grid), dim3(block), 0, hipStream_t(stream), ring, cap,
                       first, count, n, w, status, abort_step, src[0], w > 1 ? src[1] : src[0],
                       w > 2 ? src[2] : src[0], save_interval, local, off64[0], off64[1], off64[2]);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_pack(const float * src, void * dst, uint64_t nodes, uint32_t ncomp, uint32_t cstride,
                uint32_t dtype, void * stream)
{
    const uint32_t block = 256;
    const uint32_t grid = uint32_t((nodes + block - 1) / block);
    if (grid == 0) return 0;
    if (dtype == GR_F16)
        hipLaunchKernelGGL(k_pack<__half>, dim3(grid), dim3(block), 0, hipStream_t(stream), src,
                           reinterpret_cast<__half *>(dst), nodes, ncomp, cstride);
    else if (dtype == GR_BF16)
        hipLaunchKernelGGL(k_pack<bf16_t>, dim3(grid), dim3(block), 0, hipStream_t(stream), src,
                           reinterpret_cast<bf16_t *>(dst), nodes, ncomp, cstride);
    else
        hipLaunchKernelGGL(k_pack<float>, dim3(grid), dim3(block), 0, hipStream_t(stream), src,
                           reinterpret_cast<float *>(dst), nodes, ncomp, cstride);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_seed(const float * aos, float * const * soa, uint64_t n, uint32_t dim,
                const double * off64, uint32_t local, void * stream)
{
    const uint32_t block = 256;
    const uint32_t grid = uint32_t((n + block - 1) / block);
    if (grid == 0) return 0;
    hipLaunchKernelGGL(k_seed, dim3(grid), dim3(block), 0, hipStream_t(stream), aos, soa[0],
                       soa[1], soa[2], n, dim, off64[0], off64[1], off64[2], local);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_gather_positions(float * const * soa, float * aos, uint64_t n, uint32_t dim,
                            const double * off64, uint32_t local, void * stream)
{
    const uint32_t block = 256;
    const uint32_t grid = uint32_t((n + block - 1) / block);
    if (grid == 0) return 0;
    hipLaunchKernelGGL(k_gather, dim3(grid), dim3(block), 0, hipStream_t(stream), soa[0], soa[1],
                       soa[2], aos, n, dim, off64[0], off64[1], off64[2], local);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_count_alive(const int32_t * status, uint64_t n, unsigned long long * out, void * stream)
{
    const uint32_t block = 256;
    const uint32_t grid = uint32_t((n + block - 1) / block);
    if (grid == 0) return 0;
    hipLaunchKernelGGL(k_count_alive, dim3(grid), dim3(block), 0, hipStream_t(stream), status, n,
                       out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// --------------------------------------------------------------------------------------------
// locality: time-paired slices and Morton reordering of particles
// --------------------------------------------------------------------------------------------

// Padded paired copy. Output node (pi, pj, pk) of the (nx + 3) x (ny + 3) x (nz + 3) grid
// (2D: one plane) stands for grid node (pi - 1, pj - 1, pk - 1) and holds
// [back comps, front comps, zero pad] with `so` elements; outside the grid it holds zeros
// (border addressing) or the nearest edge node (clamp addressing). W is the raw element word
// (uint32_t for FP32, uint16_t for FP16 / BF16): bits are copied, not converted.
template <class W>
__global__ void k_pair(const W * __restrict__ back, const W * __restrict__ front, W * __restrict__ out,
                       uint32_t nx, uint32_t ny, uint32_t nz, uint32_t dim, uint32_t address,
                       uint32_t ncomp, uint32_t cstride, uint32_t so)
{
    const uint32_t PX = nx + 3, PY = ny + 3, PZ = dim == 3 ? nz + 3 : 1;
    const uint64_t o = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (o >= uint64_t(PX) * PY * PZ) return;
    const int pi = int(o % PX), pj = int((o / PX) % PY), pk = int(o / (uint64_t(PX) * PY));
    int i = pi - 1, j = pj - 1, k = dim == 3 ? pk - 1 : 0;
    const bool inside = i >= 0 && i < int(nx) && j >= 0 && j < int(ny) && k >= 0 && k < int(nz);
    if (address == 1)
    {
        i = min(max(i, 0), int(nx) - 1);
        j = min(max(j, 0), int(ny) - 1);
        k = min(max(k, 0), int(nz) - 1);
    }
    const bool copy = inside || address == 1;
    const uint64_t src = copy ? (uint64_t(k) * ny + uint64_t(j)) * nx + uint64_t(i) : 0;
    for (uint32_t c = 0; c < so; ++c)
    {
        W v = 0;
        if (copy && c < ncomp)
            v = back[src * cstride + c];
        else if (copy && c < 2 * ncomp)
            v = front[src * cstride + (c - ncomp)];
        out[o * so + c] = v;
    }
}

uint64_t pair_nodes(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t dim)
{
    return uint64_t(nx + 3) * (ny + 3) * (dim == 3 ? nz + 3 : 1);
}

int launch_pair(const void * back, const void * front, void * out, uint32_t nx, uint32_t ny,
                uint32_t nz, uint32_t dim, uint32_t address, uint32_t ncomp, uint32_t cstride,
                uint32_t dtype, void * stream)
{
    const uint64_t nodes = pair_nodes(nx, ny, nz, dim);
    const uint32_t block = 256;
    const uint32_t grid = uint32_t((nodes + block - 1) / block);
    if (grid == 0) return 0;
    const uint32_t so = ncomp == 3 ? 8 : 4;
    if (dtype == GR_F32)
        hipLaunchKernelGGL(k_pair<uint32_t>, dim3(grid), dim3(block), 0, hipStream_t(stream),
                           static_cast<const uint32_t *>(back), static_cast<const uint32_t *>(front),
                           static_cast<uint32_t *>(out), nx, ny, nz, dim, address, ncomp, cstride, so);
    else
        hipLaunchKernelGGL(k_pair<uint16_t>, dim3(grid), dim3(block), 0, hipStream_t(stream),
                           static_cast<const uint16_t *>(back), static_cast<const uint16_t *>(front),
                           static_cast<uint16_t *>(out), nx, ny, nz, dim, address, ncomp, cstride, so);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

__device__ __forceinline__ uint32_t spread3(uint32_t x) // 10 bits -> every third bit
{
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

__device__ __forceinline__ uint32_t spread2(uint32_t x) // 16 bits -> every second bit
{
    x &= 0xffffu;
    x = (x | (x << 8)) & 0x00ff00ffu;
    x = (x | (x << 4)) & 0x0f0f0f0fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

// Bucket of particle g of a multi-job reorder: its job, then the Morton code of its current cell
// with the low `shift` bits dropped; a job's stopped particles go to its last bucket. Counts the
// bucket's particles.
__global__ void k_bucket_keys(const SortJob * __restrict__ jobs, uint32_t njobs, uint64_t total,
                              uint32_t * __restrict__ keys, uint32_t * __restrict__ vals,
                              uint32_t * __restrict__ counts, uint32_t shift, uint32_t nb)
{
    const uint64_t g = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= total) return;
    uint32_t lo = 0, hi = njobs; // last job with base <= g
    while (hi - lo > 1)
    {
        const uint32_t mid = (lo + hi) / 2;
        if (jobs[mid].base <= g) lo = mid;
        else hi = mid;
    }
    const SortJob & J = jobs[lo];
    const uint32_t i = uint32_t(g - J.base);
    uint32_t bucket = nb - 1;
    if (J.status[i] == GR_STATUS_TIME)
    {
        const uint32_t lim = J.dim == 3 ? 1023u : 65535u;
        uint32_t c[3] = {0, 0, 0};
        for (uint32_t d = 0; d < J.dim; ++d)
        {
            float a = (J.p[d][i] - J.off[d]) * J.inv_h[d];
            a = fminf(fmaxf(a, 0.f), float(J.grid[d] - 1));
            c[d] = min(uint32_t(a), lim);
        }
        const uint32_t code = J.dim == 3 ? (spread3(c[0]) | (spread3(c[1]) << 1) | (spread3(c[2]) << 2))
                                         : (spread2(c[0]) | (spread2(c[1]) << 1));
        bucket = min(code >> shift, nb - 2);
    }
    const uint32_t key = lo * nb + bucket;
    keys[g] = key;
    vals[g] = i;
    // one atomic per distinct key of the wavefront (neighbouring particles mostly share a bucket:
    // they were regrouped by the previous reorder)
    for (;;)
    {
        const uint32_t k0 = __builtin_amdgcn_readfirstlane(key);
        const uint64_t m = __ballot(key == k0);
        if (key == k0)
        {
            if ((m & ((1ull << __lane_id()) - 1)) == 0) atomicAdd(&counts[k0], uint32_t(__popcll(m)));
            break;
        }
    }
}

// alive count of each job: the start of its stopped bucket minus its base (bucket starts from the
// exclusive scan of the counts, before the scatter advances them)
__global__ void k_bucket_alive(const SortJob * __restrict__ jobs, uint32_t njobs,
                               const uint32_t * __restrict__ starts, uint32_t nb)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= njobs) return;
    const SortJob & J = jobs[k];
    if (J.alive_out) *J.alive_out = starts[size_t(k) * nb + nb - 1] - uint32_t(J.base);
}

// each particle to the next free position of its bucket (order inside a bucket is arbitrary)
__global__ void k_bucket_scatter(const uint32_t * __restrict__ keys, const uint32_t * __restrict__ vals,
                                 uint32_t * __restrict__ starts, uint64_t total, uint32_t * __restrict__ out)
{
    const uint64_t g = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const uint32_t key = keys[g];
    // the lanes sharing a key take consecutive positions with one atomic (the lowest lane's)
    for (;;)
    {
        const uint32_t k0 = __builtin_amdgcn_readfirstlane(key);
        const uint64_t m = __ballot(key == k0);
        if (key == k0)
        {
            const uint32_t lead = uint32_t(__ffsll((unsigned long long)m) - 1);
            uint32_t base = 0;
            if (__lane_id() == lead) base = atomicAdd(&starts[k0], uint32_t(__popcll(m)));
            base = __shfl(base, int(lead), 64);
            out[base + uint32_t(__popcll(m & ((1ull << __lane_id()) - 1)))] = vals[g];
            break;
        }
    }
}

__global__ void k_fill_u32(uint32_t * __restrict__ p, uint64_t n, uint32_t v)
{
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

int launch_bucket_sort(const SortJob * jobs, uint32_t njobs, uint64_t total, uint32_t shift,
                       uint32_t nb, uint32_t * keys, uint32_t * vals, uint32_t * counts,
                       uint32_t * starts, void * temp, size_t & temp_bytes, uint32_t * out,
                       void * stream)
{
    const uint64_t nbt = uint64_t(njobs) * nb;
    if (!temp) // sizing call
        return hipcub::DeviceScan::ExclusiveSum(nullptr, temp_bytes, counts, starts, int(nbt),
                                                hipStream_t(stream)) == hipSuccess ? 0 : -1;
    const uint32_t block = 256;
    hipStream_t st = hipStream_t(stream);
    hipLaunchKernelGGL(k_fill_u32, dim3(uint32_t((nbt + block - 1) / block)), dim3(block), 0, st, counts, nbt, 0u);
    if (total)
        hipLaunchKernelGGL(k_bucket_keys, dim3(uint32_t((total + block - 1) / block)), dim3(block), 0, st,
                           jobs, njobs, total, keys, vals, counts, shift, nb);
    if (hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, counts, starts, int(nbt), st) != hipSuccess) return -1;
    hipLaunchKernelGGL(k_bucket_alive, dim3((njobs + 63) / 64), dim3(64), 0, st, jobs, njobs, starts, nb);
    if (total)
        hipLaunchKernelGGL(k_bucket_scatter, dim3(uint32_t((total + block - 1) / block)), dim3(block), 0, st,
                           keys, vals, starts, total, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Every (array, job) entry of a multi-job reorder in one launch: blockIdx.y picks the entry,
// dst[i] = src[perm[i]] for its n elements.
__global__ void k_gather_multi(const GatherEntry * __restrict__ tab)
{
    const GatherEntry & E = tab[blockIdx.y];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < E.n) E.dst[i] = E.src[E.perm ? E.perm[i] : i];
}

int launch_gather_multi(const GatherEntry * tab, uint32_t nent, uint32_t maxn, void * stream)
{
    const uint32_t block = 256;
    const uint32_t grid = (maxn + block - 1) / block;
    if (grid == 0 || nent == 0) return 0;
    hipLaunchKernelGGL(k_gather_multi, dim3(grid, nent), dim3(block), 0, hipStream_t(stream), tab);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

__global__ void k_scatter32(const uint32_t * __restrict__ in, uint32_t * __restrict__ out,
                            const uint32_t * __restrict__ orig, uint64_t n, uint32_t w)
{
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t o = orig[i];
    for (uint32_t c = 0; c < w; ++c) out[o * w + c] = in[i * w + c];
}

int launch_scatter32(const void * in, void * out, const uint32_t * orig, uint64_t n, uint32_t w,
                     void * stream)
{
    const uint32_t block = 256;
    const uint32_t grid = uint32_t((n + block - 1) / block);
    if (grid == 0) return 0;
    hipLaunchKernelGGL(k_scatter32, dim3(grid), dim3(block), 0, hipStream_t(stream),
                       static_cast<const uint32_t *>(in), static_cast<uint32_t *>(out), orig, n, w);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

} // namespace gr
