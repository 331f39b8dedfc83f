// Device code shared by the gfx950 kernel translation units: storage types, the software
// sampler, the expression interpreter and the step body (advance_one) with its launch helpers.
//
// The step kernels (gr_k_advance.hip, gr_k_batch.hip) replace the one-launch-per-step
// `granite::integrate` kernel of the reference (source/granite/TrajectoryIntegrator.tpp:108-457)
// together with the host step loop that launched it (tpp:973-994): they run `nsteps`
// integration steps of one or more windows per launch, keeping the particle state in registers
// and writing saved columns to a device ring. Sampling is done in software from node-interleaved
// FP32 / FP16 / BF16 slices (one vector load per stencil corner and slice) instead of texture
// fetches (TrajectoryIntegrator.icc:22-129).
//
// Arithmetic follows the reference's operation order and is compiled without FP contraction,
// so the FP32 form performs exactly the IEEE operations of the step (reference semantics);
// only the sampler weights differ from the texture unit (exact here, optionally quantised to
// 1/256 for emulation).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
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
// and keeps the conversion defined (NaN -> -1). One v_med3_f32: the median of (a, -1, n) is
// fminf(fmaxf(a, -1), n) for every a (a NaN input gives the minimum of the other two, -1, as
// fmaxf drops the NaN first), without the two min/max operations and their canonicalisation.
__device__ __forceinline__ void split(float a, uint32_t n, uint32_t wmode, int & i, float & w)
{
    a = __builtin_amdgcn_fmed3f(a, -1.f, float(n));
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

// The interpreter's LDS is the launch's dynamic shared memory, sized by the host for the block
// and the launch's programs (expr_lds_bytes): the full kernel stages the reference slots 4..15
// (volumes, constants; EXPR_REF_SLOTS columns) and every kernel keeps stack entries below the
// two register-held values (depth - 2 columns); one column holds one float per lane. A launch
// without programs gets none.
extern __shared__ float gr_dyn_lds[];
__device__ __forceinline__ lds_float * expr_lds() { return (lds_float *)gr_dyn_lds; }

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

// A program that is one DOT instruction, (r0 * r1) + (r2 * r3) [+ (r4 * r5)] over register
// slots r_k = arg >> 3k & 7 [then its square root]: the instruction run_program would execute,
// without the dispatch loop (same operations, same result).
__device__ __forceinline__ float run_dot(uint32_t w, const refvec_t rv)
{
    const uint32_t sub = w & 0xfu, arg = w >> 8;
    float x = (rv[arg & 7u] * rv[(arg >> 3) & 7u]) + (rv[(arg >> 6) & 7u] * rv[(arg >> 9) & 7u]);
    if ((sub & 3) == 3) x = x + (rv[(arg >> 12) & 7u] * rv[(arg >> 15) & 7u]);
    if (sub & 4) x = sqrtf(x);
    return x;
}

// rv: reference slots x y z cv u v w of the step; R: this lane's LDS reference column (volumes
// and constants, full kernel only)
template <bool FULL>
__device__ __forceinline__ float run_program(const Program & pr, const refvec_t rv,
                                             const lds_float * R, lds_float * stk, uint32_t lane,
                                             uint32_t bs)
{
    auto refv = [&](uint32_t d) -> float {
        if (!FULL) return rv[d & 7u];
        return d < 7 ? rv[d] : R[(d - 11) * bs]; // argument 7 + s reads slot s (4..15): column s - 4
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

// Feature set of a step kernel. SPEC < 0: every flag is read from the launch parameters (the
// generic kernels). SPEC >= 0: the plain feature set fixed at compile time -- exact sampler
// weights, uniform z spacing, no reverse integration, global coordinates -- with curvature mode
// SPEC & SPEC_CURVATURE and Kahan-compensated positions when SPEC_KAHAN is set (lean_spec picks
// it on the host). Fixed flags remove the uniform branches of the unused features from the step
// loop (fewer scalar instructions, no register moves at their joins, lower SGPR pressure);
// results are the same as the generic kernel's.
enum : int { SPEC_CURVATURE = 15, SPEC_KAHAN = 16 };
template <int SPEC> struct Feat
{
    static __device__ __forceinline__ bool kahan(const StepParams & P)
    {
        return SPEC < 0 ? (P.flags & F_KAHAN) != 0 : (SPEC & SPEC_KAHAN) != 0;
    }
    static __device__ __forceinline__ bool reverse(const StepParams & P) { return SPEC < 0 && (P.flags & F_REVERSE); }
    static __device__ __forceinline__ bool local(const StepParams & P) { return SPEC < 0 && (P.flags & F_LOCAL); }
    static __device__ __forceinline__ bool zlev(const StepParams & P) { return SPEC < 0 && P.wind.zlev; }
    static __device__ __forceinline__ uint32_t weights(const StepParams & P)
    {
        return SPEC < 0 ? P.weights : uint32_t(GR_WEIGHTS_EXACT);
    }
    static __device__ __forceinline__ uint32_t curvature(const StepParams & P)
    {
        return SPEC < 0 ? P.curvature : uint32_t(SPEC & SPEC_CURVATURE);
    }
};

template <int DIM, int SPEC>
__device__ __forceinline__ void texel_arg(const StepParams & P, const float * x, float * arg)
{
#pragma unroll
    for (int d = 0; d < DIM; ++d) arg[d] = (x[d] - P.off[d]) * P.inv_h[d];
    if (DIM == 3 && Feat<SPEC>::zlev(P)) arg[2] = zcoord(P.wind, x[2] - P.off[2]);
}

// The last time-paired stencil a lane loaded: its base node and the raw 32-bit words of its
// corners. 16-bit storage (FP16 / BF16): words x, y, z of the 16-byte 3D nodes [u0 v0 | w0 u1 |
// v1 w1 | pad], both words of the 8-byte 2D nodes [u0 v0 | u1 v1]; FP32: the six (3D) or four
// (2D) floats of the node's two slices. The RK4 stages of a step mostly sample the same cell (a
// particle crosses about a fifth of a cell per step), so with CACHE a lane whose base node is
// unchanged reuses the words instead of loading the stencil again -- the same data, the same
// result -- and the wavefront's gathers touch only the cache lines of the lanes that moved to
// another cell. That is what bounds the dense launches (the L1 works through the distinct lines
// of every gather), so the kernels are built with CACHE unless a caller asks otherwise; holding
// the words costs 24 (16-bit) or 48 (FP32) VGPRs in 3D.
template <int DIM, class T>
struct StencilCache
{
    static constexpr int NCOR = DIM == 3 ? 8 : 4;
    static constexpr int NW = sizeof(T) == 2 ? (DIM == 3 ? 3 : 2) : (DIM == 3 ? 6 : 4);
    uint32_t base;
    uint32_t w[NCOR][NW];
};

// the raw words of corner node `node` of a time-paired slot copy
template <int DIM, class T>
__device__ __forceinline__ void load_pair_node(const void * pair, uint32_t node,
                                               uint32_t (&w)[StencilCache<DIM, T>::NW])
{
    if constexpr (sizeof(T) == 2)
    {
        if constexpr (DIM == 3)
        {
            const uint4 h = GR_LOAD(uint4, pair, node);
            w[0] = h.x;
            w[1] = h.y;
            w[2] = h.z;
        }
        else
        {
            const uint2 h = GR_LOAD(uint2, pair, node);
            w[0] = h.x;
            w[1] = h.y;
        }
    }
    else
    {
        if constexpr (DIM == 3)
        {
            // [u0 v0 w0 u1 | v1 w1 0 0]: the padding half is not read
            const uint4 a = GR_LOAD(uint4, pair, 2 * uint64_t(node));
            const uint2 c = GR_LOAD(uint2, pair, 4 * uint64_t(node) + 2);
            w[0] = a.x;
            w[1] = a.y;
            w[2] = a.z;
            w[3] = a.w;
            w[4] = c.x;
            w[5] = c.y;
        }
        else
        {
            const uint4 a = GR_LOAD(uint4, pair, node);
            w[0] = a.x;
            w[1] = a.y;
            w[2] = a.z;
            w[3] = a.w;
        }
    }
}

// The same from the node's first byte `q` (16-byte aligned) of a paired copy smaller than 4 GiB:
// the caller forms q from the copy's (wavefront-uniform) base and a 32-bit byte offset, so the
// loads take the base + 32-bit offset + immediate form and a corner costs no 64-bit address
// arithmetic.
template <int DIM, class T>
__device__ __forceinline__ void load_pair_node_at(const char * q, uint32_t (&w)[StencilCache<DIM, T>::NW])
{
    if constexpr (sizeof(T) == 4 && DIM == 3)
    {
        const uint4 a = GR_LOAD(uint4, q, 0);
        const uint2 c = GR_LOAD(uint2, q + 16, 0);
        w[0] = a.x;
        w[1] = a.y;
        w[2] = a.z;
        w[3] = a.w;
        w[4] = c.x;
        w[5] = c.y;
    }
    else if constexpr (sizeof(T) == 2 && DIM == 2)
    {
        const uint2 h = GR_LOAD(uint2, q, 0);
        w[0] = h.x;
        w[1] = h.y;
    }
    else
    {
        const uint4 a = GR_LOAD(uint4, q, 0);
#pragma unroll
        for (int k = 0; k < StencilCache<DIM, T>::NW; ++k) w[k] = k == 0 ? a.x : k == 1 ? a.y : k == 2 ? a.z : a.w;
    }
}

// bytes of one node of the time-paired copy
template <int DIM, class T> constexpr uint32_t pair_node_bytes()
{
    return sizeof(T) == 4 ? (DIM == 3 ? 32u : 16u) : (DIM == 3 ? 16u : 8u);
}

// back (x) and front (y) slice values of a cached BF16 / FP32 corner (FP16 goes through
// combine_pair_h, which reads the words directly)
template <int DIM, class T>
__device__ __forceinline__ void decode_pair_node(const uint32_t (&w)[StencilCache<DIM, T>::NW],
                                                 float (&x)[DIM], float (&y)[DIM])
{
    if constexpr (sizeof(T) == 4)
    {
#pragma unroll
        for (int d = 0; d < DIM; ++d)
        {
            x[d] = __uint_as_float(w[d]);
            y[d] = __uint_as_float(w[DIM + d]);
        }
    }
    else if constexpr (DIM == 3)
    {
        x[0] = b_lo(w[0]); x[1] = b_hi(w[0]); x[2] = b_lo(w[1]);
        y[0] = b_hi(w[1]); y[1] = b_lo(w[2]); y[2] = b_hi(w[2]);
    }
    else
    {
        x[0] = b_lo(w[0]); x[1] = b_hi(w[0]);
        y[0] = b_lo(w[1]); y[1] = b_hi(w[1]);
    }
}

// --------------------------------------------------------------------------------------------
// field tile (k_advance_tile)
// --------------------------------------------------------------------------------------------
// Every step, a workgroup of neighbouring (cell-ordered) particles stages in LDS the box of
// time-paired nodes that covers the stencils of its live particles: the padded base nodes of
// their stage-1 samples, widened by tile_halo cells on every side (the later RK4 stages move a
// fraction of a cell), plus the +1 corner. A stencil inside the box reads its corners from LDS;
// any other (a stage that left the box, or a step whose box holds more than tile_cap nodes)
// loads them from the paired copy as the untiled kernels do. The words are the same either
// way, so results do not depend on the tile. The box keeps the copy's node layout
// (pair_node_bytes per node), x fastest.
//
// LDS area at tile_off: two sets (step parity) of four 32-bit words per wavefront for the box
// reduction, then the nodes.
typedef __attribute__((address_space(3))) char lds_char;
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) u32x2_t lds_u2;
typedef __attribute__((address_space(3))) u32x4_t lds_u4;
enum : uint32_t { TILE_SLOT_BYTES = 2 * 4 * 4 * 4, TILE_MAX_WAVES = 4 };
static_assert(TILE_SLOT_BYTES == TILE_AREA_HEADER, "tile area header");

struct TileBox
{
    uint32_t lo[3];  // first padded node of the box on each axis
    uint32_t ext[3]; // nodes on each axis; {1, 1, 1} when the step has no tile (nothing fits)
};

// the raw words of a node of the LDS box (same words as load_pair_node_at)
template <int DIM, class T>
__device__ __forceinline__ void lds_pair_node(const lds_char * q, uint32_t (&w)[StencilCache<DIM, T>::NW])
{
    if constexpr (sizeof(T) == 4 && DIM == 3)
    {
        const u32x4_t a = *(const lds_u4 *)q;
        const u32x2_t c = *(const lds_u2 *)(q + 16);
        w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w; w[4] = c.x; w[5] = c.y;
    }
    else if constexpr (sizeof(T) == 2 && DIM == 2)
    {
        const u32x2_t h = *(const lds_u2 *)q;
        w[0] = h.x;
        w[1] = h.y;
    }
    else
    {
        const u32x4_t a = *(const lds_u4 *)q;
#pragma unroll
        for (int k = 0; k < StencilCache<DIM, T>::NW; ++k) w[k] = k == 0 ? a.x : k == 1 ? a.y : k == 2 ? a.z : a.w;
    }
}

// per-half minimum of two packed 16-bit values
__device__ __forceinline__ uint32_t min_u16x2(uint32_t a, uint32_t b)
{
    const uint32_t lo = min(a & 0xffffu, b & 0xffffu), hi = min(a >> 16, b >> 16);
    return lo | (hi << 16);
}

// Forms this step's box and stages it (every thread of the workgroup calls it, with the same
// `parity` = step & 1). Returns false when no particle of the workgroup is alive (the workgroup
// is done). `live`: the lane's particle is alive at the start of the step; p: its position.
template <int DIM, class T, int SPEC>
__device__ __forceinline__ bool tile_step(const StepParams & P, const void * pair, const float * p, bool live,
                                          uint32_t parity, TileBox & tb)
{
    constexpr uint32_t NB = pair_node_bytes<DIM, T>();
    lds_char * area = (lds_char *)gr_dyn_lds + P.tile_off;
    lds_u32 * slots = (lds_u32 *)area + parity * (TILE_MAX_WAVES * 4);
    lds_char * nodes = area + TILE_SLOT_BYTES;
    const uint32_t n[3] = {P.wind.nx, P.wind.ny, P.wind.nz};
    // per axis, packed: low half the smallest padded base node of the lane's wavefront, high half
    // 0xffff minus the largest (the minimum of both halves over lanes gives the range)
    uint32_t m[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu};
    if (live)
    {
        float arg[3] = {0.f, 0.f, 0.f};
        texel_arg<DIM, SPEC>(P, p, arg);
        const uint32_t wmode = Feat<SPEC>::weights(P);
#pragma unroll
        for (int d = 0; d < DIM; ++d)
        {
            int i;
            float w;
            split(arg[d], n[d], wmode, i, w);
            const uint32_t b = uint32_t(i + 1); // in [0, n + 1]
            m[d] = b | ((0xffffu - b) << 16);
        }
    }
#pragma unroll
    for (int d = 0; d < DIM; ++d)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m[d] = min_u16x2(m[d], uint32_t(__shfl_xor(int(m[d]), o, 64)));
    const uint32_t wave = threadIdx.x >> 6, nwaves = (blockDim.x + 63u) >> 6;
    if ((threadIdx.x & 63u) == 0)
#pragma unroll
        for (int d = 0; d < DIM; ++d) slots[wave * 4 + d] = m[d];
    __syncthreads();
    uint32_t r[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu};
    for (uint32_t k = 0; k < nwaves; ++k)
#pragma unroll
        for (int d = 0; d < DIM; ++d) r[d] = min_u16x2(r[d], slots[k * 4 + d]);
    if ((r[0] & 0xffffu) + (r[0] >> 16) > 0xffffu) return false; // smallest > largest: none alive
    uint32_t vol = 1;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        if (d < DIM)
        {
            const uint32_t lo = r[d] & 0xffffu, hi = 0xffffu - (r[d] >> 16), last = n[d] + 2;
            const uint32_t a = lo > P.tile_halo ? lo - P.tile_halo : 0u;
            const uint32_t b = min(hi + 1u + P.tile_halo, last);
            tb.lo[d] = a;
            tb.ext[d] = b - a + 1u;
        }
        else
        {
            tb.lo[d] = 0;
            tb.ext[d] = 1;
        }
        vol *= tb.ext[d];
    }
    const bool fits = vol <= P.tile_cap;
    if (P.tile_stats && threadIdx.x == 0)
        atomicAdd(P.tile_stats + size_t(blockIdx.x % WORK_SLOTS) * WORK_STRIDE + (fits ? 0 : 1), 1ull);
    if (!fits)
    {
#pragma unroll
        for (int d = 0; d < 3; ++d) tb.ext[d] = 1; // no tile this step: every stencil loads globally
        return true;
    }
    // stage the box: consecutive threads copy consecutive nodes of a row (coalesced)
    const uint32_t PX = n[0] + 3, PY = DIM == 3 ? n[1] + 3 : 1u;
    const char * src = static_cast<const char *>(pair);
    for (uint32_t idx = threadIdx.x; idx < vol; idx += blockDim.x)
    {
        const uint32_t row = idx / tb.ext[0], x = idx - row * tb.ext[0];
        const uint32_t z = row / tb.ext[1], y = row - z * tb.ext[1];
        const uint64_t node = (uint64_t(tb.lo[2] + z) * PY + (tb.lo[1] + y)) * PX + (tb.lo[0] + x);
        const char * s = src + node * NB;
        lds_char * dst = nodes + idx * NB;
        if constexpr (NB == 8)
            *(lds_u2 *)dst = GR_LOAD(u32x2_t, s, 0);
        else
        {
            *(lds_u4 *)dst = GR_LOAD(u32x4_t, s, 0);
            if constexpr (NB == 32) *(lds_u2 *)(dst + 16) = GR_LOAD(u32x2_t, s + 16, 0); // not the pad
        }
    }
    __syncthreads();
    return true;
}

// Value of lane D of the calling lane's quad (lanes 4k .. 4k + 3), one DPP move.
template <int D> __device__ __forceinline__ float quad_lane(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), D * 0x55, 0xf, 0xf, false));
}

// Q = 4 (quad launches): the four lanes of a quad hold one particle and compute the same state;
// lane i of the quad interpolates velocity component min(i, DIM - 1) alone -- its back and front
// values at the eight (four) corners, one lerp tree -- and the components are exchanged inside
// the quad. The same operations on the same values as the one-lane form, a quarter of the lerps
// per lane: for launches of few particles, whose time is one lane's step chain.
// TL: the workgroup's field tile `tb` (tile_step) serves the stencils inside it.
template <int DIM, class T, int MODE, bool CACHE, int SPEC, int Q = 1, bool TL = false>
__device__ __forceinline__ void velocity(const StepParams & P, const Win & W, const float * x,
                                         float t, float * v, f16_t one, StencilCache<DIM, T> & C,
                                         const TileBox & tb)
{
    float arg[3] = {0.f, 0.f, 0.f};
    texel_arg<DIM, SPEC>(P, x, arg);
    const uint32_t wmode = Feat<SPEC>::weights(P);
    if (MODE == M_PAIRED)
    {
        // Padded paired copy: node (i, j, k), i in [-1, n + 1], at
        // ((k + 1) * PY + (j + 1)) * PX + (i + 1) with PX = nx + 3, PY = ny + 3; the halo holds
        // zeros (border addressing) or edge copies (clamp addressing), so the 8 corners are
        // fixed offsets from the base node and need no per-corner clamp or select.
        Taps S;
        int i, j, k = -1;
        split(arg[0], P.wind.nx, wmode, i, S.wx);
        split(arg[1], P.wind.ny, wmode, j, S.wy);
        S.wz = 0.f;
        if (DIM == 3) split(arg[2], P.wind.nz, wmode, k, S.wz);
        const uint32_t PX = P.wind.nx + 3, PXY = PX * (P.wind.ny + 3);
        const uint32_t base = uint32_t(k + 1) * PXY + uint32_t(j + 1) * PX + uint32_t(i + 1);
        if constexpr (Q == 4)
        {
            // this lane's component: its back value at byte comp * ES of a node, its front value
            // DIM * ES further; cached as raw words C.w[c][0] (back) and C.w[c][1] (front)
            constexpr uint32_t NB = pair_node_bytes<DIM, T>(), ES = sizeof(T), FO = DIM * ES;
            constexpr int NCOR = StencilCache<DIM, T>::NCOR;
            const uint32_t qi = threadIdx.x & 3u, comp = qi < uint32_t(DIM) ? qi : uint32_t(DIM - 1);
            if (!CACHE || base != C.base)
            {
                const char * pb = static_cast<const char *>(W.pair);
                auto word = [](const char * q) -> uint32_t {
                    if constexpr (ES == 4) return GR_LOAD(uint32_t, q, 0);
                    else return uint32_t(GR_LOAD(uint16_t, q, 0));
                };
                // quad launches address the paired copy with 32-bit byte offsets only (the host
                // runs them for copies smaller than 4 GiB)
                const uint32_t o0 = base * NB + comp * ES, oy = o0 + PX * NB, oz = o0 + PXY * NB, oyz = oz + PX * NB;
#pragma unroll
                for (int c = 0; c < NCOR; ++c)
                {
                    const char * q = pb + ((c & 4) ? ((c & 2) ? oyz : oz) : ((c & 2) ? oy : o0)) + ((c & 1) ? NB : 0u);
                    C.w[c][0] = word(q);
                    C.w[c][1] = word(q + FO);
                }
                C.base = base;
            }
            float vb[8][1], vf[8][1];
#pragma unroll
            for (int c = 0; c < NCOR; ++c)
            {
                if constexpr (ES == 4)
                {
                    vb[c][0] = __uint_as_float(C.w[c][0]);
                    vf[c][0] = __uint_as_float(C.w[c][1]);
                }
                else if constexpr (std::is_same<T, __half>::value)
                {
                    vb[c][0] = float(__builtin_bit_cast(f16_t, uint16_t(C.w[c][0])));
                    vf[c][0] = float(__builtin_bit_cast(f16_t, uint16_t(C.w[c][1])));
                }
                else
                {
                    vb[c][0] = __uint_as_float(C.w[c][0] << 16);
                    vf[c][0] = __uint_as_float(C.w[c][1] << 16);
                }
            }
            float b1[1], f1[1];
            combine<DIM, 1>(vb, S, b1);
            combine<DIM, 1>(vf, S, f1);
            const float vm = lerp(b1[0], f1[0], t);
            v[0] = quad_lane<0>(vm);
            v[1] = quad_lane<1>(vm);
            if (DIM == 3) v[2] = quad_lane<2>(vm);
            return;
        }
        if (!CACHE || base != C.base) // divergent: lanes whose cell is unchanged load nothing
        {
            // tile coordinates of the base node (unsigned: below the box wraps above it)
            const uint32_t ti = uint32_t(i + 1) - tb.lo[0], tj = uint32_t(j + 1) - tb.lo[1],
                           tk = uint32_t(k + 1) - tb.lo[2];
            if (TL && ti < tb.ext[0] - 1u && tj < tb.ext[1] - 1u && (DIM == 2 || tk < tb.ext[2] - 1u))
            {
                constexpr uint32_t NB = pair_node_bytes<DIM, T>();
                const uint32_t EX = tb.ext[0] * NB, EXY = EX * tb.ext[1];
                const lds_char * q0 = (const lds_char *)gr_dyn_lds + P.tile_off + TILE_SLOT_BYTES +
                                      ((tk * tb.ext[1] + tj) * tb.ext[0] + ti) * NB;
#pragma unroll
                for (int c = 0; c < StencilCache<DIM, T>::NCOR; ++c)
                    lds_pair_node<DIM, T>(q0 + ((c & 1) ? NB : 0u) + ((c & 2) ? EX : 0u) + ((c & 4) ? EXY : 0u),
                                          C.w[c]);
            }
            else if (P.wind.pair32) // uniform: byte offsets of the copy fit 32 bits
            {
                constexpr uint32_t NB = pair_node_bytes<DIM, T>();
                const char * pb = static_cast<const char *>(W.pair);
                const uint32_t o0 = base * NB, oy = o0 + PX * NB, oz = o0 + PXY * NB, oyz = oz + PX * NB;
#pragma unroll
                for (int c = 0; c < StencilCache<DIM, T>::NCOR; ++c)
                {
                    const uint32_t o = (c & 4) ? ((c & 2) ? oyz : oz) : ((c & 2) ? oy : o0);
                    load_pair_node_at<DIM, T>(pb + o + ((c & 1) ? NB : 0u), C.w[c]);
                }
            }
            else
            {
#pragma unroll
                for (int c = 0; c < StencilCache<DIM, T>::NCOR; ++c)
                {
                    const uint32_t node = base + ((c & 1) ? 1u : 0u) + ((c & 2) ? PX : 0u) + ((c & 4) ? PXY : 0u);
                    load_pair_node<DIM, T>(W.pair, node, C.w[c]);
                }
            }
            C.base = base;
        }
        float b[DIM], f[DIM];
        if constexpr (std::is_same<T, __half>::value)
        {
            uint4 raw[8];
#pragma unroll
            for (int c = 0; c < StencilCache<DIM, T>::NCOR; ++c)
                raw[c] = DIM == 3 ? make_uint4(C.w[c][0], C.w[c][1], C.w[c][2], 0u)
                                  : make_uint4(C.w[c][0], C.w[c][1], 0u, 0u);
            combine_pair_h<DIM>(raw, S, one, b, f);
        }
        else
        {
            float vb[8][DIM], vf[8][DIM];
#pragma unroll
            for (int c = 0; c < StencilCache<DIM, T>::NCOR; ++c) decode_pair_node<DIM, T>(C.w[c], vb[c], vf[c]);
            combine<DIM, DIM>(vb, S, b);
            combine<DIM, DIM>(vf, S, f);
        }
#pragma unroll
        for (int d = 0; d < DIM; ++d) v[d] = lerp(b[d], f[d], t);
        return;
    }
    Taps S;
    taps<DIM>(arg, P.wind.nx, P.wind.ny, P.wind.nz, P.wind.address, wmode, S);
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

template <int SPEC>
__device__ __forceinline__ float global_coord(const StepParams & P, float x, int d)
{
    return Feat<SPEC>::local(P) ? float(P.off64[d] + double(x)) : x;
}

// Advances one particle by the launch's steps; returns the number of steps it integrated
// (steps begun alive, the aborting step included).
// TL (k_advance_tile): every thread of the workgroup runs the step loop, also for a stopped
// particle or none (pid >= n), because the field tile is staged by the whole workgroup each step;
// such lanes only take part in the staging and store nothing.
template <int DIM, int INTEG, class T, int MODE, bool CACHE = false, int SPEC = -1, int Q = 1, bool TL = false>
__device__ __forceinline__ uint32_t advance_one(const StepParams & P, const uint32_t pid)
{
    // quad launches: every lane of the quad keeps the particle's state, the first one stores it
    const bool writer = Q == 1 || (threadIdx.x & 3u) == 0;
    constexpr bool FULL = MODE == M_FULL;
    static_assert(!FULL || SPEC < 0, "the full kernel reads every flag at run time");
    typedef Feat<SPEC> F;
    const uint32_t curvm = F::curvature(P);
    const bool kahan = F::kahan(P), reverse = F::reverse(P);
    const uint32_t oid = (FULL && P.orig) ? GR_LOAD(uint32_t, P.orig, pid) : pid; // caller-order index

    // window state; with nwin > 1 the launch crosses window edges (tpp:933-969 per window)
    Win W{P.wind.back, P.wind.front, P.wind.pair};
    uint32_t wi = 0;
    // the window tables are uniform and read-only: scalar loads keep their values in SGPRs
    uint32_t edge = P.nwin > 1 ? GR_CONST(uint32_t, P.win_edge, 0) : 0xffffffffu;

    // A stopped particle's state never changes again (tpp:452-455 copies it unchanged) and its
    // saved columns after the stop are left to the reader, which fills them from the final
    // state: nothing to do.
    int32_t status = (TL && pid >= P.n) ? int32_t(GR_STATUS_DOMAIN) : G(P.status)[pid];
    const bool entered = status == GR_STATUS_TIME;
    if (!TL && !entered) return 0;
    // TL: a lane without a live particle reads particle 0's state (in bounds) and never stores
    const uint32_t lp = (TL && !entered) ? 0u : pid;
    f16_t one = f16_t(1.0f);
    asm volatile("" : "+v"(one)); // opaque 1.0 for the mixed-precision lerps (lerp_h)

    float p[3] = {0.f, 0.f, 0.f}, q[3] = {0.f, 0.f, 0.f}, cmp[3] = {0.f, 0.f, 0.f};
    float ld[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < DIM; ++d) p[d] = G(P.p[d])[lp];
    const bool curv_on = curvm != GR_CURV_OFF;
    const bool curv_pos = curvm == GR_CURV_TOTAL || curvm == GR_CURV_INDIVIDUAL_AND_TOTAL;
    if (curv_pos)
#pragma unroll
        for (int d = 0; d < DIM; ++d) q[d] = G(P.q[d])[lp];
    if (curvm == GR_CURV_FAST_TOTAL)
#pragma unroll
        for (int d = 0; d < DIM; ++d) ld[d] = G(P.ld[d])[lp];
    if (kahan)
#pragma unroll
        for (int d = 0; d < DIM; ++d) cmp[d] = G(P.c[d])[lp];
    int32_t abort_step = G(P.abort_step)[lp];
    float total = curv_on ? G(P.total)[lp] : 0.f;
    float curv = curvm == GR_CURV_INDIVIDUAL_AND_TOTAL ? G(P.curv)[lp] : 0.f;
    float len = (FULL && P.abort_mode != GR_ABORT_TIME) ? G(P.len)[lp] : 0.f;
    float ex[GR_MAX_PROGRAMS];
#pragma unroll
    for (int e = 0; e < GR_MAX_PROGRAMS; ++e) ex[e] = e < int(P.nprog) ? G(P.expr[e])[lp] : 0.f;
    float vo[GR_MAX_VOLUMES];
#pragma unroll
    for (int k = 0; k < GR_MAX_VOLUMES; ++k) vo[k] = (FULL && k < int(P.nvol)) ? G(P.vol[k])[lp] : 0.f;

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
    StencilCache<DIM, T> sc;
    sc.base = 0xffffffffu; // no stencil loaded (also after every window change)
    TileBox tb = {{0u, 0u, 0u}, {1u, 1u, 1u}};
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
        if constexpr (TL)
            if (!tile_step<DIM, T, SPEC>(P, W.pair, p, status == GR_STATUS_TIME, s & 1u, tb))
                break; // no live particle left in the workgroup (the same decision in every lane)
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
                velocity<DIM, T, MODE, CACHE, SPEC, Q, TL>(P, W, p, t, k1, one, sc, tb);
                float k1c[3] = {k1[0], k1[1], k1[2]};
                if (sph) sphere(P, p[1], k1c, k1);
#pragma unroll
                for (int d = 0; d < DIM; ++d) x[d] = p[d] + k1[d] * half_dt;
                velocity<DIM, T, MODE, CACHE, SPEC, Q, TL>(P, W, x, t1, k2, one, sc, tb);
                float k2c[3] = {k2[0], k2[1], k2[2]};
                if (sph) sphere(P, x[1], k2c, k2);
#pragma unroll
                for (int d = 0; d < DIM; ++d) x[d] = p[d] + k2[d] * half_dt;
                velocity<DIM, T, MODE, CACHE, SPEC, Q, TL>(P, W, x, t1, k3, one, sc, tb);
                float k3c[3] = {k3[0], k3[1], k3[2]};
                if (sph) sphere(P, x[1], k3c, k3);
#pragma unroll
                for (int d = 0; d < DIM; ++d) x[d] = p[d] + k3[d] * dt;
                velocity<DIM, T, MODE, CACHE, SPEC, Q, TL>(P, W, x, t2, k4, one, sc, tb);
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
                velocity<DIM, T, MODE, CACHE, SPEC, Q, TL>(P, W, p, t, vc, one, sc, tb);
#pragma unroll
                for (int d = 0; d < DIM; ++d) v[d] = vc[d];
                if (sph) sphere(P, p[1], vc, v);
            }

            // dp (tpp:233-250)
            float dp[3] = {0.f, 0.f, 0.f};
            const float sgn = reverse ? -1.f : 1.f;
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
            if (kahan)
            {
#pragma unroll
                for (int d = 0; d < DIM; ++d)
                {
                    y[d] = (reverse ? -dp[d] : dp[d]) - cmp[d];
                    np[d] = p[d] + y[d];
                }
            }
            else
            {
#pragma unroll
                for (int d = 0; d < DIM; ++d) np[d] = reverse ? p[d] - dp[d] : p[d] + dp[d];
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
                        float ta[3] = {global_coord<SPEC>(P, np[0], 0) / P.h[0],
                                       global_coord<SPEC>(P, np[1], 1) / P.h[1], 0.f};
                        topo = sample_scalar<2>(P.topo.back, P.topo, 0, ta, P.weights) >
                               global_coord<SPEC>(P, np[2], 2);
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
                if (curvm == GR_CURV_FAST_TOTAL) // tpp:337-346
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
                            // (np - q) / (2 dt) and ((q + np) - 2 p) / (dt dt), each correctly
                            // rounded: one FP64 multiply by the denominator's FP64 reciprocal
                            d1[d] = float(double(np[d] - q[d]) * P.inv_2dt);
                            d2[d] = float(double((q[d] + np[d]) - 2.f * p[d]) * P.inv_dtdt);
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
                    if (curvm == GR_CURV_INDIVIDUAL_AND_TOTAL) curv = k;
                    if (!isnan(k)) total += k * ds;
                }

                if (FULL && (P.flags & (F_VOL_CONST | F_VOL_DYN))) // tpp:363-377
                {
                    float arg[3] = {0.f, 0.f, 0.f};
                    texel_arg<DIM, SPEC>(P, np, arg);
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
                    const uint32_t lane = threadIdx.x, bs = blockDim.x, m = P.refmask;
                    // column c of the reference area holds slot c + 4 (volumes 4..11, constants 12..15)
                    lds_float * s_stk = expr_lds() + (FULL ? EXPR_REF_SLOTS * bs : 0u);
                    lds_float * R = expr_lds() + lane;
                    if (FULL)
                    {
#pragma unroll
                        for (int k = 0; k < GR_MAX_VOLUMES; ++k)
                            if (m & (1u << (4 + k))) R[k * bs] = k < int(P.nvol) ? vo[k] : 0.f;
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
                            R[(8 + k) * bs] = cv;
                        }
                    }
                    float gx = 0.f, gy = 0.f, gz = 0.f;
                    if (m & 7u)
                    {
                        gx = global_coord<SPEC>(P, np[0], 0);
                        gy = global_coord<SPEC>(P, np[1], 1);
                        if (DIM == 3) gz = global_coord<SPEC>(P, np[2], 2);
                    }
                    const refvec_t rv = {gx, gy, gz, curvm == GR_CURV_INDIVIDUAL_AND_TOTAL ? curv : 0.f,
                                         vc[0], vc[1], DIM == 3 ? vc[2] : 0.f, 0.f, 0.f};
                    if (P.nprog == 1 && (P.dotmask & DOT_UVW)) // the squared speed [square root]
                    {
                        float x = (vc[0] * vc[0]) + (vc[1] * vc[1]);
                        if (DIM == 3) x = x + (vc[2] * vc[2]);
                        ex[0] = (P.prog[0].code[0] & 4u) ? sqrtf(x) : x;
                    }
                    else if (P.nprog == 1 && (P.dotmask & 1u)) // another single-DOT program
                        ex[0] = run_dot(P.prog[0].code[0], rv);
                    else
                        for (uint32_t e = 0; e < P.nprog; ++e)
                            ex[e] = run_program<FULL>(P.prog[e], rv, R, s_stk, lane, bs);
                }

                if (kahan)
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
        if (s == next_save && live && writer)
        {
            auto rp = G(P.ring_pos) + (uint64_t(col) * P.n + pid) * DIM;
#pragma unroll
            for (int d = 0; d < DIM; ++d) rp[d] = global_coord<SPEC>(P, p[d], d);
            if (curvm == GR_CURV_INDIVIDUAL_AND_TOTAL) G(P.ring_curv)[uint64_t(col) * P.n + pid] = curv;
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
        if (!TL && status != GR_STATUS_TIME) break; // stopped: nothing more to compute or save
    }

    if (!writer || (TL && !entered)) return 0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) G(P.p[d])[pid] = p[d];
    if (curv_pos)
#pragma unroll
        for (int d = 0; d < DIM; ++d) G(P.q[d])[pid] = q[d];
    if (curvm == GR_CURV_FAST_TOTAL)
#pragma unroll
        for (int d = 0; d < DIM; ++d) G(P.ld[d])[pid] = ld[d];
    if (kahan)
#pragma unroll
        for (int d = 0; d < DIM; ++d) G(P.c[d])[pid] = cmp[d];
    G(P.status)[pid] = status;
    G(P.abort_step)[pid] = abort_step;
    if (curv_on) G(P.total)[pid] = total;
    if (curvm == GR_CURV_INDIVIDUAL_AND_TOTAL) G(P.curv)[pid] = curv;
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

template <int DIM, int INTEG, class T, int MODE, bool CACHE = false, int SPEC = -1, int Q = 1, bool TL = false>
__device__ __forceinline__ void advance_block(const StepParams & P, uint32_t nxcd)
{
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x, nxcd);
    const uint32_t pid = (b * blockDim.x + threadIdx.x) / Q; // a quad per particle when Q = 4
    uint32_t w = 0;
    if (TL || pid < P.n) w = advance_one<DIM, INTEG, T, MODE, CACHE, SPEC, Q, TL>(P, pid);
    if (P.work) // block sizes are multiples of 64
        wave_add_work(P.work + size_t(blockIdx.x % WORK_SLOTS) * WORK_STRIDE, w);
}

} // namespace gr
