// Internal data contract between the host engine (gr_capi.cpp) and the gfx950 kernels
// (gr_step.h, compiled in the gr_k_*.hip translation units). Replaces the by-value SimulationData kernel argument of the reference
// (source/granite/SimulationData.hpp:21-70) with a POD of raw device pointers and uniform
// scalars; fields are node-interleaved arrays instead of texture objects.
#pragma once

#include <stdint.h>

#include "granite_hip.h"

namespace gr
{

// bytecode opcodes (low 8 bits of gr_program::code; arg in the high 24 bits)
enum Op : uint32_t
{
    OP_REF = 1,   // push ref[arg]
    OP_CONST = 2, // push imm[arg]
    OP_ADD = 3,
    OP_SUB = 4,
    OP_MUL = 5,
    OP_DIV = 6,
    OP_SQRT = 7,
    OP_EXP = 8,
    OP_LN = 9,
    OP_SIN = 10,
    OP_COS = 11,
    OP_TAN = 12,
    OP_ASIN = 13,
    OP_ACOS = 14,
    OP_ATAN = 15,
    OP_ATAN2 = 16,
    OP_ABS = 17
};

// Device form of a program (lower_program): the postfix code above with a leaf operand (a
// reference slot or an immediate) folded into the binary operation that consumes it, so that
// the two top stack values stay in registers and most programs touch no stack memory. An
// instruction is (class << 4 | sub) | (arg << 8):
//   DC_PUSH   sub 0: push ref[arg], 1: push imm[arg]
//   DC_BINARY sub = k | rev << 3 (k: 0 add, 1 sub, 2 mul, 3 div, 4 atan2): second op top, or
//             top op second when rev
//   DC_LEAF   sub = k | rev << 2 | imm << 3 (k < 4): top op leaf, or leaf op top when rev; the
//             leaf is ref[arg] or imm[arg]
//   DC_UNARY  sub: 0 sqrt 1 exp 2 ln 3 sin 4 cos 5 tan 6 asin 7 acos 8 atan 9 abs
//   DC_DOT    sub = terms | root << 2, terms = 2 or 3: push (r0 * r1) + (r2 * r3) [+ (r4 * r5)],
//             the sum formed left to right, with register slots r_k = arg >> 3k & 7 (the postfix
//             of a*b + c*d [+ e*f]); with root, its square root (the postfix followed by sqrt)
// Reference arguments are dense: 0..6 = slots x y z cv u v w (held in registers), 7 + s = slot
// s (volumes and constants, staged in LDS by the full kernel).
enum DevClass : uint32_t
{
    DC_PUSH = 0,
    DC_BINARY = 1,
    DC_LEAF = 2,
    DC_UNARY = 3,
    DC_DOT = 4
};

// Slot views of one field: node (i, j, k) of a slot starts at element
// ((k * ny + j) * nx + i) * cstride of that slot, components consecutive.
struct FieldView
{
    const void * back;
    const void * front;
    const void * pair;  // time-paired copy of (back, front), or null
    uint32_t nx, ny, nz;
    uint32_t cstride;  // elements per node (4 for 3 components, else ncomp)
    uint32_t ncomp;
    uint32_t address;  // 0 border (zero outside), 1 clamp
    uint32_t dtype;    // gr_dtype
    uint32_t pair32;   // the time-paired copy is smaller than 4 GiB: 32-bit byte offsets address it
    // stretched z axis (3D): node k at height zlev[k] (relative to Offset.Z); zlut[m] is the
    // last level at or below zlut_z0 + m / zlut_scale. Null for uniform spacing.
    uint32_t nlut;
    const float * zlev;
    const int32_t * zlut;
    float zlut_z0, zlut_scale;
};

// job flags (uniform)
enum : uint32_t
{
    F_DYNAMIC_ANY = 1u << 0,  // any Dynamic mode -> time weight t is computed (tpp:127-133)
    F_WIND_DYNAMIC = 1u << 1, // lerp between back and front wind slices
    F_REVERSE = 1u << 2,
    F_TIME64 = 1u << 3,
    F_KAHAN = 1u << 4,
    F_LOCAL = 1u << 5,        // positions are relative to off64
    F_SPHERE = 1u << 6,
    F_TOPO = 1u << 7,
    F_UPLIFT_CONST = 1u << 8,
    F_UPLIFT_DYN = 1u << 9,
    F_CONST_CONST = 1u << 10,
    F_CONST_DYN = 1u << 11,
    F_VOL_CONST = 1u << 12,
    F_VOL_DYN = 1u << 13,
};

struct Program
{
    uint32_t n;
    uint32_t code[GR_MAX_CODE];
    float imm[GR_MAX_CODE];
};

// Lowers a compiled program (gr_expr_compile) to its device form (DevOp); the reference slots
// it reads are or-ed into refmask. Returns GR_OK or GR_E_INVALID (malformed program).
int lower_program(const gr_program & in, Program & out, uint32_t & refmask);
// Host evaluation of a device-form program with the step kernels' semantics.
float eval_lowered(const Program & pr, const float * ref19);

struct StepParams
{
    uint32_t n;          // particles
    uint32_t nsteps;     // steps in this launch
    uint32_t flags;
    uint32_t border;     // gr_border
    uint32_t curvature;  // gr_curvature
    uint32_t abort_mode; // gr_abort_mode
    uint32_t weights;    // gr_weights
    uint32_t nprog, nvol, nconst;
    uint32_t refmask; // reference-vector slots read by the programs (bit per slot)
    // bit e: program e is a single DOT instruction (a sum of two or three products of register
    // slots, optionally under a square root), evaluated without the interpreter's dispatch loop;
    // DOT_UVW: program 0's DOT is u*u + v*v [+ w*w] (the squared speed, in that order)
    uint32_t dotmask;

    float dt, sphere_radius, max_length;
    // 1 / (2 dt) and 1 / (dt dt) rounded to FP64, the two float denominators (2.f * dt, dt * dt)
    // of the curvature's finite differences: float(double(x) * r) is the correctly rounded
    // x / denominator for every float x (the product's error, 2^-52, is below the distance of any
    // float quotient from a rounding boundary), at the cost of one FP64 multiply
    double inv_2dt, inv_dtdt;
    float h[3], inv_h[3], off[3], dom[3];
    double off64[3];

    // time of this window
    float tb, tf, t_start;
    double tb64, tf64, t_start64;
    uint64_t counter_start;
    uint64_t save_interval; // 0 = no in-kernel saves
    uint64_t save_first;    // step index (within this launch) of the first save, or ~0
    uint64_t save_g0;       // global save index of that first save
    uint32_t ring_cap;      // columns in the device ring

    // launches that cross window edges: window i of the launch (i = 0 is `wind` above) reads
    // win_ptr[3 i + {0, 1, 2}] = (back, front, pair) and starts at step win_edge[i - 1];
    // its time stamps are (win_first + i) * dtd and (win_first + i + 1) * dtd
    uint32_t nwin;
    uint32_t win_first;
    float dtd;
    const void * const * win_ptr;
    const uint32_t * win_edge;
    // per step of this launch (required for every launch that has steps): the time weights
    // (t, t + dt/2, t + dt mapped into the step's window) and 1 / 0 for "the step starts after
    // time 0", four floats per step, formed on the host with the reference's FP32 / FP64
    // operations
    const float * ttab;

    // particle state (SoA)
    float * p[3];
    float * q[3];    // previous position (curvature)
    float * c[3];    // Kahan compensation
    float * ld[3];   // last direction (fast total curvature)
    int32_t * status;
    int32_t * abort_step;
    float * total;
    float * curv;
    float * len;
    float * expr[GR_MAX_PROGRAMS];
    float * vol[GR_MAX_VOLUMES];

    // saved-column ring
    float * ring_pos;  // [cap][n][dim]
    float * ring_curv; // [cap][n]
    float * ring_expr; // [nprog][cap][n]
    float * ring_vol;  // [nvol][cap][n]

    // fields
    FieldView wind;
    FieldView vols;
    FieldView topo;

    // original particle index of each (reordered) state slot, or null for identity; used for
    // per-particle inputs that stay in the caller's order
    const uint32_t * orig;

    // per-particle data
    const float * lift_b;
    const float * lift_f;
    const float * cst_b[GR_MAX_CONSTANTS];
    const float * cst_f[GR_MAX_CONSTANTS];

    Program prog[GR_MAX_PROGRAMS];

    // integrated particle-steps (steps begun alive) are added here, or null
    unsigned long long * work;

    // field tile of the tiled time-paired kernel (k_advance_tile): byte offset of its area in the
    // workgroup's dynamic LDS (after the interpreter's columns; 16-byte aligned), the paired-copy
    // nodes it holds, and the cells added on every side of the workgroup's stage-1 cells
    uint32_t tile_off, tile_cap, tile_halo;
    // workgroup-steps with a tile (slot element 0) and without one (element 1), counted in
    // WORK_SLOTS slots WORK_STRIDE apart like the work counter; null: not counted
    unsigned long long * tile_stats;
};
enum : uint32_t { DOT_UVW = 1u << 16 };
// bytes at the start of the tile area before its nodes (the box reduction's words)
enum : uint32_t { TILE_AREA_HEADER = 128 };

// LDS of the expression interpreter, in columns of one float per lane of the workgroup: the
// full kernel stages reference slots 4..15 (EXPR_REF_SLOTS columns), and every kernel holds the
// stack entries below the two register-held values (maximum stack depth - 2 columns).
enum : uint32_t { EXPR_REF_SLOTS = 12 };

// kernel modes (see gr_advance.h)
enum : uint32_t { MODE_PAIRED = 0, MODE_SEPARATE = 1, MODE_FULL = 2 };

// Compile-time feature set of a lean launch (the step kernels' SPEC parameter) when the launch
// uses the plain features only -- exact sampler weights, uniform z spacing, no reverse
// integration, global coordinates -- and a specialised build exists for its curvature mode (Off,
// IndividualAndTotal): that mode, plus 16 with Kahan-compensated positions (the kernels'
// SPEC_KAHAN bit); -1 (generic build) otherwise. The specialised sets are 0, 3, 16 and 19.
inline int lean_spec(const StepParams & P, uint32_t mode)
{
    if (mode == MODE_FULL) return -1;
    if (P.flags & (F_REVERSE | F_LOCAL)) return -1;
    if (P.wind.zlev || P.weights != GR_WEIGHTS_EXACT) return -1;
    if (P.curvature != GR_CURV_OFF && P.curvature != GR_CURV_INDIVIDUAL_AND_TOTAL) return -1;
    return int(P.curvature) + ((P.flags & F_KAHAN) ? 16 : 0);
}

// one work item of the persistent batched kernel: `count` particles of job `job` from `first`
// on (<= 64, one wavefront; <= QUAD_ITEM for quad items)
struct BatchItem
{
    uint32_t job;
    uint32_t first;
    uint32_t count;
    uint32_t variant; // batch_variant(dim, integrator, dtype, spec), + BATCH_QUAD for quad items
};
// BATCH_ITEM: particles per item (one wavefront). Queue data of a launch (qdata): nq counters
// QUEUE_STRIDE words apart (zero at launch), nq + 1 queue starts, ceil(nq / 32) mask words (bit
// set for every queue holding items). The work counter is WORK_SLOTS 64-bit slots WORK_STRIDE
// apart (summed by the host). CU_KEY_BITS: width of cu_key() values (the CU lookup table size).
enum : uint32_t
{
    BATCH_ITEM = 64,
    MAX_QUEUES = 1024,
    QUEUE_STRIDE = 16,
    WORK_SLOTS = 256,
    WORK_STRIDE = 8,
    CU_KEY_BITS = 11
};
// item variant: 12 * specialisation (0 generic, 1 plain with curvature off, 2 plain with
// individual-and-total curvature) + 6 * (dim == 3) + 3 * (RK4) + storage type
int batch_variant(uint32_t dim, uint32_t integrator, uint32_t dtype, int spec);
// quad items: variant BATCH_QUAD + v (v a specialised variant, 12 .. 59), QUAD_ITEM particles
// each, run only by the stencil-reuse (narrow) build
enum : uint32_t { BATCH_VARIANTS = 60, BATCH_QUAD = 64, QUAD_ITEM = 16 };
// culut[cu_key] = dense index of that CU (0xffffffff: none); a CU's queue is index / group;
// work: the work counter slots
// waves: 0 / 1 = the compiler's register allocation, 4 = at most 128 VGPRs (4 wavefronts/SIMD);
// narrow: every item's wavefront fits on the device at once (selects the build that reuses FP16
// stencils across RK4 stages, at two wavefronts per SIMD)
// lds_bytes: dynamic shared memory of each workgroup (the interpreter's columns, possibly
// padded to cap the workgroups per CU, see launch_lds_bytes)
int launch_batch(const StepParams * jobs, const BatchItem * items, uint32_t * qdata, uint32_t nq,
                 const uint32_t * culut, uint32_t group, unsigned long long * work, uint32_t grid,
                 uint32_t block, uint32_t waves, bool narrow, uint32_t lds_bytes, void * stream);
// the per-variant batched kernel of specialised variant `var` (12 .. 59): for launches whose
// items all have that variant (or its quad form); four wavefronts per SIMD, stencil reuse.
// Returns -2 for a variant without one.
int launch_batch_uni(uint32_t var, const StepParams * jobs, const BatchItem * items, uint32_t * qdata, uint32_t nq,
                     const uint32_t * culut, uint32_t group, unsigned long long * work, uint32_t grid,
                     uint32_t block, uint32_t lds_bytes, void * stream);
int batch_uni_occupancy(uint32_t var, uint32_t block, uint32_t lds_bytes, int * blocks_per_cu);
// writes cu_key() of the CU running each of n workgroups
int launch_cu_probe(uint32_t * out, uint32_t n, void * stream);
int batch_occupancy(uint32_t block, uint32_t waves, uint32_t lds_bytes, int * blocks_per_cu);
// Dynamic shared memory of a workgroup of `block` threads: the interpreter's `lds_cols` columns,
// and, with cu_waves > 0, at least enough that no more than max(1, cu_waves / (block / 64))
// workgroups fit in the CU's `lds_per_cu` bytes (an occupancy cap: fewer wavefronts per CU keep
// their stencil gathers inside the CU's cache). Padding is rounded down to 1 KiB so that the cap
// holds whatever the allocation granule.
inline uint32_t launch_lds_bytes(uint32_t block, uint32_t lds_cols, uint32_t cu_waves, uint32_t lds_per_cu)
{
    uint32_t need = lds_cols * block * uint32_t(sizeof(float));
    if (cu_waves && lds_per_cu)
    {
        const uint32_t wpb = block / 64 ? block / 64 : 1;
        const uint32_t blocks = cu_waves / wpb ? cu_waves / wpb : 1;
        const uint32_t pad = (lds_per_cu / blocks) & ~1023u;
        if (pad > need) need = pad;
    }
    return need;
}

// launchers (gr_k_advance.hip, gr_k_support.hip)
// nxcd: XCDs of the device (workgroups are renumbered so each XCD runs one contiguous range);
// nrun (<= P.n): the launch covers particles [0, nrun) -- every particle from nrun on is stopped;
// wide: the launch holds many more wavefronts than the device runs at once (selects a register
// budget of four wavefronts per SIMD where a kernel variant has one); narrow: all its wavefronts
// fit at two per SIMD (selects the FP16 build that reuses stencils across RK4 stages)
// spec: the kernel's compile-time feature set (lean_spec, or -1 for the generic build);
// lds_bytes: as for launch_batch
// quad: the launch runs a quad of lanes per particle (k_advance_quad; spec >= 0 and the paired
// mode only, fewer than 2^29 particles)
// tile: the 3D time-paired kernel with the workgroup's field tile (k_advance_tile; P.tile_* set,
// lds_bytes covering the tile area; not with quad)
int launch_advance(const StepParams & P, uint32_t dim, uint32_t integrator, uint32_t mode, int spec,
                   uint32_t block, uint32_t nxcd, uint64_t nrun, bool wide, bool narrow, bool quad,
                   uint32_t lds_bytes, void * stream, bool tile = false);
// Scratch of the no-allocation entry points (gr_k_noalloc.hip): NOALLOC_MAX_STEPS steps of time
// table in a file-scope device array. noalloc_setup takes its address on the current device
// (once per device, at context creation); noalloc_scratch returns it (null before setup).
// launch_time_table forms the time table of a single-window launch P in `out` on the device.
enum : uint32_t { NOALLOC_MAX_STEPS = GR_NOALLOC_MAX_STEPS };
int noalloc_setup(int device);
void * noalloc_scratch(int device);
int launch_time_table(const StepParams & P, float * out, void * stream);

// builds the padded time-paired layout of (back, front) slots: (nx+3)(ny+3)(nz+3) nodes (2D: one
// plane), elements per node 8 (3D) / 4 (2D)
uint64_t pair_nodes(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t dim);
int launch_pair(const void * back, const void * front, void * out, uint32_t nx, uint32_t ny,
                uint32_t nz, uint32_t dim, uint32_t address, uint32_t ncomp, uint32_t cstride,
                uint32_t dtype, void * stream);
// one job of a multi-job reorder: its particles occupy keys [base, base + n); after the sort
// the number of its alive particles (all placed first) is stored to *alive_out
struct SortJob
{
    const float * p[3];
    const int32_t * status;
    uint32_t * alive_out;
    uint64_t base;
    uint32_t n, dim;
    float off[3], inv_h[3];
    uint32_t grid[3];
};
// dst[i] = src[perm[i]] (perm null: dst[i] = src[i]), i < n (4-byte elements)
struct GatherEntry
{
    const uint32_t * src;
    uint32_t * dst;
    const uint32_t * perm;
    uint32_t n;
    uint32_t pad;
};
// Regroups the particles [base, base + n) of every job by the Morton code of their current cell
// with the low `shift` bits dropped (stopped particles last), keeping each job in its range: out
// [base + r] is the in-job index of the job's r-th particle in that order. A counting sort over
// njobs * nb buckets (nb - 1 cell buckets and the stopped bucket per job); also stores each job's
// alive count to *alive_out. keys, vals: total entries of scratch; counts, starts: njobs * nb;
// temp: scan storage sized by a first call with temp == null (temp_bytes set, nothing launched).
int launch_bucket_sort(const SortJob * jobs, uint32_t njobs, uint64_t total, uint32_t shift,
                       uint32_t nb, uint32_t * keys, uint32_t * vals, uint32_t * counts,
                       uint32_t * starts, void * temp, size_t & temp_bytes, uint32_t * out,
                       void * stream);
// copies `bytes` (a multiple of 16, both pointers 16-byte aligned) from device-visible pinned host
// memory to device memory with a kernel on `stream`
int launch_stage_copy(const void * src, void * dst, size_t bytes, void * stream);
// grid-stride 16-byte copy of `bytes` (a multiple of 16) with `grid` workgroups of 256
int launch_stream_copy(const void * src, void * dst, size_t bytes, uint32_t grid, void * stream);
// all entries in one launch (maxn: largest entry n)
int launch_gather_multi(const GatherEntry * tab, uint32_t nent, uint32_t maxn, void * stream);
// out[orig[i] * w + c] = in[i * w + c] for 4-byte elements, w = 1..3
int launch_scatter32(const void * in, void * out, const uint32_t * orig, uint64_t n, uint32_t w,
                     void * stream);
// fills the saved-column entries of stopped particles (columns saved after their stop step)
// with their final state: ring holds `cap` columns of n x w floats, columns [first, first +
// count) are completed; src[0..w-1] are the final values; off64 is added when local != 0
// p[i] = i for i < n (n < 2^32)
int launch_iota(uint32_t * p, uint64_t n, void * stream);
int launch_fill_frozen(float * ring, uint32_t cap, uint64_t first, uint32_t count, uint64_t n,
                       uint32_t w, const int32_t * status, const int32_t * abort_step,
                       const float * const * src, uint64_t save_interval, uint32_t local,
                       const double * off64, void * stream);
int launch_pack(const float * src, void * dst, uint64_t nodes, uint32_t ncomp, uint32_t cstride,
                uint32_t dtype, void * stream);
int launch_seed(const float * aos, float * const * soa, uint64_t n, uint32_t dim,
                const double * off64, uint32_t local, void * stream);
int launch_count_alive(const int32_t * status, uint64_t n, unsigned long long * out,
                       void * stream);
int launch_gather_positions(float * const * soa, float * aos, uint64_t n, uint32_t dim,
                            const double * off64, uint32_t local, void * stream);

} // namespace gr
