/*
 * granite_hip.h -- C ABI of the MI355X trajectory engine (libgranite_hip.so).
 *
 * This is the boundary between the pygranite-compatible host layer (the `pygranite` Python
 * module in pygranite_amd/, or any other C/C++/FFI caller) and the gfx950 HIP kernels. It
 * carries plain pointers, sizes and POD descriptors only; no exceptions cross it. Every
 * function returns 0 (GR_OK) on success or a negative GR_E* code; gr_last_error() returns the
 * calling thread's last message.
 *
 * The reference (pygranite 1.5.0) has no C ABI: its boundary is the pybind11 module
 * (source/pygranite.cpp:3-152) over one C++ class that owns CUDA textures and launches one
 * kernel per time step. Each entry point below names the reference code it replaces.
 *
 * Ownership: device memory is owned by the context (fields, jobs). Host pointers are borrowed
 * for the duration of the call. A context, and every field and job created from it, is used
 * by one submitting thread at a time.
 */
#ifndef GRANITE_HIP_H
#define GRANITE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GR_ABI_VERSION 2

/* error codes */
#define GR_OK 0
#define GR_E_INVALID -1   /* bad argument / unsupported combination */
#define GR_E_HIP -2       /* HIP runtime error */
#define GR_E_NOMEM -3     /* allocation failed */
#define GR_E_PARSE -4     /* expression did not parse / validate */
#define GR_E_STATE -5     /* call out of order (e.g. capacity exceeded) */

/* limits (source/granite/Constants.hpp:7-13, Language.hpp:75) */
#define GR_MAX_VOLUMES 8
#define GR_MAX_PROGRAMS 4
#define GR_MAX_CONSTANTS 4
#define GR_MAX_CODE 64
#define GR_MAX_STACK 8
/* steps of one gr_job_advance_noalloc call */
#define GR_NOALLOC_MAX_STEPS 16384
#define GR_ABORT_VALUE 1e-9f

/* storage types of field slices */
typedef enum { GR_F32 = 0, GR_F16 = 1, GR_BF16 = 2 } gr_dtype;

/* enum values match source/granite/IntegratorSettings.hpp:16-76 and AbortReason.hpp:8-15 */
typedef enum { GR_EULER = 0, GR_RK4 = 1 } gr_integrator;
typedef enum { GR_BLOCK = 0, GR_LOOP_X = 1, GR_LOOP_Y = 2, GR_LOOP_XY = 3 } gr_border;
typedef enum {
    GR_CURV_OFF = 0,
    GR_CURV_TOTAL = 1,
    GR_CURV_FAST_TOTAL = 2,
    GR_CURV_INDIVIDUAL_AND_TOTAL = 3
} gr_curvature;
typedef enum { GR_ABORT_TIME = 0, GR_ABORT_LENGTH = 1, GR_ABORT_FIT_LENGTH = 2 } gr_abort_mode;
typedef enum { GR_MODE_OFF = 0, GR_MODE_CONSTANT = 1, GR_MODE_DYNAMIC = 2 } gr_mode;
typedef enum {
    GR_STATUS_TIME = 0,
    GR_STATUS_LENGTH = 1,
    GR_STATUS_DOMAIN = 2,
    GR_STATUS_TOPOGRAPHY = 3,
    GR_STATUS_WIND = 4
} gr_status;
/* texture-weight emulation: exact FP32 weights, or weights quantised to 1/256 */
typedef enum { GR_WEIGHTS_EXACT = 0, GR_WEIGHTS_Q8_ROUND = 1, GR_WEIGHTS_Q8_TRUNC = 2 } gr_weights;

typedef struct gr_ctx gr_ctx;
typedef struct gr_field gr_field;
typedef struct gr_job gr_job;

/* ---- library / context ---------------------------------------------------------------- */

int gr_abi_version(void);
/* message of the calling thread's last failed call ("" if none) */
const char * gr_last_error(void);
/* number of visible HIP devices (0 without a GPU; never fails) */
int gr_device_count(int * count);

/* Creates a context on `device`: one compute stream, one copy stream.
 * Replaces the implicit device 0 + default stream of TrajectoryIntegrator.tpp:787-796. */
int gr_ctx_create(int device, gr_ctx ** out);
int gr_ctx_destroy(gr_ctx * ctx);
int gr_ctx_synchronize(gr_ctx * ctx);
/* Integrated particle-steps (steps a particle began alive, its aborting step included) of
 * every launch on this context since the previous call, then resets the count. Synchronizes
 * the compute stream. No reference counterpart: it is the work measure of throughput figures
 * (dead particles cost no sampling). */
int gr_ctx_take_work(gr_ctx * ctx, uint64_t * particle_steps);
/* Workgroup-steps of field-tile launches (GR_OPT_TILE) since the last call: with a tile
 * (*tiled) and without one, because the box of the workgroup's live cells did not fit
 * (*untiled); resets both. Synchronises the compute stream. No counterpart in the reference
 * (its stencils come from the texture unit, source/granite/TrajectoryIntegrator.icc:22-40). */
int gr_ctx_take_tile_stats(gr_ctx * ctx, uint64_t * tiled, uint64_t * untiled);
/* Summed device time (HIP events on the compute stream) and number of the step-kernel launches
 * (gr_job_advance, gr_batch_run) since the previous call, then resets both. Waits for the last
 * launch; launches are not serialised by the timing. No reference counterpart. */
int gr_ctx_take_kernel_ms(gr_ctx * ctx, double * ms, uint64_t * launches);
/* Device-to-device copy bandwidth of the context's GPU in GB/s (bytes read plus bytes written per
 * second), the best of `trials` copies of `bytes` (a multiple of 16) between two buffers by a
 * grid-stride streaming kernel: the measured counterpart of the published HBM peak. No
 * counterpart in the reference. Allocates and frees the two buffers. */
int gr_ctx_copy_bandwidth(gr_ctx * ctx, uint64_t bytes, uint32_t trials, double * gbs);
/* device name and HBM bytes of the context's GPU */
int gr_ctx_info(gr_ctx * ctx, char * name, size_t name_len, uint64_t * hbm_bytes,
                int * compute_units);

/* ---- fields (wind, additional volumes, topography) ---------------------------------------
 * A field holds `nslots` resident time slices of an nx*ny*nz grid with `ncomp` components,
 * stored node-interleaved in `dtype`. Replaces the CUDA texture objects and arrays of
 * TrajectoryIntegrator.cpp:444-609 (loadTexture3D / loadTexture2D) and the back/front/cache
 * rotation of TrajectoryIntegrator.tpp:894-902. Addressing: 0 = border (zero outside, the
 * reference's 3D textures), 1 = clamp to edge (the reference's 2D textures). */
typedef struct {
    uint32_t nx, ny, nz; /* nz = 1 for 2D */
    uint32_t ncomp;      /* 1..3 */
    uint32_t dtype;      /* gr_dtype */
    uint32_t nslots;     /* resident slices, >= 1 */
    uint32_t address;    /* 0 border, 1 clamp */
    uint32_t reserved;
} gr_field_desc;

int gr_field_create(gr_ctx * ctx, const gr_field_desc * desc, gr_field ** out);
int gr_field_destroy(gr_field * field);
/* Uploads one slice: comps[c] points to nz*ny*nx host floats (x fastest). The host data is
 * staged into pinned memory before return; the H2D copy and the on-device FP32 -> dtype
 * conversion run on the copy stream. With async = 0 the call returns after the slot is
 * resident. Kernels that read the slot wait for it on the device. */
int gr_field_upload(gr_field * field, uint32_t slot, const float * const * comps, int async);
/* Same from device memory (comps[c] are device pointers, FP32), copied on the copy stream: the
 * work that writes them must be complete, and they must stay unchanged until gr_field_wait. */
int gr_field_upload_device(gr_field * field, uint32_t slot, const float * const * dcomps);
/* Blocks until the slot's last upload finished. */
int gr_field_wait(gr_field * field, uint32_t slot);
int gr_field_bytes(const gr_field * field, uint64_t * bytes_per_slot);
/* Stretched vertical axis (3D): node k lies at height levels[k] relative to Offset.Z
 * (strictly increasing, n = nz). Sampling then maps a height to k + (z - z_k)/(z_{k+1} - z_k).
 * Jobs on such a field set domain[2] to the top level so that leaving [0, top] terminates the
 * particle (Domain). No counterpart in the reference (uniform GridScale only, icc:112-120). */
int gr_field_set_z_levels(gr_field * field, const float * levels, uint32_t n);

/* ---- expressions -------------------------------------------------------------------------
 * Compiles one AdditionalCompute program (grammar of TrajectoryIntegrator.cpp:10-26,
 * identifier binding cpp:744-768, validation tpp:653-697) into postfix bytecode for the
 * wavefront-uniform interpreter fused into the step kernel (replaces flatten cpp:770-798 and
 * the recursive device interpret hpp:539-627). `names` are the loader's constant keys
 * (bound to C0..C3) and volume keys (bound to F0..F7). */
typedef struct {
    uint32_t n;                  /* number of instructions */
    uint32_t depth;              /* maximum stack depth */
    uint32_t code[GR_MAX_CODE];  /* op | (arg << 8) */
    float imm[GR_MAX_CODE];      /* immediates referenced by OP_CONST */
} gr_program;

#define GR_EXPR_ALLOW_ABS 1u /* accept abs(...) (extension) */
typedef struct {
    uint32_t dim;           /* 2 or 3: z and w are rejected in 2D */
    uint32_t curvature;     /* cv needs GR_CURV_INDIVIDUAL_AND_TOTAL */
    uint32_t nconst;        /* constant names, bound to C0.. */
    uint32_t nvol;          /* volume names, bound to F0.. */
    const char * const * const_names;
    const char * const * vol_names;
    uint32_t flags;
} gr_expr_env;

int gr_expr_compile(const char * src, const gr_expr_env * env, gr_program * out);
/* CPU evaluation of compiled bytecode on a 19-slot reference vector (for checking). */
int gr_expr_eval_host(const gr_program * prog, const float * ref19, float * out);
/* CPU evaluation of the device form the step kernels run (leaf operands folded into their
 * operations, the top two stack values in registers), for checking the lowering: the result
 * equals gr_expr_eval_host's bit for bit. No counterpart in the reference (its interpreter
 * walks the flattened tree, TrajectoryIntegrator.tpp:379-447). */
int gr_expr_eval_device_form(const gr_program * prog, const float * ref19, float * out);

/* ---- jobs ----------------------------------------------------------------------------------
 * A job is one particle set advanced through windows of a field. It replaces compute_ /
 * computeLoop_ / computeLoopGPU_ (TrajectoryIntegrator.tpp:467-995) and the SimulationData
 * kernel argument (SimulationData.hpp:21-70). Particle state stays resident in HBM between
 * calls. */
typedef struct {
    uint32_t dim;            /* 2 or 3 */
    uint32_t integrator;     /* gr_integrator */
    uint32_t border;         /* gr_border (2D only; 3D is always Block) */
    uint32_t curvature;      /* gr_curvature */
    uint32_t abort_mode;     /* gr_abort_mode */
    uint32_t uplift_mode;    /* gr_mode (3D only) */
    uint32_t windfield_mode; /* 0 Constant, 1 Dynamic */
    uint32_t constants_mode; /* gr_mode */
    uint32_t volume_mode;    /* gr_mode */
    uint32_t reverse;        /* Reverse */
    uint32_t weights;        /* gr_weights */
    uint32_t time64;         /* accumulate simulation time in FP64 (precision form) */
    uint32_t kahan;          /* Kahan-compensated positions (precision form) */
    uint32_t local_coords;   /* keep positions relative to `offset` (FP64 geo-origin) */
    uint32_t nvol;           /* additional volumes (<= 8) */
    uint32_t nconst;         /* constants (<= 4) */
    float dt;                /* DeltaT */
    float dtd;               /* DataTimeDistance */
    float sphere_radius;     /* > 0 enables sphere coordinates (2D) */
    float max_length;        /* MaximumLength */
    float grid_scale[3];     /* expanded GridScale (tpp:521-551) */
    float domain[3];         /* grid * GridScale (tpp:526-546) */
    double offset[3];        /* expanded Offset (tpp:553-574) */
    uint64_t save_interval;  /* SaveInterval (0 = only last) */
    uint32_t nprog;          /* number of programs */
    uint32_t reserved;
    gr_program prog[GR_MAX_PROGRAMS];
} gr_job_desc;

/* One window of the host schedule: the slots to read and the time stamps
 * (TrajectoryIntegrator.tpp:933-969), and the steps to run (tpp:973-994). */
typedef struct {
    gr_field * wind;          /* required */
    uint32_t wind_back, wind_front;
    gr_field * volumes;       /* additional volumes as one field with nvol components, or NULL */
    uint32_t vol_back, vol_front;
    gr_field * topography;    /* 2D scalar field (clamp), or NULL */
    uint32_t uplift_back, uplift_front;   /* particle-data slots */
    uint32_t const_back, const_front;     /* particle-data slots */
    float time_back, time_front;          /* COUNTER*DTD, (COUNTER+1)*DTD in FP32 */
    double time_back64, time_front64;     /* same in FP64 (time64 jobs) */
    float time_start;                     /* data.Time at the first step */
    double time_start64;
    uint64_t counter_start;               /* _simulation_counter at the first step */
    uint32_t nsteps;                      /* steps to run */
    uint32_t reserved;
} gr_window;

/* output kinds for gr_job_read / gr_job_read_saved */
typedef enum {
    GR_OUT_POSITIONS = 0,       /* N x dim float (global coordinates) */
    GR_OUT_STATUS = 1,          /* N int32 (gr_status) */
    GR_OUT_TOTAL_CURVATURE = 2, /* N float */
    GR_OUT_LENGTHS = 3,         /* N float */
    GR_OUT_CURVATURE = 4,       /* N float (per-step curvature) */
    GR_OUT_ABORT_STEP = 5,      /* N int32: step counter at which the status left Time, or -1 */
    GR_OUT_EXPR0 = 8,           /* + e: N float, latest value of program e */
    GR_OUT_VOLUME0 = 16         /* + k: N float, latest value of volume k */
} gr_output;

/* Creates a job for `n` particles; `seeds` are n*dim floats (AoS, the last row of the
 * reference's TrajectorySet, cpp:424-442) in host memory or in device memory of the context's
 * GPU (then copied on the compute stream, in order, without waiting: the work that writes them
 * must be complete, and they must stay unchanged until the job's first launch or
 * gr_ctx_synchronize). `save_capacity` = saved
 * columns held on the device between reads. The state is initialised on the compute stream: a
 * call does not wait for launches already queued there. */
int gr_job_create(gr_ctx * ctx, const gr_job_desc * desc, const float * seeds, uint64_t n,
                  uint32_t save_capacity, gr_job ** out);
int gr_job_destroy(gr_job * job);
/* Device bytes gr_job_create allocates for a job of `desc` with n particles and save_capacity
 * saved columns, plus (with_reorder != 0) what its first reorder adds (index map and shadow
 * copies). For admission by memory budget; no counterpart in the reference. */
int gr_job_bytes(const gr_job_desc * desc, uint64_t n, uint32_t save_capacity, int with_reorder,
                 uint64_t * bytes);
/* Per-particle data with slots: kind 0 = uplift (n*dim floats), 1 + k = constant k
 * (n floats). Replaces the uplift/constant buffers of TrajectoryIntegrator.cpp:195-245 and
 * 305-359. */
int gr_job_set_particle_data(gr_job * job, uint32_t kind, uint32_t slot, const float * host);
/* Runs w->nsteps steps on the compute stream (asynchronously). Saved columns (steps whose
 * counter % save_interval == 0) go to the device column ring. */
int gr_job_advance(gr_job * job, const gr_window * w);
/* gr_job_advance without any allocation, for callers that cannot allocate device or host
 * memory: code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc. No
 * counterpart in the reference (one job per TrajectoryIntegrator, allocations in compute_,
 * TrajectoryIntegrator.tpp:467-715). Same window, same results as gr_job_advance; the launch's
 * time table is formed on the device in a fixed per-GPU scratch array of GR_NOALLOC_MAX_STEPS
 * steps (w->nsteps beyond it: GR_E_INVALID), the time-paired copy of the window's slots is used
 * only if it is already resident (otherwise the slots are sampled separately, with identical
 * results), and the launch is not timed (gr_job_last_advance_ms, gr_ctx_take_kernel_ms do not
 * see it). Every call on one GPU shares the scratch array: calls queued on one context are
 * ordered by its stream; calls that overlap on one GPU from several contexts race. The job's
 * state, ring and the fields are allocated beforehand (gr_job_create, gr_field_create). */
int gr_job_advance_noalloc(gr_job * job, const gr_window * w);
/* Reads the current per-particle state (blocking). */
int gr_job_read(gr_job * job, uint32_t kind, void * dst, uint64_t bytes);
/* Copies saved columns [first, first + count) of `kind` (positions, curvature, expressions,
 * volumes) to host memory laid out [count][N(][dim)] (blocking). Columns are numbered by
 * save index since job creation; a column stays readable until `save_capacity` newer
 * columns were written. */
int gr_job_read_saved(gr_job * job, uint32_t kind, uint64_t first, uint64_t count, void * dst);
/* gr_job_read_saved without waiting: queues the completion of the columns, their return to
 * the caller's particle order and the device-to-host copy (into pinned staging, on the copy
 * stream) behind the launches that wrote them, and returns a ticket. dst_cols[c] receives
 * column first + c (N(*dim) floats) when gr_job_read_wait(job, ticket) returns; the pointers
 * must stay valid until then. Later launches of the job wait only for copies of ring columns
 * they would overwrite, so a caller that reads each chunk of saves while the next chunk
 * computes overlaps the copies with integration (the ring must hold both chunks). Replaces the
 * synchronous per-save copies of TrajectoryIntegrator.tpp:997-1049 (copyParticleInfo). */
int gr_job_read_saved_async(gr_job * job, uint32_t kind, uint64_t first, uint64_t count,
                            void * const * dst_cols, uint64_t * ticket);
/* Blocks until the read `ticket` is complete and its columns are in the caller's memory. */
int gr_job_read_wait(gr_job * job, uint64_t ticket);
/* Number of particles still at status Time (replaces count_if, ToolKernel.tpp:4-25). */
int gr_job_count_alive(gr_job * job, uint64_t * alive);
/* Live range without a wait: every particle slot from *range on holds a stopped particle, and
 * step launches cover only [0, *range). A reorder places the alive particles first and counts
 * them; the count is taken up once that reorder has completed on the device (without waiting
 * here, or at the job's next reorder, which waits for it). Equals the particle count until
 * then. No counterpart in the reference (its launches always cover every particle,
 * TrajectoryIntegrator.tpp:982). */
int gr_job_live_range(gr_job * job, uint64_t * range);
/* Number of saved columns written so far. */
int gr_job_saved_columns(gr_job * job, uint64_t * cols);
/* Device time of the last gr_job_advance, measured with HIP events on the compute stream. */
int gr_job_last_advance_ms(gr_job * job, float * ms);
/* Block size used for the job's step kernel (0 = automatic). */
int gr_job_set_launch(gr_job * job, uint32_t block);

/* Job options. GR_OPT_BLOCK: as gr_job_set_launch. GR_OPT_PAIRED (default 1): sample dynamic
 * wind from a time-paired copy of the (back, front) slots -- one vector load per stencil corner
 * for both slices; 0 samples the two slots separately. */
#define GR_OPT_BLOCK 1u
#define GR_OPT_PAIRED 2u
/* GR_OPT_CU_WAVES (0 .. 32, default 0): cap on the step kernel's wavefronts per compute unit,
 * applied by reserving LDS per workgroup; 0 = as many as the registers allow. Fewer wavefronts
 * per CU keep their stencil gathers inside the CU's cache. Results do not depend on it. */
#define GR_OPT_CU_WAVES 3u
/* GR_OPT_SPECIALIZE (default 1): launches whose features are the plain set (exact sampler
 * weights, uniform z, no Kahan, no reverse, global coordinates) with curvature Off or
 * IndividualAndTotal run kernels built for that set; 0 runs the generic kernels, which read every
 * flag at run time. Results are identical either way. */
#define GR_OPT_SPECIALIZE 4u
/* GR_OPT_REUSE (default 1): time-paired launches run the kernel that keeps each lane's last
 * stencil and reloads it only when the lane's cell changes (fewer gathers; 24 or 48 more VGPRs).
 * 0: only launches whose wavefronts all fit on the device at once do. Results do not depend on
 * it. */
#define GR_OPT_REUSE 5u
/* GR_OPT_QUAD (0 .. 2^29 - 1, default 128 x compute units): launches of a job whose live range
 * holds at most this many particles, in the time-paired layout (a copy smaller than 4 GiB) with a
 * plain feature set, run four lanes per particle, each lane interpolating one velocity component
 * (a quarter of the lerps per lane: a shorter step chain for launches of few particles); 0:
 * never. In gr_batch_run the count is the live ranges of all the launch's spans together.
 * Results do not depend on it. */
#define GR_OPT_QUAD 6u
/* GR_OPT_TILE (0 or 1024 .. 65536 bytes, default 0): 3D time-paired launches (not quad ones) run
 * the kernel in which every workgroup, each step, stages in LDS the box of time-paired nodes
 * around its live particles' cells (this many bytes of LDS per workgroup, beyond the
 * interpreter's) and samples the stencils inside it from there; stencils outside the box, and
 * steps whose box does not fit, load from the copy. 0: never. GR_OPT_TILE_HALO (0 .. 8, default
 * 1): cells added to the box on every side. Results do not depend on either. */
#define GR_OPT_TILE 7u
#define GR_OPT_TILE_HALO 8u
int gr_job_set_option(gr_job * job, uint32_t option, int64_t value);

/* Reorders the resident particle state by the Morton code of each particle's current cell
 * (stopped particles last), so that neighbouring lanes gather neighbouring stencils, and counts
 * the alive particles it places first (see gr_job_live_range). Results stay addressed by the
 * caller's particle index: gr_job_read and gr_job_read_saved return them in the original order.
 * Saved columns written before the call must be read before it. Waits until the job's previous
 * reorder has run on the device (so a caller queueing many launches stays at most one reorder
 * ahead). No counterpart in the reference (its particle order is fixed). */
int gr_job_reorder(gr_job * job);
/* gr_job_reorder for several jobs of one context in one pass (one counting sort by cell over
 * all their particles, one gather launch for all their state arrays). */
int gr_jobs_reorder(gr_job * const * jobs, uint32_t njobs);

/* ---- batches: many jobs in one persistent launch -------------------------------------------
 * Each span advances one job through consecutive windows of a wind field whose slices are all
 * resident (slot s holds slice s). Window edges follow the reference's host schedule
 * (FP32 time, max_windfield_time += DataTimeDistance per window, TrajectoryIntegrator.tpp:
 * 852-994); a span ends after max_steps steps, when the step counter reaches
 * max_sim_time / dt, or when no further slice exists (the loader's step() == false). Jobs of
 * different dimension, integrator and storage type run in the same launch: every wavefront
 * pulls (job, 64 particles) items from one queue per XCD (items split by their position in the
 * job's particle order, longest spans first inside a queue) and moves on to the other queues
 * once its own is drained. Jobs must use the
 * lean feature set (no uplift, constants, volumes, topography, loop borders, sphere or length
 * aborts). No counterpart in the reference (one job per TrajectoryIntegrator). */
typedef struct {
    gr_job * job;
    gr_field * wind;
    uint32_t first_window;  /* COUNTER of the first window */
    uint32_t max_steps;
    float time_start;       /* data.Time at the first step */
    float max_sim_time;     /* MaximumSimulationTime (<= 0: no limit) */
    double time_start64;
    uint64_t counter_start; /* _simulation_counter at the first step */
} gr_span;

/* Where a span stopped: the next span of the same job starts from these values. */
typedef struct {
    uint32_t steps;     /* steps run */
    uint32_t window;    /* window COUNTER of the next step */
    float time;         /* data.Time of the next step */
    uint32_t done;      /* 1: no further step is possible (slices exhausted or the step
                           counter reached max_sim_time / dt) */
    double time64;
    uint64_t counter;   /* _simulation_counter of the next step */
} gr_span_end;

/* Runs the spans; ends[i] (may be NULL) receives where span i stopped. `block` = 0 picks the
 * context's batch block size (GR_CTX_BATCH_BLOCK). The call returns after queueing; *ms (may be
 * NULL) receives the device time of the launch and makes the call wait for it. */
int gr_batch_run(gr_ctx * ctx, const gr_span * spans, uint32_t nspans, uint32_t block,
                 gr_span_end * ends, float * ms);

/* Context options of the batched kernel (no counterpart in the reference). Results do not
 * depend on them, only speed.
 * GR_CTX_BATCH_BLOCK: workgroup size when gr_batch_run gets block = 0 (64, 128, 192 or 256;
 *   default 256).
 * GR_CTX_BATCH_WAVES: register budget of the batched kernel: 0 = the compiler's allocation
 *   (default), 4 = at most 128 VGPRs so that 4 wavefronts fit per SIMD.
 * GR_CTX_BATCH_GROUP: compute units that share one work queue (1 .. number of CUs; default:
 *   the CUs of one XCD). Each queue holds one compact region of the particle sets.
 * GR_CTX_BATCH_CU_WAVES: cap on the batched kernel's wavefronts per compute unit (0 .. 32;
 *   default 0 = as many as the registers allow), applied by reserving LDS per workgroup.
 * GR_CTX_BATCH_REUSE: 1 (default) = every batched launch uses the build that keeps each lane's
 *   last time-paired stencil across RK4 stages (fewer gathers, more registers); 0 = only launches
 *   whose wavefronts all fit on the device at once.
 * GR_CTX_BATCH_QUAD_ITEMS (read-only): items of the last gr_batch_run that ran a quad of lanes
 *   per particle. A launch whose spans hold at most a job's quad limit (GR_OPT_QUAD) of live
 *   particles in all runs the jobs that have a specialised feature set in the time-paired
 *   layout this way, in items of 16 particles, with the stencil-reuse build.
 * GR_CTX_BATCH_UNIFORM: 1 (default) = a stencil-reuse launch whose spans all share one
 *   specialised variant (dimension, integrator, storage type, feature set) runs that variant's
 *   own kernel, held to four wavefronts per SIMD; 0 = always the kernel of every variant (whose
 *   register demand is the largest variant's).
 * GR_CTX_BATCH_LAST_KERNEL (read-only): build of the last gr_batch_run launch: 0 none (no
 *   items), 1 the default build, 2 the 128-VGPR build, 3 the stencil-reuse build, 4 a
 *   per-variant kernel. */
#define GR_CTX_BATCH_BLOCK 1u
#define GR_CTX_BATCH_WAVES 2u
#define GR_CTX_BATCH_GROUP 3u
#define GR_CTX_BATCH_CU_WAVES 4u
#define GR_CTX_BATCH_REUSE 5u
#define GR_CTX_BATCH_QUAD_ITEMS 6u
#define GR_CTX_BATCH_UNIFORM 7u
#define GR_CTX_BATCH_LAST_KERNEL 8u
/* GR_CTX_POOL_BYTES (>= 0, default 8 GiB): device memory of destroyed jobs' arrays that the
 * context keeps for later jobs (by size class: powers of two from 4 KiB below 2 MiB, 2 MiB
 * granules above), so that creating and destroying jobs does not call hipMalloc / hipFree (which
 * waits for the device) each time; 0 frees at once. A failing allocation empties the cache and
 * retries; gr_ctx_destroy frees it. GR_CTX_POOL_CACHED (read-only): bytes cached now. */
#define GR_CTX_POOL_BYTES 9u
#define GR_CTX_POOL_CACHED 10u
int gr_ctx_set_option(gr_ctx * ctx, uint32_t option, int64_t value);
int gr_ctx_get_option(gr_ctx * ctx, uint32_t option, int64_t * value);

/* Autotuner of the batched kernel (no counterpart in the reference). Times every candidate
 * launch configuration (workgroup size 64 / 128 / 256 x compute units per queue: one XCD, 8 x
 * wavefronts per CU: register-limited, 4, 6, 8) on njobs probe jobs of `desc` with n seeds each
 * spread over the domain, advanced `steps` steps through `wind` from window 0 (best of `trials`
 * runs each, all from identical state), and sets the fastest as the context's batch options. results[0 ..
 * *nresults) receive every candidate's device time (may be NULL). The context's work counter
 * and kernel-time sums are unchanged. */
typedef struct {
    uint32_t block;
    uint32_t waves;
    uint32_t group;
    float ms;
    uint32_t cu_waves;
} gr_tune_entry;
int gr_batch_autotune(gr_ctx * ctx, gr_field * wind, const gr_job_desc * desc, uint32_t n,
                      uint32_t njobs, uint32_t steps, uint32_t trials, gr_tune_entry * results,
                      uint32_t max_results, uint32_t * nresults);

/* Autotuner of single-job launches (gr_job_advance; no counterpart in the reference, which
 * launches one fixed configuration, TrajectoryIntegrator.tpp:790-796). Times every candidate --
 * workgroup size 64 / 128 / 256 x lanes per particle 1 / 4 (GR_OPT_QUAD) x LDS field tile
 * 0 / 16 / 32 KiB (GR_OPT_TILE; 3D time-paired launches with one lane per particle) -- on a
 * probe job of `job`'s description with min(n, 2^20) seeds spread over the domain (regrouped by
 * cell), advanced `steps` steps through `wind`'s slots 0 and 1 from time 0 (best of `trials`
 * runs each, all from identical state), and sets the fastest candidate's options on `job`:
 * its block size, its tile, and, when four lanes per particle won, a quad limit covering the
 * whole job (one lane keeps the job's limit, which serves its last few live particles). The
 * job's particle state is not touched; results do not depend on the options. Jobs with
 * uplift, constants or additional volumes are refused (the probe has no per-particle data).
 * results[0 .. *nresults) receive every candidate's device time (may be NULL). */
typedef struct {
    uint32_t block;
    uint32_t lanes;      /* lanes per particle: 1 or 4 */
    uint32_t tile_bytes; /* 0: no field tile */
    float ms;
} gr_job_tune_entry;
int gr_job_autotune(gr_job * job, gr_field * wind, uint32_t steps, uint32_t trials,
                    gr_job_tune_entry * results, uint32_t max_results, uint32_t * nresults);

#ifdef __cplusplus
}
#endif

#endif /* GRANITE_HIP_H */
