// Host engine behind the C ABI (include/granite_hip.h): contexts (device + streams), fields
// (resident time slices in node-interleaved FP32/FP16/BF16), and jobs (resident particle state
// advanced window by window by the gfx950 step kernel).
//
// Replaces the reference's host data plane and driver: texture/array management and
// synchronous uploads (source/granite/TrajectoryIntegrator.cpp:146-360, 444-609), buffer
// setup (TrajectoryIntegrator.tpp:467-715), the per-step launch + cudaDeviceSynchronize loop
// (tpp:973-994) and the per-save D2H (tpp:997-1049). Uploads run on a copy stream and are
// ordered against kernels with events, so the next slice streams in while a window computes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "gr_internal.h"

namespace
{
thread_local std::string g_err;

int set_err(int code, const char * fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define GR_HIP(call)                                                                              \
    do                                                                                            \
    {                                                                                             \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return set_err(GR_E_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),       \
                           __FILE__, __LINE__);                                                   \
    } while (0)

template <class T> int dalloc(T ** p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) return GR_OK;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(p), bytes);
    if (e != hipSuccess)
    {
        *p = nullptr;
        (void)hipGetLastError();
        return set_err(GR_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    return GR_OK;
}

void dfree(void * p)
{
    if (p) (void)hipFree(p);
}

size_t dtype_size(uint32_t t) { return t == GR_F32 ? 4 : 2; }

// Events that only time device work: recorded without the system-scope fence (cache writeback
// and invalidation) that a default event performs around every launch it brackets.
constexpr unsigned TIMING_EVENT = hipEventDisableSystemFence;
} // namespace

extern "C" void gr_set_error(const char * msg) { g_err = msg ? msg : ""; }

struct gr_ctx
{
    int device{0};
    hipStream_t compute{nullptr};
    hipStream_t copy{nullptr};
    unsigned long long * dcount{nullptr}; // alive count scratch
    unsigned long long * dwork{nullptr};  // integrated particle-steps, WORK_SLOTS x WORK_STRIDE
    unsigned long long * dtile{nullptr};  // field-tile workgroup-steps (with, without), same slots
    uint32_t * d_culut{nullptr};          // cu_key -> queue of the batched kernel
    uint32_t ncu{1};                      // compute units found (dense indices in d_culut)
    // persistent batched launches
    std::vector<gr::BatchItem> h_items;  // the last item list (reused while the spans match)
    uint64_t items_quad{0};               // its quad items
    uint32_t batch_last_kernel{0};        // build of the last batched launch (GR_CTX_BATCH_LAST_KERNEL)
    std::vector<uint32_t> qstart;         // its queue starts
    std::vector<uint64_t> items_key;
    bool items_valid{false};
    gr::BatchItem * d_items{nullptr};
    size_t items_cap{0};
    struct Stage
    {
        void * h{nullptr};  // pinned, mapped into the device's address space
        void * dh{nullptr}; // the device's address of h
        size_t cap{0};
        hipEvent_t done{nullptr};
        bool pending{false};
    };
    static constexpr int NSTAGE = 4;
    Stage stage[NSTAGE];
    int stage_i{0};
    char * d_stage{nullptr}; // device copy of the staging block (queues, params, window tables)
    size_t d_stage_cap{0};
    // reorders: keys / values of all particles of the jobs being reordered, sort temp storage,
    // job and gather tables
    uint64_t * skeys[2]{};
    uint32_t * svals[2]{};
    size_t scap{0};
    void * stmp{nullptr};
    size_t stmp_bytes{0};
    gr::SortJob * d_sjobs{nullptr};
    size_t sjobs_cap{0};
    uint32_t * bcount{nullptr}; // bucket counts / starts of the reorder's counting sort
    uint32_t * bstart{nullptr};
    size_t bcap{0};
    gr::GatherEntry * d_gtab{nullptr};
    size_t gtab_cap{0};
    hipEvent_t bev0{nullptr}, bev1{nullptr};
    int cus{0};
    uint32_t lds_per_cu{0}; // LDS bytes of one compute unit (occupancy caps)
    uint32_t nxcd{1}; // XCDs of the device
    uint32_t batch_block{256};
    uint32_t batch_waves{0};
    uint32_t batch_group{32}; // compute units per queue of the batched kernel
    uint32_t batch_cu_waves{0}; // cap on the batched kernel's wavefronts per CU (0: none)
    uint32_t batch_reuse{1};    // stencil-reuse build: 1 always (default), 0 for narrow launches only
    uint32_t batch_uniform{1};  // per-variant kernels for narrow launches of one variant (default 1)
    // device time of every step-kernel launch (event pairs, summed by gr_ctx_take_kernel_ms)
    std::vector<hipEvent_t> kev;
    size_t kpending{0};
    double kms_acc{0.0};
    uint64_t klaunch_acc{0};
    // alive counts of reordered jobs: the count kernel of a reorder stores each job's count into
    // its slot of host-coherent pinned memory; reorder number k records alive_ev[k % ALIVE_EVENTS]
    // after it, so a job whose count is pending knows which event to query
    static constexpr uint32_t ALIVE_EVENTS = 8;
    std::vector<uint32_t *> alive_pages; // 4 KiB pages of count slots
    std::vector<uint32_t *> alive_free;  // free slots (host pointers)
    hipEvent_t alive_ev[ALIVE_EVENTS]{};
    uint64_t alive_seq{0};  // reorders that stored counts
    uint64_t alive_done{0}; // all of the first alive_done of them are known complete
    // live fields and jobs: the context outlives them (gr_ctx_destroy refuses otherwise)
    int children{0};
    // end of the last no-allocation launch (created with the context, recorded by each such
    // launch): later uploads into the slots it read wait for it
    hipEvent_t noalloc_done{nullptr};
    // Cache of released job arrays (state, shadows, index maps, rings), keyed by size class:
    // a job created after another's release reuses its device memory instead of calling
    // hipMalloc / hipFree (which waits for the device) again. At most pool_cap bytes are kept
    // (GR_CTX_POOL_BYTES); a failing allocation empties the cache and retries.
    std::multimap<size_t, void *> pool;
    size_t pool_cached{0};
    size_t pool_cap{size_t(8) << 30};
};

struct gr_field
{
    gr_ctx * ctx{nullptr};
    gr_field_desc d{};
    uint32_t cstride{0};
    uint64_t nodes{0};
    uint64_t slot_bytes{0};
    std::vector<void *> slots;
    float * dstage{nullptr};   // FP32 component-major staging on the device
    float * hstage{nullptr};   // pinned host staging
    hipEvent_t stage_done{nullptr};
    bool stage_pending{false};
    std::vector<hipEvent_t> ready, inuse;
    std::vector<char> has_ready, has_inuse;
    std::vector<char> waited; // the compute stream already waits for the slot's current ready event
    // time-paired copies of (back, front) slot pairs, built on the compute stream on demand
    struct Pair
    {
        void * buf{nullptr};
        uint32_t b{0}, f{0};
        bool valid{false};
        uint64_t used{0};
    };
    std::vector<Pair> pairs; // at most max(2, nslots - 1): every window of a resident dataset
    uint64_t pair_clock{0};
    // stretched z axis
    float * zlev{nullptr};
    int32_t * zlut{nullptr};
    uint32_t nlut{0};
    float zlut_z0{0.f}, zlut_scale{0.f};
};

// serial numbers of jobs: a job created where a closed one lived gets a new number, so caches
// keyed by job (the batched launch's item list) never mistake one for the other
static std::atomic<uint64_t> g_job_serial{0};

struct gr_job
{
    gr_ctx * ctx{nullptr};
    uint64_t serial{++g_job_serial};
    gr_job_desc d{};
    uint64_t n{0};
    uint32_t full{0};
    uint32_t block{0};
    // interpreter stack columns in LDS: the deepest program's stack depth - 2 (the two top values
    // live in registers), from the validated bytecode
    uint32_t expr_cols{0};
    uint32_t cu_waves{0}; // cap on the step kernel's wavefronts per CU (0: none)
    uint32_t specialize{1}; // use the compile-time feature sets when they apply
    uint32_t reuse{1};      // stencil-reuse build for every launch (0: narrow launches only)
    uint64_t quad_max{0};   // launches of at most this many live particles run a quad per particle
    uint32_t tile_bytes{0}; // LDS field tile per workgroup of 3D time-paired launches (0: none)
    uint32_t tile_halo{1};  // cells added to the tile's box on every side
    float * p[3]{};
    float * q[3]{};
    float * c[3]{};
    float * ld[3]{};
    int32_t * status{nullptr};
    int32_t * abort_step{nullptr};
    float * total{nullptr};
    float * curv{nullptr};
    float * len{nullptr};
    float * expr[GR_MAX_PROGRAMS]{};
    float * vol[GR_MAX_VOLUMES]{};
    uint32_t cap{0};
    float * ring_pos{nullptr};
    float * ring_curv{nullptr};
    float * ring_expr{nullptr};
    float * ring_vol{nullptr};
    float * scratch_aos{nullptr};
    float * uplift[3]{};
    float * cst[3][GR_MAX_CONSTANTS]{};
    uint64_t saved{0};
    hipEvent_t ev0{nullptr}, ev1{nullptr};
    bool timed{false};
    // layout / reordering
    uint32_t paired{1};          // use the time-paired layout when the wind is dynamic
    uint32_t grid[3]{0, 0, 0};   // cells per axis of the last wind field (Morton keys)
    uint32_t * orig{nullptr};    // caller index of each state slot
    std::vector<void *> shadow;  // second copy of each state array (reorders gather into it)
    bool reordered{false};
    uint64_t reorder_col0{0};    // first saved column written in the current order
    // Every particle from run_n on is stopped: a reorder places the alive particles first and
    // counts them, and stopped particles never run again, so the count of any completed reorder
    // bounds the launches queued after it. Launches cover [0, run_n) only.
    uint64_t run_n{0};
    uint32_t * h_alive{nullptr}; // count slot (host-coherent pinned memory)
    uint32_t * d_alive{nullptr}; // its device address
    uint64_t alive_seq{0};       // number of the reorder whose count is pending (0: none)
    float * scratch2{nullptr};   // n * dim floats
    // Reads of saved columns in flight (gr_job_read_saved_async) and finished ones kept for
    // their buffers: pinned host memory the copy stream fills, a device staging copy in the
    // caller's particle order (reordered jobs), and the events that order them.
    struct Read
    {
        uint64_t ticket{0};
        bool from_ring{false};        // the copy reads ring columns [first, first + count)
        uint64_t first{0}, count{0};
        hipEvent_t ready{nullptr};    // the columns are complete (compute stream)
        hipEvent_t done{nullptr};     // the device-to-host copy has finished (copy stream)
        void * host{nullptr};
        size_t host_cap{0};
        void * dstage{nullptr};
        size_t dstage_cap{0};
        size_t col_bytes{0};
        std::vector<void *> dst;
    };
    std::vector<Read> reads;
    std::vector<Read> read_pool;
    uint64_t next_ticket{1};
};

// Size class of a cached job array: powers of two from 4 KiB below 2 MiB (the runtime packs
// such allocations), whole 2 MiB granules above.
static size_t pool_class(size_t bytes)
{
    const size_t G = size_t(2) << 20;
    if (bytes >= G) return (bytes + G - 1) / G * G;
    size_t c = 4096;
    while (c < bytes) c <<= 1;
    return c;
}

static void pool_flush(gr_ctx * c)
{
    for (auto & e : c->pool) (void)hipFree(e.second);
    c->pool.clear();
    c->pool_cached = 0;
}

// Device memory of `bytes` (0: none) for a job array: a cached block of its size class, or a new
// one. Callers fill what must start with defined contents.
template <class T> static int pool_alloc(gr_ctx * c, T ** p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) return GR_OK;
    const size_t cls = pool_class(bytes);
    auto it = c->pool.find(cls);
    if (it != c->pool.end())
    {
        *p = static_cast<T *>(it->second);
        c->pool.erase(it);
        c->pool_cached -= cls;
        return GR_OK;
    }
    void * q = nullptr;
    hipError_t e = hipMalloc(&q, cls);
    if (e != hipSuccess && !c->pool.empty())
    {
        (void)hipGetLastError();
        pool_flush(c);
        e = hipMalloc(&q, cls);
    }
    if (e != hipSuccess)
    {
        (void)hipGetLastError();
        return set_err(GR_E_NOMEM, "hipMalloc(%zu) failed: %s", cls, hipGetErrorString(e));
    }
    *p = static_cast<T *>(q);
    return GR_OK;
}

// Releases a job array that no queued work uses any more (the caller has drained the streams).
static void pool_free(gr_ctx * c, void * p, size_t bytes)
{
    if (!p) return;
    const size_t cls = pool_class(bytes);
    if (c && c->pool_cached + cls <= c->pool_cap)
    {
        c->pool.emplace(cls, p);
        c->pool_cached += cls;
    }
    else
        (void)hipFree(p);
}

// --------------------------------------------------------------------------------------------
// library / context
// --------------------------------------------------------------------------------------------
extern "C" int gr_abi_version(void) { return GR_ABI_VERSION; }

extern "C" const char * gr_last_error(void) { return g_err.c_str(); }

extern "C" int gr_device_count(int * count)
{
    if (!count) return set_err(GR_E_INVALID, "gr_device_count: null");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
    {
        (void)hipGetLastError();
        n = 0;
    }
    *count = n;
    return GR_OK;
}

extern "C" int gr_ctx_destroy(gr_ctx * c);

// Which compute units run workgroups: a probe kernel records cu_key() for many small workgroups;
// every key seen gets a dense index (ordered by XCC, shader engine, CU), which becomes the CU's
// home queue in the batched kernel. CUs never seen map to no queue (their wavefronts fall back to
// workgroup index modulo the queue count).
static int map_compute_units(gr_ctx * c)
{
    const uint32_t nprobe = 16384;
    const size_t nkeys = size_t(1) << gr::CU_KEY_BITS;
    uint32_t * d = nullptr;
    int rc = dalloc(&d, nprobe * sizeof(uint32_t));
    if (rc != GR_OK) return rc;
    std::vector<uint32_t> h(nprobe);
    if (gr::launch_cu_probe(d, nprobe, c->compute) != 0 ||
        hipMemcpyAsync(h.data(), d, nprobe * sizeof(uint32_t), hipMemcpyDeviceToHost, c->compute) != hipSuccess ||
        hipStreamSynchronize(c->compute) != hipSuccess)
    {
        (void)hipGetLastError();
        dfree(d);
        return set_err(GR_E_HIP, "compute-unit probe failed");
    }
    dfree(d);
    std::vector<char> seen(nkeys, 0);
    for (uint32_t k : h) seen[k & (nkeys - 1)] = 1;
    std::vector<uint32_t> lut(nkeys, 0xffffffffu);
    uint32_t nq = 0;
    for (size_t k = 0; k < nkeys; ++k)
        if (seen[k] && nq < gr::MAX_QUEUES) lut[k] = nq++;
    c->ncu = std::max(1u, nq);
    c->batch_group = std::max(1u, c->ncu / std::max(1u, c->nxcd)); // one queue per XCD
    if ((rc = dalloc(&c->d_culut, nkeys * sizeof(uint32_t))) != GR_OK) return rc;
    GR_HIP(hipMemcpy(c->d_culut, lut.data(), nkeys * sizeof(uint32_t), hipMemcpyHostToDevice));
    return GR_OK;
}

extern "C" int gr_ctx_create(int device, gr_ctx ** out)
{
    if (!out) return set_err(GR_E_INVALID, "gr_ctx_create: null out");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    {
        (void)hipGetLastError();
        return set_err(GR_E_HIP, "no HIP device available");
    }
    if (device < 0 || device >= n) return set_err(GR_E_INVALID, "device %d out of range", device);
    GR_HIP(hipSetDevice(device));
    gr_ctx * c = new (std::nothrow) gr_ctx;
    if (!c) return set_err(GR_E_NOMEM, "gr_ctx_create: out of host memory");
    c->device = device;
    // default-priority streams only
    if (hipStreamCreateWithFlags(&c->compute, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->copy, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&c->dcount), sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&c->dwork), size_t(gr::WORK_SLOTS) * gr::WORK_STRIDE * 8) != hipSuccess ||
        hipMemsetAsync(c->dwork, 0, size_t(gr::WORK_SLOTS) * gr::WORK_STRIDE * 8, c->compute) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&c->dtile), size_t(gr::WORK_SLOTS) * gr::WORK_STRIDE * 8) != hipSuccess ||
        hipMemsetAsync(c->dtile, 0, size_t(gr::WORK_SLOTS) * gr::WORK_STRIDE * 8, c->compute) != hipSuccess)
    {
        (void)hipGetLastError();
        if (c->compute) (void)hipStreamDestroy(c->compute);
        if (c->copy) (void)hipStreamDestroy(c->copy);
        dfree(c->dcount);
        dfree(c->dwork);
        dfree(c->dtile);
        delete c;
        return set_err(GR_E_HIP, "gr_ctx_create: stream / buffer creation failed");
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess)
    {
        c->cus = prop.multiProcessorCount;
        c->lds_per_cu = uint32_t(prop.maxSharedMemoryPerMultiProcessor);
    }
    int nx = 1;
    if (hipDeviceGetAttribute(&nx, hipDeviceAttributeNumberOfXccs, device) == hipSuccess && nx >= 1)
        c->nxcd = uint32_t(nx);
    (void)hipGetLastError();
    int rc = map_compute_units(c);
    // the no-allocation entry points: this device's scratch address, and their launch event
    if (rc == GR_OK && gr::noalloc_setup(device) != 0)
        rc = set_err(GR_E_HIP, "gr_ctx_create: scratch address of device %d not found", device);
    if (rc == GR_OK && hipEventCreateWithFlags(&c->noalloc_done, hipEventDisableTiming) != hipSuccess)
        rc = set_err(GR_E_HIP, "gr_ctx_create: event creation failed");
    if (rc != GR_OK)
    {
        gr_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return GR_OK;
}

extern "C" int gr_ctx_destroy(gr_ctx * c)
{
    if (!c) return GR_OK;
    if (c->children > 0)
        return set_err(GR_E_STATE, "gr_ctx_destroy: %d fields/jobs of this context are still alive",
                       c->children);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->compute);
    (void)hipStreamSynchronize(c->copy);
    for (hipEvent_t e : c->kev) (void)hipEventDestroy(e);
    dfree(c->dcount);
    dfree(c->dwork);
    dfree(c->dtile);
    dfree(c->d_culut);
    dfree(c->d_stage);
    dfree(c->d_items);
    for (auto & st : c->stage)
    {
        if (st.h) (void)hipHostFree(st.h);
        if (st.done) (void)hipEventDestroy(st.done);
    }
    for (int k = 0; k < 2; ++k)
    {
        dfree(c->skeys[k]);
        dfree(c->svals[k]);
    }
    dfree(c->stmp);
    dfree(c->d_sjobs);
    dfree(c->bcount);
    dfree(c->bstart);
    dfree(c->d_gtab);
    if (c->bev0) (void)hipEventDestroy(c->bev0);
    if (c->bev1) (void)hipEventDestroy(c->bev1);
    if (c->noalloc_done) (void)hipEventDestroy(c->noalloc_done);
    for (hipEvent_t & e : c->alive_ev)
        if (e) (void)hipEventDestroy(e);
    for (uint32_t * pg : c->alive_pages) (void)hipHostFree(pg);
    pool_flush(c);
    (void)hipStreamDestroy(c->compute);
    (void)hipStreamDestroy(c->copy);
    delete c;
    return GR_OK;
}

extern "C" int gr_ctx_synchronize(gr_ctx * c)
{
    if (!c) return set_err(GR_E_INVALID, "null ctx");
    GR_HIP(hipSetDevice(c->device));
    GR_HIP(hipStreamSynchronize(c->copy));
    GR_HIP(hipStreamSynchronize(c->compute));
    return GR_OK;
}

extern "C" int gr_ctx_set_option(gr_ctx * c, uint32_t option, int64_t value)
{
    if (!c) return set_err(GR_E_INVALID, "null ctx");
    switch (option)
    {
        case GR_CTX_BATCH_BLOCK:
            if (value <= 0 || value % 64 != 0 || value > 256)
                return set_err(GR_E_INVALID, "batch block must be a multiple of 64 up to 256");
            c->batch_block = uint32_t(value);
            return GR_OK;
        case GR_CTX_BATCH_WAVES:
            if (value != 0 && value != 4) return set_err(GR_E_INVALID, "batch waves must be 0 or 4");
            c->batch_waves = uint32_t(value);
            return GR_OK;
        case GR_CTX_BATCH_GROUP:
            if (value < 1 || value > int64_t(c->ncu))
                return set_err(GR_E_INVALID, "batch group must be 1 .. %u compute units", c->ncu);
            c->batch_group = uint32_t(value);
            return GR_OK;
        case GR_CTX_BATCH_CU_WAVES:
            if (value < 0 || value > 32) return set_err(GR_E_INVALID, "batch CU waves must be 0 .. 32");
            c->batch_cu_waves = uint32_t(value);
            return GR_OK;
        case GR_CTX_BATCH_REUSE:
            if (value != 0 && value != 1) return set_err(GR_E_INVALID, "batch reuse must be 0 or 1");
            c->batch_reuse = uint32_t(value);
            return GR_OK;
        case GR_CTX_BATCH_UNIFORM:
            if (value != 0 && value != 1) return set_err(GR_E_INVALID, "batch uniform must be 0 or 1");
            c->batch_uniform = uint32_t(value);
            return GR_OK;
        case GR_CTX_POOL_BYTES:
            if (value < 0) return set_err(GR_E_INVALID, "pool bytes must be >= 0");
            c->pool_cap = size_t(value);
            if (c->pool_cached > c->pool_cap) pool_flush(c);
            return GR_OK;
        case GR_CTX_BATCH_QUAD_ITEMS:
        case GR_CTX_BATCH_LAST_KERNEL:
        case GR_CTX_POOL_CACHED: return set_err(GR_E_INVALID, "context option %u is read-only", option);
        default: return set_err(GR_E_INVALID, "unknown context option %u", option);
    }
}

extern "C" int gr_ctx_get_option(gr_ctx * c, uint32_t option, int64_t * value)
{
    if (!c || !value) return set_err(GR_E_INVALID, "gr_ctx_get_option: null argument");
    switch (option)
    {
        case GR_CTX_BATCH_BLOCK: *value = c->batch_block; return GR_OK;
        case GR_CTX_BATCH_WAVES: *value = c->batch_waves; return GR_OK;
        case GR_CTX_BATCH_GROUP: *value = c->batch_group; return GR_OK;
        case GR_CTX_BATCH_CU_WAVES: *value = c->batch_cu_waves; return GR_OK;
        case GR_CTX_BATCH_REUSE: *value = c->batch_reuse; return GR_OK;
        case GR_CTX_BATCH_QUAD_ITEMS: *value = int64_t(c->items_valid ? c->items_quad : 0); return GR_OK;
        case GR_CTX_BATCH_UNIFORM: *value = c->batch_uniform; return GR_OK;
        case GR_CTX_POOL_BYTES: *value = int64_t(c->pool_cap); return GR_OK;
        case GR_CTX_POOL_CACHED: *value = int64_t(c->pool_cached); return GR_OK;
        case GR_CTX_BATCH_LAST_KERNEL: *value = c->batch_last_kernel; return GR_OK;
        default: return set_err(GR_E_INVALID, "unknown context option %u", option);
    }
}

// sums the recorded launch times into the accumulator (waits for the last launch)
static int kt_fold(gr_ctx * c)
{
    if (!c->kpending) return GR_OK;
    GR_HIP(hipEventSynchronize(c->kev[2 * c->kpending - 1]));
    for (size_t i = 0; i < c->kpending; ++i)
    {
        float ms = 0.f;
        GR_HIP(hipEventElapsedTime(&ms, c->kev[2 * i], c->kev[2 * i + 1]));
        c->kms_acc += ms;
    }
    c->klaunch_acc += c->kpending;
    c->kpending = 0;
    return GR_OK;
}

// Records the start or end event of a step-kernel launch (end: *done = the event, which also
// marks the end of everything the launch read).
static int kt_mark(gr_ctx * c, hipStream_t s, bool end, hipEvent_t * done = nullptr)
{
    if (!end && c->kpending == 1024)
    {
        int rc = kt_fold(c);
        if (rc != GR_OK) return rc;
    }
    const size_t i = 2 * c->kpending + (end ? 1 : 0);
    while (c->kev.size() <= i)
    {
        hipEvent_t e;
        GR_HIP(hipEventCreateWithFlags(&e, TIMING_EVENT));
        c->kev.push_back(e);
    }
    GR_HIP(hipEventRecord(c->kev[i], s));
    if (end) ++c->kpending;
    if (done) *done = c->kev[i];
    return GR_OK;
}

extern "C" int gr_ctx_take_kernel_ms(gr_ctx * c, double * ms, uint64_t * launches)
{
    if (!c || !ms) return set_err(GR_E_INVALID, "gr_ctx_take_kernel_ms: null argument");
    GR_HIP(hipSetDevice(c->device));
    int rc = kt_fold(c);
    if (rc != GR_OK) return rc;
    *ms = c->kms_acc;
    if (launches) *launches = c->klaunch_acc;
    c->kms_acc = 0.0;
    c->klaunch_acc = 0;
    return GR_OK;
}

extern "C" int gr_ctx_take_work(gr_ctx * c, uint64_t * particle_steps)
{
    if (!c || !particle_steps) return set_err(GR_E_INVALID, "gr_ctx_take_work: null argument");
    GR_HIP(hipSetDevice(c->device));
    std::vector<unsigned long long> h(size_t(gr::WORK_SLOTS) * gr::WORK_STRIDE);
    GR_HIP(hipMemcpyAsync(h.data(), c->dwork, h.size() * 8, hipMemcpyDeviceToHost, c->compute));
    GR_HIP(hipMemsetAsync(c->dwork, 0, h.size() * 8, c->compute));
    GR_HIP(hipStreamSynchronize(c->compute));
    uint64_t sum = 0;
    for (uint32_t k = 0; k < gr::WORK_SLOTS; ++k) sum += h[size_t(k) * gr::WORK_STRIDE];
    *particle_steps = sum;
    return GR_OK;
}

extern "C" int gr_ctx_take_tile_stats(gr_ctx * c, uint64_t * tiled, uint64_t * untiled)
{
    if (!c || !tiled || !untiled) return set_err(GR_E_INVALID, "gr_ctx_take_tile_stats: null argument");
    GR_HIP(hipSetDevice(c->device));
    std::vector<unsigned long long> h(size_t(gr::WORK_SLOTS) * gr::WORK_STRIDE);
    GR_HIP(hipMemcpyAsync(h.data(), c->dtile, h.size() * 8, hipMemcpyDeviceToHost, c->compute));
    GR_HIP(hipMemsetAsync(c->dtile, 0, h.size() * 8, c->compute));
    GR_HIP(hipStreamSynchronize(c->compute));
    uint64_t a = 0, b = 0;
    for (uint32_t k = 0; k < gr::WORK_SLOTS; ++k)
    {
        a += h[size_t(k) * gr::WORK_STRIDE];
        b += h[size_t(k) * gr::WORK_STRIDE + 1];
    }
    *tiled = a;
    *untiled = b;
    return GR_OK;
}

extern "C" int gr_ctx_copy_bandwidth(gr_ctx * c, uint64_t bytes, uint32_t trials, double * gbs)
{
    if (!c || !gbs) return set_err(GR_E_INVALID, "gr_ctx_copy_bandwidth: null argument");
    if (bytes == 0 || (bytes & 15)) return set_err(GR_E_INVALID, "gr_ctx_copy_bandwidth: bytes must be a positive multiple of 16");
    if (trials == 0) trials = 5;
    GR_HIP(hipSetDevice(c->device));
    void * a = nullptr;
    void * b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = GR_OK;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
        hipMemsetAsync(a, 0, bytes, c->compute) != hipSuccess)
        rc = set_err(GR_E_HIP, "gr_ctx_copy_bandwidth: allocation failed");
    // enough workgroups to keep every CU busy: 16 per CU
    const uint32_t grid = uint32_t(std::max(1, c->cus)) * 16u;
    float best = 3.4e38f;
    for (uint32_t t = 0; rc == GR_OK && t <= trials; ++t) // the first copy warms up
    {
        if (hipEventRecord(e0, c->compute) != hipSuccess || gr::launch_stream_copy(a, b, bytes, grid, c->compute) != 0 ||
            hipEventRecord(e1, c->compute) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
        {
            rc = set_err(GR_E_HIP, "gr_ctx_copy_bandwidth: copy failed");
            break;
        }
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            rc = set_err(GR_E_HIP, "gr_ctx_copy_bandwidth: timing failed");
        else if (t > 0)
            best = std::min(best, ms);
    }
    (void)hipStreamSynchronize(c->compute);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dfree(a);
    dfree(b);
    if (rc != GR_OK) return rc;
    *gbs = 2.0 * double(bytes) / (double(best) * 1e-3) / 1e9;
    return GR_OK;
}

extern "C" int gr_ctx_info(gr_ctx * c, char * name, size_t name_len, uint64_t * hbm_bytes,
                           int * compute_units)
{
    if (!c) return set_err(GR_E_INVALID, "null ctx");
    hipDeviceProp_t prop;
    GR_HIP(hipGetDeviceProperties(&prop, c->device));
    if (name && name_len)
    {
        std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
    if (compute_units) *compute_units = prop.multiProcessorCount;
    return GR_OK;
}

// --------------------------------------------------------------------------------------------
// fields
// --------------------------------------------------------------------------------------------
static void field_free(gr_field * f)
{
    if (!f) return;
    for (void * s : f->slots) dfree(s);
    for (auto & p : f->pairs) dfree(p.buf);
    dfree(f->zlev);
    dfree(f->zlut);
    dfree(f->dstage);
    if (f->hstage) (void)hipHostFree(f->hstage);
    if (f->stage_done) (void)hipEventDestroy(f->stage_done);
    for (hipEvent_t e : f->ready)
        if (e) (void)hipEventDestroy(e);
    delete f;
}

extern "C" int gr_field_create(gr_ctx * ctx, const gr_field_desc * desc, gr_field ** out)
{
    if (!ctx || !desc || !out) return set_err(GR_E_INVALID, "gr_field_create: null argument");
    *out = nullptr;
    const gr_field_desc & d = *desc;
    if (d.nx == 0 || d.ny == 0 || d.nz == 0 || d.ncomp == 0 || d.ncomp > GR_MAX_VOLUMES ||
        d.nslots == 0 || d.nslots > 4096 || d.dtype > GR_BF16 || d.address > 1)
        return set_err(GR_E_INVALID, "gr_field_create: invalid descriptor");
    const uint64_t nodes = uint64_t(d.nx) * d.ny * d.nz;
    if (nodes >= (1ull << 32)) return set_err(GR_E_INVALID, "gr_field_create: grid too large");
    GR_HIP(hipSetDevice(ctx->device));
    gr_field * f = new (std::nothrow) gr_field;
    if (!f) return set_err(GR_E_NOMEM, "gr_field_create: out of host memory");
    f->ctx = ctx;
    f->d = d;
    f->cstride = d.ncomp == 3 ? 4 : d.ncomp;
    f->nodes = nodes;
    f->slot_bytes = nodes * f->cstride * dtype_size(d.dtype);
    f->slots.assign(d.nslots, nullptr);
    f->ready.assign(d.nslots, nullptr);
    f->inuse.assign(d.nslots, nullptr);
    f->has_ready.assign(d.nslots, 0);
    f->has_inuse.assign(d.nslots, 0);
    f->waited.assign(d.nslots, 0);
    int rc = GR_OK;
    for (uint32_t s = 0; s < d.nslots && rc == GR_OK; ++s)
    {
        rc = dalloc(&f->slots[s], f->slot_bytes);
        // on the copy stream, ahead of the slot's uploads (which run there)
        if (rc == GR_OK && hipMemsetAsync(f->slots[s], 0, f->slot_bytes, ctx->copy) != hipSuccess)
            rc = set_err(GR_E_HIP, "gr_field_create: memset failed");
        if (rc == GR_OK && hipEventCreateWithFlags(&f->ready[s], hipEventDisableTiming) != hipSuccess)
            rc = set_err(GR_E_HIP, "gr_field_create: event creation failed");
    }
    if (rc == GR_OK) rc = dalloc(&f->dstage, nodes * d.ncomp * sizeof(float));
    if (rc == GR_OK && hipHostMalloc(reinterpret_cast<void **>(&f->hstage),
                                     nodes * d.ncomp * sizeof(float), hipHostMallocDefault) != hipSuccess)
    {
        f->hstage = nullptr;
        rc = set_err(GR_E_NOMEM, "gr_field_create: pinned staging allocation failed");
    }
    if (rc == GR_OK && hipEventCreateWithFlags(&f->stage_done, hipEventDisableTiming) != hipSuccess)
        rc = set_err(GR_E_HIP, "gr_field_create: event creation failed");
    if (rc != GR_OK)
    {
        (void)hipGetLastError();
        field_free(f);
        return rc;
    }
    ctx->children++;
    *out = f;
    return GR_OK;
}

extern "C" int gr_field_destroy(gr_field * f)
{
    if (!f) return GR_OK;
    gr_ctx * c = f->ctx;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->copy);
    (void)hipStreamSynchronize(c->compute);
    field_free(f);
    c->children--;
    return GR_OK;
}

static int field_pack(gr_field * f, uint32_t slot)
{
    gr_ctx * c = f->ctx;
    for (auto & p : f->pairs)
        if (p.b == slot || p.f == slot) p.valid = false; // stale once the slot changes
    if (gr::launch_pack(f->dstage, f->slots[slot], f->nodes, f->d.ncomp, f->cstride, f->d.dtype,
                        c->copy) != 0)
        return set_err(GR_E_HIP, "pack kernel launch failed: %s", hipGetErrorString(hipGetLastError()));
    GR_HIP(hipEventRecord(f->ready[slot], c->copy));
    f->has_ready[slot] = 1;
    f->waited[slot] = 0;
    return GR_OK;
}

extern "C" int gr_field_upload(gr_field * f, uint32_t slot, const float * const * comps, int async)
{
    if (!f || !comps) return set_err(GR_E_INVALID, "gr_field_upload: null argument");
    if (slot >= f->d.nslots) return set_err(GR_E_INVALID, "gr_field_upload: slot %u out of range", slot);
    for (uint32_t k = 0; k < f->d.ncomp; ++k)
        if (!comps[k]) return set_err(GR_E_INVALID, "gr_field_upload: component %u is null", k);
    gr_ctx * c = f->ctx;
    GR_HIP(hipSetDevice(c->device));
    // the pinned staging may still feed the previous H2D copy
    if (f->stage_pending) GR_HIP(hipEventSynchronize(f->stage_done));
    const size_t bytes = f->nodes * sizeof(float);
    for (uint32_t k = 0; k < f->d.ncomp; ++k) std::memcpy(f->hstage + k * f->nodes, comps[k], bytes);
    // do not overwrite a slot a queued kernel still reads
    if (f->has_inuse[slot]) GR_HIP(hipStreamWaitEvent(c->copy, f->inuse[slot], 0));
    GR_HIP(hipMemcpyAsync(f->dstage, f->hstage, bytes * f->d.ncomp, hipMemcpyHostToDevice, c->copy));
    GR_HIP(hipEventRecord(f->stage_done, c->copy));
    f->stage_pending = true;
    int rc = field_pack(f, slot);
    if (rc != GR_OK) return rc;
    if (!async) GR_HIP(hipStreamSynchronize(c->copy));
    return GR_OK;
}

extern "C" int gr_field_upload_device(gr_field * f, uint32_t slot, const float * const * dcomps)
{
    if (!f || !dcomps) return set_err(GR_E_INVALID, "gr_field_upload_device: null argument");
    if (slot >= f->d.nslots) return set_err(GR_E_INVALID, "slot %u out of range", slot);
    gr_ctx * c = f->ctx;
    GR_HIP(hipSetDevice(c->device));
    if (f->has_inuse[slot]) GR_HIP(hipStreamWaitEvent(c->copy, f->inuse[slot], 0));
    const size_t bytes = f->nodes * sizeof(float);
    for (uint32_t k = 0; k < f->d.ncomp; ++k)
        GR_HIP(hipMemcpyAsync(f->dstage + k * f->nodes, dcomps[k], bytes, hipMemcpyDeviceToDevice, c->copy));
    return field_pack(f, slot);
}

extern "C" int gr_field_wait(gr_field * f, uint32_t slot)
{
    if (!f || slot >= f->d.nslots) return set_err(GR_E_INVALID, "gr_field_wait: bad argument");
    GR_HIP(hipSetDevice(f->ctx->device));
    if (f->has_ready[slot]) GR_HIP(hipEventSynchronize(f->ready[slot]));
    return GR_OK;
}

extern "C" int gr_field_set_z_levels(gr_field * f, const float * levels, uint32_t n)
{
    if (!f || !levels) return set_err(GR_E_INVALID, "gr_field_set_z_levels: null argument");
    if (n != f->d.nz || n < 2) return set_err(GR_E_INVALID, "need nz = %u levels", f->d.nz);
    double min_dz = 1e300;
    for (uint32_t k = 0; k + 1 < n; ++k)
    {
        const double dz = double(levels[k + 1]) - double(levels[k]);
        if (!(dz > 0.0)) return set_err(GR_E_INVALID, "z levels must increase strictly");
        min_dz = std::min(min_dz, dz);
    }
    const double span = double(levels[n - 1]) - double(levels[0]);
    // bins at most half the thinnest layer, so the search from the table is a step or two
    const uint32_t nlut = uint32_t(std::min<double>(1 << 20, std::ceil(2.0 * span / min_dz) + 1));
    const float scale = float(double(nlut) / span);
    std::vector<int32_t> lut(nlut);
    int32_t k = 0;
    for (uint32_t m = 0; m < nlut; ++m)
    {
        const double zb = double(levels[0]) + double(m) / double(scale);
        while (k + 2 < int32_t(n) && double(levels[k + 1]) <= zb) ++k;
        lut[m] = k;
    }
    GR_HIP(hipSetDevice(f->ctx->device));
    dfree(f->zlev);
    dfree(f->zlut);
    f->zlev = nullptr;
    f->zlut = nullptr;
    int rc;
    if ((rc = dalloc(&f->zlev, n * sizeof(float))) != GR_OK) return rc;
    if ((rc = dalloc(&f->zlut, nlut * sizeof(int32_t))) != GR_OK) return rc;
    GR_HIP(hipMemcpy(f->zlev, levels, n * sizeof(float), hipMemcpyHostToDevice));
    GR_HIP(hipMemcpy(f->zlut, lut.data(), nlut * sizeof(int32_t), hipMemcpyHostToDevice));
    f->nlut = nlut;
    f->zlut_z0 = levels[0];
    f->zlut_scale = scale;
    return GR_OK;
}

extern "C" int gr_field_bytes(const gr_field * f, uint64_t * bytes)
{
    if (!f || !bytes) return set_err(GR_E_INVALID, "gr_field_bytes: null argument");
    *bytes = f->slot_bytes;
    return GR_OK;
}

// Returns (building it on stream `s` if needed) the time-paired copy of slots (b, f). The caller
// has already made `s` wait for both slots' uploads; the slots are marked in use after the
// build so later uploads wait for it.
static int field_pair(gr_field * f, uint32_t b, uint32_t fr, hipStream_t s, const void ** out)
{
    *out = nullptr;
    if (f->d.ncomp != 2 && f->d.ncomp != 3) return set_err(GR_E_INVALID, "pairing needs 2 or 3 components");
    gr_field::Pair * hit = nullptr;
    for (auto & p : f->pairs)
        if (p.valid && p.b == b && p.f == fr) hit = &p;
    if (!hit)
    {
        const size_t max_pairs = std::max<size_t>(2, f->d.nslots > 1 ? f->d.nslots - 1 : 1);
        if (f->pairs.size() < max_pairs)
        {
            // a stale entry is reused before the cache grows
            for (auto & p : f->pairs)
                if (!p.valid) hit = &p;
            if (!hit)
            {
                f->pairs.emplace_back();
                hit = &f->pairs.back();
            }
        }
        else
        {
            // stale entries first, then the least recently used one
            hit = &f->pairs[0];
            for (auto & p : f->pairs)
                if ((!p.valid && hit->valid) || (p.valid == hit->valid && p.used < hit->used)) hit = &p;
        }
        const uint32_t so = f->d.ncomp == 3 ? 8 : 4;
        const uint32_t dim = f->d.ncomp == 3 ? 3 : 2;
        if (!hit->buf)
        {
            const uint64_t pn = gr::pair_nodes(f->d.nx, f->d.ny, f->d.nz, dim);
            int rc = dalloc(&hit->buf, pn * so * dtype_size(f->d.dtype));
            if (rc != GR_OK) return rc;
        }
        if (gr::launch_pair(f->slots[b], f->slots[fr], hit->buf, f->d.nx, f->d.ny, f->d.nz, dim,
                            f->d.address, f->d.ncomp, f->cstride, f->d.dtype, s) != 0)
            return set_err(GR_E_HIP, "pair kernel launch failed");
        hit->b = b;
        hit->f = fr;
        hit->valid = true;
    }
    hit->used = ++f->pair_clock;
    *out = hit->buf;
    return GR_OK;
}

// --------------------------------------------------------------------------------------------
// jobs
// --------------------------------------------------------------------------------------------

static int zalloc(float ** p, size_t n)
{
    int rc = dalloc(p, n * sizeof(float));
    if (rc == GR_OK && *p && hipMemset(*p, 0, n * sizeof(float)) != hipSuccess)
        return set_err(GR_E_HIP, "memset failed");
    return rc;
}

// Every device array of a job at creation: where its pointer goes, its size, and the byte its
// contents start as (-1: left uninitialised). `state` marks the per-particle state arrays that a
// reorder gathers (each gets a shadow copy then). gr_job_create allocates exactly these and
// gr_job_bytes sums them, so the two cannot drift apart.
struct JobArray
{
    void ** slot;
    size_t bytes;
    int fill;
    bool state;
};
static std::vector<JobArray> job_arrays(gr_job * j)
{
    const gr_job_desc & d = j->d;
    const size_t N = size_t(j->n ? j->n : 1), D = d.dim, cap = j->cap;
    std::vector<JobArray> a;
    auto add = [&](auto *& p, size_t bytes, int fill, bool state) {
        a.push_back(JobArray{reinterpret_cast<void **>(&p), bytes, fill, state});
    };
    const bool curv_pos = d.curvature == GR_CURV_TOTAL || d.curvature == GR_CURV_INDIVIDUAL_AND_TOTAL;
    for (size_t k = 0; k < D; ++k) add(j->p[k], N * 4, 0, true);
    for (size_t k = 0; k < D && curv_pos; ++k) add(j->q[k], N * 4, 0, true);
    for (size_t k = 0; k < D && d.kahan; ++k) add(j->c[k], N * 4, 0, true);
    for (size_t k = 0; k < D && d.curvature == GR_CURV_FAST_TOTAL; ++k) add(j->ld[k], N * 4, 0, true);
    add(j->status, N * 4, 0, true);
    add(j->abort_step, N * 4, 0xff, true);
    if (d.curvature != GR_CURV_OFF) add(j->total, N * 4, 0, true);
    if (d.curvature == GR_CURV_INDIVIDUAL_AND_TOTAL) add(j->curv, N * 4, 0, true);
    if (d.abort_mode != GR_ABORT_TIME) add(j->len, N * 4, 0, true);
    for (uint32_t e = 0; e < d.nprog; ++e) add(j->expr[e], N * 4, 0, true);
    for (uint32_t k = 0; k < d.nvol; ++k) add(j->vol[k], N * 4, 0, true);
    if (d.save_interval != 0)
    {
        add(j->ring_pos, cap * N * D * 4, 0, false);
        if (d.curvature == GR_CURV_INDIVIDUAL_AND_TOTAL) add(j->ring_curv, cap * N * 4, 0, false);
        if (d.nprog) add(j->ring_expr, cap * N * d.nprog * 4, 0, false);
        if (d.nvol) add(j->ring_vol, cap * N * d.nvol * 4, 0, false);
    }
    add(j->scratch_aos, N * D * 4, -1, false);
    add(j->scratch2, N * D * 4, -1, false);
    return a;
}

static void job_free(gr_job * j)
{
    if (!j) return;
    gr_ctx * c = j->ctx;
    // the arrays of job_arrays (a reorder swaps state arrays with shadows of the same size), the
    // index map and the shadows go back to the context's cache
    for (const JobArray & a : job_arrays(j))
    {
        pool_free(c, *a.slot, a.bytes);
        *a.slot = nullptr;
    }
    pool_free(c, j->orig, size_t(j->n) * 4);
    for (void * a : j->shadow) pool_free(c, a, size_t(j->n) * 4);
    for (int d = 0; d < 3; ++d)
    {
        dfree(j->uplift[d]);
        for (int k = 0; k < GR_MAX_CONSTANTS; ++k) dfree(j->cst[d][k]);
    }
    if (j->ev0) (void)hipEventDestroy(j->ev0);
    if (j->ev1) (void)hipEventDestroy(j->ev1);
    for (auto * v : {&j->reads, &j->read_pool})
        for (gr_job::Read & r : *v)
        {
            if (r.host) (void)hipHostFree(r.host);
            dfree(r.dstage);
            if (r.ready) (void)hipEventDestroy(r.ready);
            if (r.done) (void)hipEventDestroy(r.done);
        }
    // the slot is reused only after the compute stream has drained (gr_job_destroy)
    if (j->h_alive && c) c->alive_free.push_back(j->h_alive);
    delete j;
}

extern "C" int gr_job_bytes(const gr_job_desc * desc, uint64_t n, uint32_t save_capacity,
                            int with_reorder, uint64_t * bytes)
{
    if (!desc || !bytes) return set_err(GR_E_INVALID, "gr_job_bytes: null argument");
    if (desc->dim != 2 && desc->dim != 3) return set_err(GR_E_INVALID, "dim must be 2 or 3");
    if (desc->nprog > GR_MAX_PROGRAMS || desc->nvol > GR_MAX_VOLUMES)
        return set_err(GR_E_INVALID, "gr_job_bytes: too many programs / volumes");
    gr_job j;
    j.d = *desc;
    j.n = n;
    j.cap = save_capacity ? save_capacity : 1;
    uint64_t total = 0, nstate = 0;
    for (const JobArray & a : job_arrays(&j))
    {
        total += a.bytes;
        nstate += a.state ? 1 : 0;
    }
    // a reorder adds the index map and a shadow copy of every state array and of the map
    if (with_reorder && n >= 2) total += (nstate + 2) * uint64_t(n) * 4;
    *bytes = total;
    return GR_OK;
}

extern "C" int gr_job_create(gr_ctx * ctx, const gr_job_desc * desc, const float * seeds,
                             uint64_t n, uint32_t save_capacity, gr_job ** out)
{
    if (!ctx || !desc || !out || (n && !seeds)) return set_err(GR_E_INVALID, "gr_job_create: null argument");
    *out = nullptr;
    const gr_job_desc & d = *desc;
    if (d.dim != 2 && d.dim != 3) return set_err(GR_E_INVALID, "dim must be 2 or 3");
    if (d.integrator > GR_RK4 || d.border > GR_LOOP_XY || d.curvature > GR_CURV_INDIVIDUAL_AND_TOTAL ||
        d.abort_mode > GR_ABORT_FIT_LENGTH || d.uplift_mode > 2 || d.windfield_mode > 1 ||
        d.constants_mode > 2 || d.volume_mode > 2 || d.weights > GR_WEIGHTS_Q8_TRUNC)
        return set_err(GR_E_INVALID, "gr_job_create: enum out of range");
    if (d.nprog > GR_MAX_PROGRAMS || d.nvol > GR_MAX_VOLUMES || d.nconst > GR_MAX_CONSTANTS)
        return set_err(GR_E_INVALID, "gr_job_create: too many programs / volumes / constants");
    if (n >= (1ull << 31)) return set_err(GR_E_INVALID, "gr_job_create: too many particles");
    if (!(d.dt == d.dt)) return set_err(GR_E_INVALID, "gr_job_create: DeltaT is NaN");
    uint32_t expr_cols = 0;
    for (uint32_t e = 0; e < d.nprog; ++e)
    {
        const gr_program & pr = d.prog[e];
        if (pr.n == 0 || pr.n > GR_MAX_CODE)
            return set_err(GR_E_INVALID, "gr_job_create: program %u is empty or too long", e);
        // the device interpreter trusts the bytecode: check opcodes, operands and stack bounds
        int sp = 0;
        for (uint32_t i = 0; i < pr.n; ++i)
        {
            const uint32_t op = pr.code[i] & 0xffu, arg = pr.code[i] >> 8;
            int pops, pushes = 1;
            if (op == gr::OP_REF || op == gr::OP_CONST)
            {
                pops = 0;
                if ((op == gr::OP_REF && arg >= 19) || (op == gr::OP_CONST && arg >= GR_MAX_CODE))
                    return set_err(GR_E_INVALID, "program %u: operand out of range at %u", e, i);
            }
            else if (op == gr::OP_ADD || op == gr::OP_SUB || op == gr::OP_MUL || op == gr::OP_DIV ||
                     op == gr::OP_ATAN2)
                pops = 2;
            else if (op >= gr::OP_SQRT && op <= gr::OP_ABS)
                pops = 1;
            else
                return set_err(GR_E_INVALID, "program %u: bad opcode %u at %u", e, op, i);
            if (sp < pops) return set_err(GR_E_INVALID, "program %u: stack underflow at %u", e, i);
            sp += pushes - pops;
            if (sp > GR_MAX_STACK) return set_err(GR_E_INVALID, "program %u: stack overflow at %u", e, i);
            if (sp > 2) expr_cols = std::max(expr_cols, uint32_t(sp - 2));
        }
        if (sp != 1) return set_err(GR_E_INVALID, "program %u leaves %d values on the stack", e, sp);
    }
    GR_HIP(hipSetDevice(ctx->device));
    gr_job * j = new (std::nothrow) gr_job;
    if (!j) return set_err(GR_E_NOMEM, "gr_job_create: out of host memory");
    j->ctx = ctx;
    j->d = d;
    j->n = n;
    j->run_n = n;
    j->expr_cols = expr_cols;
    j->cap = save_capacity ? save_capacity : 1;
    // quad launches below 128 live particles per compute unit (measured on the ABC workload:
    // 64 .. 512 per CU tried, 128 best for FP32 and FP16 storage)
    j->quad_max = 128ull * uint64_t(std::max(1, ctx->cus));
    j->full = (d.dim == 2 && (d.border != GR_BLOCK || d.sphere_radius > 0.f)) ||
              d.abort_mode != GR_ABORT_TIME || d.uplift_mode != GR_MODE_OFF ||
              d.constants_mode != GR_MODE_OFF || d.volume_mode != GR_MODE_OFF || d.nvol > 0 ||
              d.nconst > 0;
    const uint32_t D = d.dim;
    int rc = GR_OK;
    const bool curv_pos = d.curvature == GR_CURV_TOTAL || d.curvature == GR_CURV_INDIVIDUAL_AND_TOTAL;
    for (const JobArray & a : job_arrays(j))
    {
        if (rc != GR_OK) break;
        rc = pool_alloc(ctx, a.slot, a.bytes);
        // filled on the compute stream, ahead of the seed kernel and every launch of the job
        // (the compute stream does not wait for work on the null stream)
        if (rc == GR_OK && a.fill >= 0 && hipMemsetAsync(*a.slot, a.fill, a.bytes, ctx->compute) != hipSuccess)
            rc = set_err(GR_E_HIP, "gr_job_create: memset failed");
    }
    for (int k = 0; k < 3; ++k)
        j->grid[k] = d.grid_scale[k] > 0.f ? uint32_t(std::lround(double(d.domain[k]) / d.grid_scale[k])) : 1;
    if (rc == GR_OK && (hipEventCreateWithFlags(&j->ev0, TIMING_EVENT) != hipSuccess ||
                        hipEventCreateWithFlags(&j->ev1, TIMING_EVENT) != hipSuccess))
        rc = set_err(GR_E_HIP, "event creation failed");
    if (rc == GR_OK && n)
    {
        // seeds: AoS -> device AoS staging -> SoA state (optionally relative to the origin).
        // Device seeds are copied on the compute stream (ordered, nothing waits); host seeds
        // on the copy stream, waiting for that copy alone (not for kernels queued on the compute
        // stream), so that the caller's buffer is free when the call returns.
        hipPointerAttribute_t pa;
        std::memset(&pa, 0, sizeof(pa));
        const bool on_device = hipPointerGetAttributes(&pa, seeds) == hipSuccess &&
                               (pa.type == hipMemoryTypeDevice || pa.type == hipMemoryTypeManaged);
        (void)hipGetLastError();
        const size_t bytes = n * D * sizeof(float);
        if (on_device)
        {
            if (hipMemcpyAsync(j->scratch_aos, seeds, bytes, hipMemcpyDeviceToDevice, ctx->compute) != hipSuccess)
                rc = set_err(GR_E_HIP, "seed copy failed");
        }
        else
        {
            if (hipMemcpyAsync(j->scratch_aos, seeds, bytes, hipMemcpyHostToDevice, ctx->copy) != hipSuccess ||
                hipEventRecord(j->ev0, ctx->copy) != hipSuccess || hipEventSynchronize(j->ev0) != hipSuccess ||
                hipStreamWaitEvent(ctx->compute, j->ev0, 0) != hipSuccess)
                rc = set_err(GR_E_HIP, "seed upload failed");
        }
        float * soa[3] = {j->p[0], j->p[1], D == 3 ? j->p[2] : nullptr};
        if (rc == GR_OK && gr::launch_seed(j->scratch_aos, soa, n, D, d.offset, d.local_coords, ctx->compute) != 0)
            rc = set_err(GR_E_HIP, "seed kernel failed");
        for (uint32_t k = 0; k < D && rc == GR_OK && curv_pos; ++k)
            if (hipMemcpyAsync(j->q[k], j->p[k], n * sizeof(float), hipMemcpyDeviceToDevice, ctx->compute) != hipSuccess)
                rc = set_err(GR_E_HIP, "state init failed");
    }
    if (rc != GR_OK)
    {
        if (g_err.empty()) set_err(rc, "gr_job_create failed");
        (void)hipGetLastError();
        (void)hipStreamSynchronize(ctx->compute); // queued fills end before the arrays are reused
        job_free(j);
        return rc;
    }
    ctx->children++;
    *out = j;
    return GR_OK;
}

extern "C" int gr_job_destroy(gr_job * j)
{
    if (!j) return GR_OK;
    gr_ctx * c = j->ctx;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->compute);
    if (!j->reads.empty()) (void)hipStreamSynchronize(c->copy); // copies of its columns in flight
    job_free(j);
    c->children--;
    return GR_OK;
}

extern "C" int gr_job_set_launch(gr_job * j, uint32_t block)
{
    if (!j) return set_err(GR_E_INVALID, "null job");
    if (block != 0 && (block % 64 != 0 || block > 256))
        return set_err(GR_E_INVALID, "block must be 0 or a multiple of 64 up to 256");
    j->block = block;
    return GR_OK;
}

extern "C" int gr_job_set_option(gr_job * j, uint32_t option, int64_t value)
{
    if (!j) return set_err(GR_E_INVALID, "null job");
    switch (option)
    {
        case GR_OPT_BLOCK: return gr_job_set_launch(j, uint32_t(value));
        case GR_OPT_PAIRED: j->paired = value ? 1 : 0; return GR_OK;
        case GR_OPT_CU_WAVES:
            if (value < 0 || value > 32) return set_err(GR_E_INVALID, "CU waves must be 0 .. 32");
            j->cu_waves = uint32_t(value);
            return GR_OK;
        case GR_OPT_SPECIALIZE: j->specialize = value ? 1 : 0; return GR_OK;
        case GR_OPT_REUSE: j->reuse = value ? 1 : 0; return GR_OK;
        case GR_OPT_QUAD:
            if (value < 0 || value >= (1ll << 29)) return set_err(GR_E_INVALID, "quad limit must be 0 .. 2^29 - 1");
            j->quad_max = uint64_t(value);
            return GR_OK;
        case GR_OPT_TILE:
            if (value != 0 && (value < 1024 || value > 65536))
                return set_err(GR_E_INVALID, "tile bytes must be 0 or 1024 .. 65536");
            j->tile_bytes = uint32_t(value);
            return GR_OK;
        case GR_OPT_TILE_HALO:
            if (value < 0 || value > 8) return set_err(GR_E_INVALID, "tile halo must be 0 .. 8");
            j->tile_halo = uint32_t(value);
            return GR_OK;
        default: return set_err(GR_E_INVALID, "unknown option %u", option);
    }
}

// the state arrays of a job (4-byte elements, n each) in a fixed order, as slots so that a
// reorder can swap each with its shadow copy; the index map comes last
static std::vector<void **> state_slots(gr_job * j)
{
    std::vector<void **> a;
    auto add = [&](auto *& p) {
        if (p) a.push_back(reinterpret_cast<void **>(&p));
    };
    for (int k = 0; k < 3; ++k)
    {
        add(j->p[k]);
        add(j->q[k]);
        add(j->c[k]);
        add(j->ld[k]);
    }
    add(j->status);
    add(j->abort_step);
    add(j->total);
    add(j->curv);
    add(j->len);
    for (float *& e : j->expr) add(e);
    for (float *& v : j->vol) add(v);
    add(j->orig);
    return a;
}

template <class T> static int grow(T *& buf, size_t & cap, size_t need)
{
    if (need <= cap) return GR_OK;
    dfree(buf);
    buf = nullptr;
    cap = 0;
    int rc = dalloc(&buf, need * sizeof(T));
    if (rc == GR_OK) cap = need;
    return rc;
}

// pinned staging ring of launch and reorder tables (defined with the batched launches)
static int stage_reserve(gr_ctx * ctx, int k, size_t bytes);
static int stage_commit(gr_ctx * ctx, int k, void * dst, size_t bytes, hipStream_t s);
static int stage_upload(gr_ctx * ctx, void * dst, const void * src, size_t bytes, hipStream_t s);
static_assert(sizeof(gr::SortJob) % 16 == 0 && sizeof(gr::GatherEntry) % 16 == 0 &&
                  sizeof(gr::BatchItem) % 16 == 0,
              "tables staged by launch_stage_copy are whole 16-byte words");

// A count slot for job j (host-coherent pinned memory, device-visible): taken from the free list,
// which grows by 4 KiB pages.
static int alive_slot(gr_job * j)
{
    if (j->h_alive) return GR_OK;
    gr_ctx * c = j->ctx;
    if (c->alive_free.empty())
    {
        void * pg = nullptr;
        GR_HIP(hipHostMalloc(&pg, 4096, hipHostMallocMapped | hipHostMallocCoherent));
        c->alive_pages.push_back(static_cast<uint32_t *>(pg));
        for (int k = 1023; k >= 0; --k) c->alive_free.push_back(static_cast<uint32_t *>(pg) + k);
    }
    uint32_t * h = c->alive_free.back();
    void * d = nullptr;
    GR_HIP(hipHostGetDevicePointer(&d, h, 0));
    c->alive_free.pop_back();
    j->h_alive = h;
    j->d_alive = static_cast<uint32_t *>(d);
    return GR_OK;
}

// Takes the job's pending alive count once its reorder has completed (with wait: blocks until it
// has). Events complete in stream order, so a later reorder's event standing in for an
// overwritten one also proves completion.
static int alive_resolve(gr_job * j, bool wait)
{
    if (!j->alive_seq) return GR_OK;
    gr_ctx * c = j->ctx;
    if (j->alive_seq > c->alive_done)
    {
        hipEvent_t e = c->alive_ev[j->alive_seq % gr_ctx::ALIVE_EVENTS];
        if (wait)
            GR_HIP(hipEventSynchronize(e));
        else
        {
            const hipError_t q = hipEventQuery(e);
            if (q == hipErrorNotReady) return GR_OK;
            if (q != hipSuccess) return set_err(GR_E_HIP, "alive count event: %s", hipGetErrorString(q));
        }
        // the event stands for reorder alive_seq or a later one (which implies it)
        c->alive_done = std::max(c->alive_done, j->alive_seq);
    }
    const uint32_t a = *static_cast<volatile uint32_t *>(j->h_alive);
    j->run_n = std::min<uint64_t>(j->run_n, a);
    j->alive_seq = 0;
    return GR_OK;
}

// Reorders several jobs of one context by the Morton code of each particle's current cell
// (stopped particles last): one counting sort over all particles (buckets are per job, which
// keeps every job in its own range) and one gather launch for every state array of every job,
// each gathered into its shadow copy, which then takes the array's place.
// Only a job's live range [0, run_n) takes part (later slots hold stopped particles, which
// stay where they are): when it is at most half the job, the gathered part is copied back
// into the arrays; otherwise the stopped tail is copied into the shadows before the swap.
extern "C" int gr_jobs_reorder(gr_job * const * jobs, uint32_t njobs)
{
    if (njobs && !jobs) return set_err(GR_E_INVALID, "gr_jobs_reorder: null argument");
    std::vector<gr_job *> all;
    gr_ctx * ctx = nullptr;
    for (uint32_t i = 0; i < njobs; ++i)
    {
        gr_job * j = jobs[i];
        if (!j) return set_err(GR_E_INVALID, "gr_jobs_reorder: null job %u", i);
        if (ctx && j->ctx != ctx) return set_err(GR_E_INVALID, "gr_jobs_reorder: jobs of different contexts");
        ctx = j->ctx;
        if (j->n >= 2) all.push_back(j);
    }
    if (all.empty()) return GR_OK;
    if (all.size() > 0xffffu) return set_err(GR_E_INVALID, "gr_jobs_reorder: too many jobs");
    GR_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->compute;
    int rc;
    // the previous reorder's counts first: keeps the host at most one reorder ahead of the
    // device, so the counts arrive while they still bound the launches being queued
    std::vector<gr_job *> js;
    for (gr_job * j : all)
    {
        if ((rc = alive_resolve(j, true)) != GR_OK) return rc;
        if (j->run_n >= 2) js.push_back(j); // fewer live slots: nothing to regroup
    }
    if (js.empty()) return GR_OK;
    for (gr_job * j : js)
        if ((rc = alive_slot(j)) != GR_OK) return rc;
    if (!ctx->alive_ev[0])
        for (hipEvent_t & e : ctx->alive_ev) GR_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    // index maps and shadow arrays on first use
    for (gr_job * j : js)
    {
        if (!j->orig)
        {
            if ((rc = pool_alloc(ctx, &j->orig, size_t(j->n) * 4)) != GR_OK) return rc;
            if (gr::launch_iota(j->orig, j->n, s) != 0) return set_err(GR_E_HIP, "index map kernel launch failed");
        }
        const size_t na = state_slots(j).size();
        while (j->shadow.size() < na)
        {
            void * b = nullptr;
            if ((rc = pool_alloc(ctx, &b, size_t(j->n) * 4)) != GR_OK) return rc;
            j->shadow.push_back(b);
        }
    }
    // job table
    std::vector<gr::SortJob> tab(js.size());
    uint64_t total = 0;
    for (size_t k = 0; k < js.size(); ++k)
    {
        gr_job * j = js[k];
        gr::SortJob & t = tab[k];
        std::memset(&t, 0, sizeof(t));
        for (int d = 0; d < 3; ++d)
        {
            t.p[d] = j->p[d];
            t.off[d] = j->d.local_coords ? 0.f : float(j->d.offset[d]);
            t.inv_h[d] = j->d.grid_scale[d] > 0.f ? 1.f / j->d.grid_scale[d] : 1.f;
            t.grid[d] = std::max(1u, j->grid[d]);
        }
        t.status = j->status;
        t.alive_out = j->d_alive;
        t.base = total;
        t.n = uint32_t(j->run_n);
        t.dim = j->d.dim;
        total += j->run_n;
    }
    if (total > ctx->scap)
    {
        for (int k = 0; k < 2; ++k)
        {
            dfree(ctx->skeys[k]);
            dfree(ctx->svals[k]);
            ctx->skeys[k] = nullptr;
            ctx->svals[k] = nullptr;
        }
        ctx->scap = 0;
        for (int k = 0; k < 2; ++k)
            if ((rc = dalloc(&ctx->skeys[k], total * sizeof(uint64_t))) != GR_OK ||
                (rc = dalloc(&ctx->svals[k], total * sizeof(uint32_t))) != GR_OK)
                return rc;
        ctx->scap = total;
    }
    // Buckets: the Morton code of the cell (a coordinate along an axis of g cells needs
    // bitlen(g - 1) bits, at most 10 in 3D and 16 in 2D; a code dim times that) with its low
    // `shift` bits dropped, so that a job has at most 2^18 cell buckets, all jobs at most 2^24,
    // and the largest live range about 16 particles per bucket: a wavefront of 64 particles then
    // spans a few buckets, blocks of neighbouring cells, as close together as a full sort would
    // place them, while coarser buckets mean fewer counters to clear and scan and more lanes of
    // a wavefront adding to the same counter (one atomic for them). One more bucket per job
    // takes its stopped particles.
    uint32_t mbits = 1;
    for (const gr::SortJob & t : tab)
        for (uint32_t d = 0; d < t.dim; ++d)
        {
            uint32_t b = 0;
            while (b < (t.dim == 3 ? 10u : 16u) && (t.grid[d] - 1) >> b) ++b;
            mbits = std::max(mbits, t.dim * b);
        }
    uint64_t nmax = 0;
    for (gr_job * j : js) nmax = std::max<uint64_t>(nmax, j->run_n);
    uint32_t bbits = 18;
    while (bbits > 1 && ((uint64_t(js.size()) << bbits) > (uint64_t(1) << 24) || (uint64_t(16) << bbits) > nmax * 2))
        --bbits;
    const uint32_t shift = mbits > bbits ? mbits - bbits : 0;
    const uint32_t nb = (1u << (mbits - shift)) + 1;
    const uint64_t nbt = uint64_t(js.size()) * nb;
    if (nbt > ctx->bcap)
    {
        dfree(ctx->bcount);
        dfree(ctx->bstart);
        ctx->bcount = ctx->bstart = nullptr;
        ctx->bcap = 0;
        if ((rc = dalloc(&ctx->bcount, nbt * sizeof(uint32_t))) != GR_OK ||
            (rc = dalloc(&ctx->bstart, nbt * sizeof(uint32_t))) != GR_OK)
            return rc;
        ctx->bcap = nbt;
    }
    uint32_t * keys = reinterpret_cast<uint32_t *>(ctx->skeys[0]);
    size_t need = 0;
    if (gr::launch_bucket_sort(ctx->d_sjobs, uint32_t(tab.size()), total, shift, nb, keys, ctx->svals[0],
                               ctx->bcount, ctx->bstart, nullptr, need, ctx->svals[1], s) != 0)
        return set_err(GR_E_HIP, "bucket sort sizing failed");
    if (need > ctx->stmp_bytes || !ctx->stmp)
    {
        dfree(ctx->stmp);
        ctx->stmp = nullptr;
        ctx->stmp_bytes = 0;
        need = std::max<size_t>(need, 256);
        if ((rc = dalloc(&ctx->stmp, need)) != GR_OK) return rc;
        ctx->stmp_bytes = need;
    }
    if ((rc = grow(ctx->d_sjobs, ctx->sjobs_cap, tab.size())) != GR_OK) return rc;
    if ((rc = stage_upload(ctx, ctx->d_sjobs, tab.data(), tab.size() * sizeof(gr::SortJob), s)) != GR_OK) return rc;
    size_t tb = ctx->stmp_bytes;
    if (gr::launch_bucket_sort(ctx->d_sjobs, uint32_t(tab.size()), total, shift, nb, keys, ctx->svals[0],
                               ctx->bcount, ctx->bstart, ctx->stmp, tb, ctx->svals[1], s) != 0)
        return set_err(GR_E_HIP, "bucket sort failed");
    ++ctx->alive_seq;
    GR_HIP(hipEventRecord(ctx->alive_ev[ctx->alive_seq % gr_ctx::ALIVE_EVENTS], s));
    for (gr_job * j : js) j->alive_seq = ctx->alive_seq;
    // gather the live part of every state array (and the index map) into its shadow; then
    // either copy it back (second table) or complete the shadow with the stopped tail and swap
    std::vector<gr::GatherEntry> g, back;
    std::vector<char> swap_mode(js.size(), 0);
    uint32_t maxn = 0, maxb = 0;
    for (size_t k = 0; k < js.size(); ++k)
    {
        gr_job * j = js[k];
        const uint32_t m = uint32_t(j->run_n), n = uint32_t(j->n);
        swap_mode[k] = 2ull * m > n;
        const std::vector<void **> slots = state_slots(j);
        for (size_t a = 0; a < slots.size(); ++a)
        {
            uint32_t * arr = static_cast<uint32_t *>(*slots[a]);
            uint32_t * sh = static_cast<uint32_t *>(j->shadow[a]);
            g.push_back(gr::GatherEntry{arr, sh, ctx->svals[1] + tab[k].base, m, 0u});
            if (swap_mode[k] && m < n) g.push_back(gr::GatherEntry{arr + m, sh + m, nullptr, n - m, 0u});
            if (!swap_mode[k]) back.push_back(gr::GatherEntry{sh, arr, nullptr, m, 0u});
        }
        maxn = std::max(maxn, swap_mode[k] ? std::max(m, n - m) : m);
        if (!swap_mode[k]) maxb = std::max(maxb, m);
    }
    if (g.size() + back.size() > 65535) return set_err(GR_E_INVALID, "gr_jobs_reorder: too many arrays in one call");
    if ((rc = grow(ctx->d_gtab, ctx->gtab_cap, g.size() + back.size())) != GR_OK) return rc;
    g.insert(g.end(), back.begin(), back.end());
    if ((rc = stage_upload(ctx, ctx->d_gtab, g.data(), g.size() * sizeof(gr::GatherEntry), s)) != GR_OK) return rc;
    const uint32_t ng = uint32_t(g.size() - back.size());
    if (gr::launch_gather_multi(ctx->d_gtab, ng, maxn, s) != 0 ||
        (!back.empty() && gr::launch_gather_multi(ctx->d_gtab + ng, uint32_t(back.size()), maxb, s) != 0))
        return set_err(GR_E_HIP, "gather kernel failed");
    for (size_t k = 0; k < js.size(); ++k)
    {
        gr_job * j = js[k];
        if (swap_mode[k])
        {
            const std::vector<void **> slots = state_slots(j);
            for (size_t a = 0; a < slots.size(); ++a) std::swap(*slots[a], j->shadow[a]);
        }
        j->reordered = true;
        j->reorder_col0 = j->saved;
    }
    return GR_OK;
}

extern "C" int gr_job_reorder(gr_job * j)
{
    if (!j) return set_err(GR_E_INVALID, "null job");
    return gr_jobs_reorder(&j, 1);
}

extern "C" int gr_job_set_particle_data(gr_job * j, uint32_t kind, uint32_t slot, const float * host)
{
    if (!j || !host) return set_err(GR_E_INVALID, "gr_job_set_particle_data: null argument");
    if (slot >= 3) return set_err(GR_E_INVALID, "particle-data slot %u out of range", slot);
    GR_HIP(hipSetDevice(j->ctx->device));
    float ** dst;
    size_t count;
    if (kind == 0)
    {
        dst = &j->uplift[slot];
        count = size_t(j->n) * j->d.dim;
    }
    else if (kind >= 1 && kind <= GR_MAX_CONSTANTS)
    {
        dst = &j->cst[slot][kind - 1];
        count = size_t(j->n);
    }
    else
        return set_err(GR_E_INVALID, "particle-data kind %u out of range", kind);
    if (!*dst)
    {
        int rc = zalloc(dst, count ? count : 1);
        if (rc != GR_OK) return rc;
    }
    if (count)
    {
        // ordered behind queued kernels on the compute stream; pageable source -> blocking copy
        GR_HIP(hipStreamSynchronize(j->ctx->compute));
        GR_HIP(hipMemcpy(*dst, host, count * sizeof(float), hipMemcpyHostToDevice));
    }
    return GR_OK;
}

static void view(gr::FieldView & v, const gr_field * f, uint32_t back, uint32_t front)
{
    std::memset(&v, 0, sizeof(v));
    if (!f) return;
    v.back = f->slots[back];
    v.front = f->slots[front];
    v.nx = f->d.nx;
    v.ny = f->d.ny;
    v.nz = f->d.nz;
    v.cstride = f->cstride;
    v.ncomp = f->d.ncomp;
    v.address = f->d.address;
    v.dtype = f->d.dtype;
    {
        const uint32_t dim = f->d.ncomp == 3 ? 3 : 2;
        const uint64_t bytes = gr::pair_nodes(f->d.nx, f->d.ny, f->d.nz, dim) * (f->d.ncomp == 3 ? 8 : 4) *
                               dtype_size(f->d.dtype);
        v.pair32 = bytes < (1ull << 32) ? 1u : 0u;
    }
    v.zlev = f->zlev;
    v.zlut = f->zlut;
    v.nlut = f->nlut;
    v.zlut_z0 = f->zlut_z0;
    v.zlut_scale = f->zlut_scale;
}

// Validates the window, fills the kernel parameters, makes the compute stream wait for the
// slices the window reads and builds the paired copy when the paired layout is used.
// makes the compute stream `s` wait for the last upload of a slot (once per upload: later work on
// the stream is ordered behind the first wait)
static int f_wait(gr_field * f, uint32_t slot, hipStream_t s)
{
    if (f && slot < f->d.nslots && f->has_ready[slot] && !f->waited[slot])
    {
        GR_HIP(hipStreamWaitEvent(s, f->ready[slot], 0));
        f->waited[slot] = 1;
    }
    return GR_OK;
}

// With allow_alloc = false nothing is allocated: the time-paired copy is used only when it is
// already resident, else the window's slots are sampled separately (the same results).
static int prepare_advance(gr_job * j, const gr_window * w, gr::StepParams & P, uint32_t & mode,
                           uint64_t & nsaves, bool allow_alloc = true)
{
    if (!j || !w || !w->wind) return set_err(GR_E_INVALID, "gr_job_advance: null argument");
    const gr_job_desc & d = j->d;
    gr_field * wf = w->wind;
    if (wf->d.ncomp != d.dim) return set_err(GR_E_INVALID, "wind field has %u components, job is %uD", wf->d.ncomp, d.dim);
    if (d.dim == 2 && wf->d.nz != 1) return set_err(GR_E_INVALID, "2D job needs nz = 1");
    if (w->wind_back >= wf->d.nslots || w->wind_front >= wf->d.nslots)
        return set_err(GR_E_INVALID, "wind slot out of range");
    if (d.nvol && (!w->volumes || w->volumes->d.ncomp < d.nvol || w->vol_back >= w->volumes->d.nslots ||
                   w->vol_front >= w->volumes->d.nslots))
        return set_err(GR_E_INVALID, "volume field missing or too small");
    if (w->topography && (w->topography->d.nx != wf->d.nx || w->topography->d.ny != wf->d.ny))
        return set_err(GR_E_INVALID, "topography grid does not match the wind grid");
    GR_HIP(hipSetDevice(j->ctx->device));

    std::memset(&P, 0, sizeof(P));
    P.n = uint32_t(j->n);
    P.work = j->ctx->dwork;
    P.nsteps = w->nsteps;
    uint32_t fl = 0;
    const bool dyn = d.windfield_mode == 1 || d.uplift_mode == 2 || d.constants_mode == 2 || d.volume_mode == 2;
    if (dyn) fl |= gr::F_DYNAMIC_ANY;
    if (d.windfield_mode == 1) fl |= gr::F_WIND_DYNAMIC;
    if (d.reverse) fl |= gr::F_REVERSE;
    if (d.time64) fl |= gr::F_TIME64;
    if (d.kahan) fl |= gr::F_KAHAN;
    if (d.local_coords) fl |= gr::F_LOCAL;
    if (d.dim == 2 && d.sphere_radius > 0.f) fl |= gr::F_SPHERE;
    if (d.dim == 3 && w->topography) fl |= gr::F_TOPO;
    if (d.dim == 3 && d.uplift_mode == 1) fl |= gr::F_UPLIFT_CONST;
    if (d.dim == 3 && d.uplift_mode == 2) fl |= gr::F_UPLIFT_DYN;
    if (d.constants_mode == 1) fl |= gr::F_CONST_CONST;
    if (d.constants_mode == 2) fl |= gr::F_CONST_DYN;
    if (d.nvol && d.volume_mode == 1) fl |= gr::F_VOL_CONST;
    if (d.nvol && d.volume_mode == 2) fl |= gr::F_VOL_DYN;
    P.flags = fl;
    P.border = d.dim == 2 ? d.border : GR_BLOCK;
    P.curvature = d.curvature;
    P.abort_mode = d.abort_mode;
    P.weights = d.weights;
    P.nprog = d.nprog;
    P.nvol = d.nvol;
    P.nconst = d.nconst;
    P.dt = d.dt;
    {
        const float two_dt = 2.f * d.dt, dtdt = d.dt * d.dt; // the kernel's float denominators
        P.inv_2dt = 1.0 / double(two_dt);
        P.inv_dtdt = 1.0 / double(dtdt);
    }
    P.sphere_radius = d.sphere_radius;
    P.max_length = d.max_length;
    for (int k = 0; k < 3; ++k)
    {
        P.h[k] = d.grid_scale[k];
        P.inv_h[k] = 1.f / d.grid_scale[k];
        P.off[k] = d.local_coords ? 0.f : float(d.offset[k]);
        P.dom[k] = d.domain[k];
        P.off64[k] = d.offset[k];
    }
    P.tb = w->time_back;
    P.tf = w->time_front;
    P.t_start = w->time_start;
    P.tb64 = w->time_back64;
    P.tf64 = w->time_front64;
    P.t_start64 = w->time_start64;
    P.counter_start = w->counter_start;
    P.save_interval = d.save_interval;
    P.save_first = ~0ull;
    nsaves = 0;
    if (d.save_interval)
    {
        const uint64_t r = w->counter_start % d.save_interval;
        const uint64_t first = r ? d.save_interval - r : 0;
        if (first < w->nsteps)
        {
            P.save_first = first;
            P.save_g0 = (w->counter_start + first) / d.save_interval;
            nsaves = (w->nsteps - 1 - first) / d.save_interval + 1;
            if (nsaves > j->cap)
                return set_err(GR_E_STATE, "window saves %llu columns but the ring holds %u",
                               (unsigned long long)nsaves, j->cap);
        }
    }
    P.ring_cap = j->cap;
    for (int k = 0; k < 3; ++k)
    {
        P.p[k] = j->p[k];
        P.q[k] = j->q[k];
        P.c[k] = j->c[k];
        P.ld[k] = j->ld[k];
    }
    P.status = j->status;
    P.abort_step = j->abort_step;
    P.total = j->total;
    P.curv = j->curv;
    P.len = j->len;
    for (int e = 0; e < GR_MAX_PROGRAMS; ++e) P.expr[e] = j->expr[e];
    for (int k = 0; k < GR_MAX_VOLUMES; ++k) P.vol[k] = j->vol[k];
    P.ring_pos = j->ring_pos;
    P.ring_curv = j->ring_curv;
    P.ring_expr = j->ring_expr;
    P.ring_vol = j->ring_vol;
    view(P.wind, wf, w->wind_back, d.windfield_mode == 1 ? w->wind_front : w->wind_back);
    if (d.nvol) view(P.vols, w->volumes, w->vol_back, w->vol_front);
    if (w->topography) view(P.topo, w->topography, 0, 0);
    if (d.dim == 3 && d.uplift_mode != GR_MODE_OFF)
    {
        const uint32_t b = d.uplift_mode == 1 ? 0 : w->uplift_back;
        const uint32_t f = d.uplift_mode == 1 ? 0 : w->uplift_front;
        if (b >= 3 || f >= 3 || !j->uplift[b] || !j->uplift[f])
            return set_err(GR_E_STATE, "uplift data for slots %u/%u not set", b, f);
        P.lift_b = j->uplift[b];
        P.lift_f = j->uplift[f];
    }
    for (uint32_t k = 0; k < d.nconst; ++k)
    {
        const uint32_t b = d.constants_mode == 2 ? w->const_back : 0;
        const uint32_t f = d.constants_mode == 2 ? w->const_front : 0;
        if (b >= 3 || f >= 3 || !j->cst[b][k] || !j->cst[f][k])
            return set_err(GR_E_STATE, "constant %u for slots %u/%u not set", k, b, f);
        P.cst_b[k] = j->cst[b][k];
        P.cst_f[k] = j->cst[f][k];
    }
    for (uint32_t e = 0; e < d.nprog; ++e)
        if (gr::lower_program(d.prog[e], P.prog[e], P.refmask) != GR_OK)
            return set_err(GR_E_INVALID, "program %u is malformed", e);
    for (uint32_t e = 0; e < d.nprog; ++e)
        if (P.prog[e].n == 1 && ((P.prog[e].code[0] >> 4) & 0xfu) == gr::DC_DOT) P.dotmask |= 1u << e;
    if (P.dotmask & 1u)
    {
        // slots 4 5 6 are u v w: arg holds the slot pairs (4, 4), (5, 5) [, (6, 6)], 3 bits each
        const uint32_t w = P.prog[0].code[0], terms = w & 3u, arg = w >> 8;
        const uint32_t uvw = 4u | 4u << 3 | 5u << 6 | 5u << 9 | 6u << 12 | 6u << 15;
        const uint32_t mask = terms == 3 ? 0x3ffffu : 0xfffu;
        if ((terms == 3 || d.dim == 2) && (arg & mask) == (uvw & mask)) P.dotmask |= gr::DOT_UVW;
    }

    hipStream_t s = j->ctx->compute;
    // wait for the slices this window reads
    auto wait_ready = [&](gr_field * f, uint32_t slot) -> int { return f_wait(f, slot, s); };
    int rc;
    if ((rc = wait_ready(wf, w->wind_back)) || (rc = wait_ready(wf, w->wind_front))) return rc;
    if (d.nvol && ((rc = wait_ready(w->volumes, w->vol_back)) || (rc = wait_ready(w->volumes, w->vol_front))))
        return rc;
    if (w->topography && (rc = wait_ready(w->topography, 0))) return rc;

    mode = j->full ? gr::MODE_FULL
                                  : ((d.windfield_mode == 1 && j->paired) ? gr::MODE_PAIRED : gr::MODE_SEPARATE);
    if (mode == gr::MODE_PAIRED && !allow_alloc)
    {
        for (auto & pr : wf->pairs)
            if (pr.valid && pr.b == w->wind_back && pr.f == w->wind_front) P.wind.pair = pr.buf;
        if (!P.wind.pair) mode = gr::MODE_SEPARATE;
    }
    else if (mode == gr::MODE_PAIRED)
        if ((rc = field_pair(wf, w->wind_back, w->wind_front, s, &P.wind.pair)) != GR_OK) return rc;
    P.orig = j->reordered ? j->orig : nullptr;
    j->grid[0] = wf->d.nx;
    j->grid[1] = wf->d.ny;
    j->grid[2] = wf->d.nz;

    return GR_OK;
}

// A slot read by a launch is in use until the launch's end event (borrowed from the context's
// kernel-timing events, no record of its own). Those events are recorded again for later
// launches once folded; an upload that waits on one then waits for a later launch of the same
// stream, which implies the one that read the slot.
static void mark_inuse(gr_field * f, uint32_t slot, hipEvent_t done)
{
    if (f && done)
    {
        f->inuse[slot] = done;
        f->has_inuse[slot] = 1;
    }
}

// after the launch: later uploads into the slots read by the window wait for its end event
// `done` (the batched path marks each slot once per launch itself: done = nullptr)
static int finish_advance(gr_job * j, const gr_window * w, uint64_t nsaves, hipEvent_t done)
{
    if (!done)
    {
        j->saved += nsaves;
        return GR_OK;
    }
    mark_inuse(w->wind, w->wind_back, done);
    mark_inuse(w->wind, w->wind_front, done);
    if (j->d.nvol)
    {
        mark_inuse(w->volumes, w->vol_back, done);
        mark_inuse(w->volumes, w->vol_front, done);
    }
    if (w->topography) mark_inuse(w->topography, 0, done);
    j->saved += nsaves;
    return GR_OK;
}

static void time_table(const gr::StepParams & P, const uint32_t * edges, float * out);
static int wait_ring_reads(gr_job * j, uint64_t nsaves, hipStream_t s);

// A launch runs a quad of lanes per particle when its features are a specialised plain set in
// the time-paired layout (a paired copy smaller than 4 GiB) and its live range is at most the
// job's quad limit.
static bool use_quad(const gr_job * j, const gr::StepParams & P, int spec, uint32_t mode)
{
    return spec >= 0 && mode == gr::MODE_PAIRED && P.wind.pair32 && j->quad_max && j->run_n <= j->quad_max;
}

// A 3D time-paired launch runs the field-tile kernel when the job asks for a tile, the launch is
// not a quad one, the padded grid's node coordinates fit 16 bits (the box reduction packs them)
// and the interpreter's columns (rounded up to 16 bytes) plus the tile fit 64 KiB of LDS. Then
// sets P.tile_* and raises *lds (the launch's dynamic LDS) to cover the tile area.
static bool tile_setup(const gr_job * j, gr::StepParams & P, uint32_t mode, bool quad, uint32_t block,
                       uint32_t * lds)
{
    if (!j->tile_bytes || quad || mode != gr::MODE_PAIRED || j->d.dim != 3 || P.wind.nx + 3 > 0xffffu ||
        P.wind.ny + 3 > 0xffffu || P.wind.nz + 3 > 0xffffu)
        return false;
    const uint32_t lds_cols = j->d.nprog ? j->expr_cols : 0u;
    const uint32_t off = (lds_cols * block * uint32_t(sizeof(float)) + 15u) & ~15u;
    if (off + j->tile_bytes > 65536u) return false;
    const uint32_t nb = P.wind.dtype == GR_F32 ? 32u : 16u; // bytes per 3D paired node
    P.tile_off = off;
    P.tile_cap = (j->tile_bytes - gr::TILE_AREA_HEADER) / nb;
    P.tile_halo = j->tile_halo;
    P.tile_stats = j->ctx->dtile;
    *lds = std::max(*lds, off + j->tile_bytes);
    return true;
}

extern "C" int gr_job_advance(gr_job * j, const gr_window * w)
{
    if (!j || !w || !w->wind) return set_err(GR_E_INVALID, "gr_job_advance: null argument");
    if (w->nsteps == 0) return GR_OK;
    gr::StepParams P;
    uint32_t mode = 0;
    uint64_t nsaves = 0;
    int rc = prepare_advance(j, w, P, mode, nsaves);
    if (rc != GR_OK) return rc;
    hipStream_t s = j->ctx->compute;
    if ((rc = alive_resolve(j, false)) != GR_OK) return rc;
    if ((rc = wait_ring_reads(j, nsaves, s)) != GR_OK) return rc;
    // the launch's time table through the pinned staging ring (stream-ordered device copy)
    gr_ctx * ctx = j->ctx;
    const size_t tbytes = size_t(P.nsteps) * 4 * sizeof(float);
    const int k = ctx->stage_i;
    if ((rc = stage_reserve(ctx, k, tbytes)) != GR_OK) return rc;
    if ((rc = grow(ctx->d_stage, ctx->d_stage_cap, tbytes)) != GR_OK) return rc;
    time_table(P, nullptr, static_cast<float *>(ctx->stage[k].h));
    if ((rc = stage_commit(ctx, k, ctx->d_stage, tbytes, s)) != GR_OK) return rc;
    P.ttab = reinterpret_cast<const float *>(ctx->d_stage);
    GR_HIP(hipEventRecord(j->ev0, s));
    if ((rc = kt_mark(j->ctx, s, false)) != GR_OK) return rc;
    // wide: at least two rounds of the wavefronts the device holds at three per SIMD; narrow:
    // every wavefront fits at two per SIMD
    const uint64_t cus = uint64_t(std::max(1, j->ctx->cus));
    const bool wide = j->run_n >= 2ull * cus * 4 * 3 * 64;
    const bool narrow = j->reuse || j->run_n <= cus * 4 * 2 * 64;
    const uint32_t lds_cols = j->d.nprog ? (mode == gr::MODE_FULL ? gr::EXPR_REF_SLOTS : 0u) + j->expr_cols : 0u;
    const uint32_t blk = (j->block == 0 || j->block > 256) ? 256u : j->block;
    const int spec = j->specialize ? gr::lean_spec(P, mode) : -1;
    const bool quad = use_quad(j, P, spec, mode);
    uint32_t lds = gr::launch_lds_bytes(blk, lds_cols, j->cu_waves, j->ctx->lds_per_cu);
    const bool tile = tile_setup(j, P, mode, quad, blk, &lds);
    if (gr::launch_advance(P, j->d.dim, j->d.integrator, mode, spec, blk, j->ctx->nxcd, j->run_n, wide, narrow,
                           quad, lds, s, tile) != 0)
        return set_err(GR_E_HIP, "step kernel launch failed: %s", hipGetErrorString(hipGetLastError()));
    hipEvent_t done = nullptr;
    if ((rc = kt_mark(j->ctx, s, true, &done)) != GR_OK) return rc;
    GR_HIP(hipEventRecord(j->ev1, s));
    j->timed = true;
    return finish_advance(j, w, nsaves, done);
}

extern "C" int gr_job_advance_noalloc(gr_job * j, const gr_window * w)
{
    if (!j || !w || !w->wind) return set_err(GR_E_INVALID, "gr_job_advance_noalloc: null argument");
    if (w->nsteps == 0) return GR_OK;
    if (w->nsteps > GR_NOALLOC_MAX_STEPS)
        return set_err(GR_E_INVALID, "gr_job_advance_noalloc: %u steps exceed the scratch of %u steps",
                       w->nsteps, unsigned(GR_NOALLOC_MAX_STEPS));
    gr::StepParams P;
    uint32_t mode = 0;
    uint64_t nsaves = 0;
    int rc = prepare_advance(j, w, P, mode, nsaves, false);
    if (rc != GR_OK) return rc;
    gr_ctx * ctx = j->ctx;
    hipStream_t s = ctx->compute;
    if ((rc = alive_resolve(j, false)) != GR_OK) return rc;
    if ((rc = wait_ring_reads(j, nsaves, s)) != GR_OK) return rc;
    float * tab = static_cast<float *>(gr::noalloc_scratch(ctx->device));
    if (!tab) return set_err(GR_E_STATE, "gr_job_advance_noalloc: no scratch on device %d", ctx->device);
    // the whole table of this launch is written before the step kernel reads it (stream order)
    if (gr::launch_time_table(P, tab, s) != 0)
        return set_err(GR_E_HIP, "time-table kernel launch failed: %s", hipGetErrorString(hipGetLastError()));
    P.ttab = tab;
    const uint64_t cus = uint64_t(std::max(1, ctx->cus));
    const bool wide = j->run_n >= 2ull * cus * 4 * 3 * 64;
    const bool narrow = j->reuse || j->run_n <= cus * 4 * 2 * 64;
    const uint32_t lds_cols = j->d.nprog ? (mode == gr::MODE_FULL ? gr::EXPR_REF_SLOTS : 0u) + j->expr_cols : 0u;
    const uint32_t blk = (j->block == 0 || j->block > 256) ? 256u : j->block;
    const int spec = j->specialize ? gr::lean_spec(P, mode) : -1;
    const bool quad = use_quad(j, P, spec, mode);
    uint32_t lds = gr::launch_lds_bytes(blk, lds_cols, j->cu_waves, ctx->lds_per_cu);
    const bool tile = tile_setup(j, P, mode, quad, blk, &lds);
    if (gr::launch_advance(P, j->d.dim, j->d.integrator, mode, spec, blk, ctx->nxcd, j->run_n, wide, narrow,
                           quad, lds, s, tile) != 0)
        return set_err(GR_E_HIP, "step kernel launch failed: %s", hipGetErrorString(hipGetLastError()));
    GR_HIP(hipEventRecord(ctx->noalloc_done, s));
    j->timed = false;
    return finish_advance(j, w, nsaves, ctx->noalloc_done);
}

// --------------------------------------------------------------------------------------------
// batches
// --------------------------------------------------------------------------------------------


// Steps per window of a span under the reference's host schedule (tpp:852-994): FP32 (or FP64)
// time, max_windfield_time accumulated by += DataTimeDistance, stop at counter >= max_sim / dt
// or when slice c + 1 does not exist.
static void plan_span(const gr_job_desc & d, const gr_span & sp, uint32_t nslices,
                      std::vector<uint32_t> & win_steps, gr_span_end * end = nullptr)
{
    win_steps.clear();
    float mwt = d.dtd;
    double mwt64 = d.dtd;
    for (uint32_t c = 0; c < sp.first_window; ++c)
    {
        mwt += d.dtd;
        mwt64 += double(d.dtd);
    }
    const bool capped = sp.max_sim_time > 0.f;
    const float maxs_f = capped ? sp.max_sim_time / d.dt : 0.f;
    float t = sp.time_start;
    double t64 = sp.time_start64;
    uint64_t counter = sp.counter_start;
    uint64_t total = 0;
    uint32_t c = sp.first_window;
    while (total < sp.max_steps && c + 1 < nslices)
    {
        uint32_t k = 0;
        bool window_end = false;
        while (total + k < sp.max_steps && (!capped || float(counter + k) < maxs_f))
        {
            ++k;
            t += d.dt;
            t64 += double(d.dt);
            if (d.time64 ? (t64 >= mwt64) : (t >= mwt))
            {
                window_end = true;
                break;
            }
        }
        if (k == 0) break;
        win_steps.push_back(k);
        total += k;
        counter += k;
        if (!window_end) break;
        mwt += d.dtd;
        mwt64 += double(d.dtd);
        ++c;
    }
    if (end)
    {
        end->steps = uint32_t(total);
        end->window = c;
        end->time = t;
        end->time64 = t64;
        end->counter = counter;
        end->done = (c + 1 >= nslices) || (capped && !(float(counter) < maxs_f)) ? 1u : 0u;
    }
}

// Pinned host staging of launch tables: a ring of NSTAGE blocks, so the host fills the next
// launch's tables while earlier ones are still queued. A block is handed to the device by
// launch_stage_copy on the compute stream (a kernel, ordered with the kernels around it), and
// reused only once that copy has completed.
static int stage_reserve(gr_ctx * ctx, int k, size_t bytes)
{
    gr_ctx::Stage & st = ctx->stage[k];
    if (st.pending)
    {
        GR_HIP(hipEventSynchronize(st.done));
        st.pending = false;
    }
    if (bytes <= st.cap) return GR_OK;
    if (st.h) (void)hipHostFree(st.h);
    st.h = st.dh = nullptr;
    st.cap = 0;
    const size_t want = std::max<size_t>(bytes, 64 << 10);
    // coherent: the copy kernel reads what the host wrote since the block's last use, never a
    // line the device cached then
    if (hipHostMalloc(&st.h, want, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
    {
        (void)hipGetLastError();
        return set_err(GR_E_NOMEM, "pinned staging of %zu bytes failed", want);
    }
    if (hipHostGetDevicePointer(&st.dh, st.h, 0) != hipSuccess)
    {
        (void)hipGetLastError();
        (void)hipHostFree(st.h);
        st.h = nullptr;
        return set_err(GR_E_HIP, "pinned staging has no device address");
    }
    st.cap = want;
    // the host only needs to know that the copy has read the staging block: no fence
    if (!st.done) GR_HIP(hipEventCreateWithFlags(&st.done, hipEventDisableTiming | hipEventDisableSystemFence));
    return GR_OK;
}

// Hands staging block k (its first `bytes` bytes, a multiple of 16) to device memory `dst` on
// stream s and advances the ring.
static int stage_commit(gr_ctx * ctx, int k, void * dst, size_t bytes, hipStream_t s)
{
    if (gr::launch_stage_copy(ctx->stage[k].dh, dst, bytes, s) != 0)
        return set_err(GR_E_HIP, "staging copy of %zu bytes failed", bytes);
    GR_HIP(hipEventRecord(ctx->stage[k].done, s));
    ctx->stage[k].pending = true;
    ctx->stage_i = (k + 1) % gr_ctx::NSTAGE;
    return GR_OK;
}

// Copies a small host table (bytes a multiple of 16) to device memory `dst` through the staging
// ring: the caller's buffer is free when this returns, and the copy is ordered on stream s.
static int stage_upload(gr_ctx * ctx, void * dst, const void * src, size_t bytes, hipStream_t s)
{
    const int k = ctx->stage_i;
    int rc = stage_reserve(ctx, k, bytes);
    if (rc != GR_OK) return rc;
    std::memcpy(ctx->stage[k].h, src, bytes);
    return stage_commit(ctx, k, dst, bytes, s);
}

static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// The time weights of every step of a launch, which the step kernels read instead of forming
// them (FP32 time, or FP64 with F_TIME64): t at the step's start, t + dt/2 and t + dt, each
// mapped into the
// step's window (unit_interval, tpp:127-144), and 1 / 0 for "the step starts after time 0"
// (tpp:347); four floats per step. edges: the window edges of a launch that crosses them.
static void time_table(const gr::StepParams & P, const uint32_t * edges, float * out)
{
    const float dt = P.dt, half_dt = 0.5f * dt;
    const bool t64 = (P.flags & gr::F_TIME64) != 0;
    float tb = P.tb, tf = P.tf;
    double tb64 = P.tb64, tf64 = P.tf64;
    float time32 = P.t_start;
    double time64 = P.t_start64;
    uint32_t wi = 0;
    uint32_t edge = P.nwin > 1 ? edges[0] : 0xffffffffu;
    auto unit = [&](float add) -> float {
        if (t64) return float((time64 + double(add) - tb64) / (tf64 - tb64));
        const float x = time32 + add;
        return (x - tb) / (tf - tb);
    };
    for (uint32_t s = 0; s < P.nsteps; ++s)
    {
        if (s == edge)
        {
            ++wi;
            const uint32_t c = P.win_first + wi;
            tb = float(c) * P.dtd;
            tf = float(c + 1) * P.dtd;
            tb64 = double(c) * double(P.dtd);
            tf64 = double(c + 1) * double(P.dtd);
            edge = wi + 1 < P.nwin ? edges[wi] : 0xffffffffu;
        }
        float t = 0.f;
        if (P.flags & gr::F_DYNAMIC_ANY) t = unit(0.f);
        float t1 = t, t2 = t;
        if (P.flags & gr::F_WIND_DYNAMIC)
        {
            t1 = unit(half_dt);
            t2 = unit(dt);
        }
        out[4 * size_t(s) + 0] = t;
        out[4 * size_t(s) + 1] = t1;
        out[4 * size_t(s) + 2] = t2;
        out[4 * size_t(s) + 3] = (t64 ? time64 > 0.0 : time32 > 0.f) ? 1.f : 0.f;
        time32 += dt;
        time64 += double(dt);
    }
}

extern "C" int gr_batch_run(gr_ctx * ctx, const gr_span * spans, uint32_t nspans, uint32_t block,
                            gr_span_end * ends, float * ms)
{
    if (!ctx || (nspans && !spans)) return set_err(GR_E_INVALID, "gr_batch_run: null argument");
    if (block == 0) block = ctx->batch_block;
    if (block % 64 != 0 || block > 256) return set_err(GR_E_INVALID, "block must be a multiple of 64 up to 256");
    GR_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->compute;
    if (!ctx->bev0)
    {
        GR_HIP(hipEventCreateWithFlags(&ctx->bev0, TIMING_EVENT));
        GR_HIP(hipEventCreateWithFlags(&ctx->bev1, TIMING_EVENT));
    }
    std::vector<gr::StepParams> params(nspans);
    std::vector<uint64_t> saves(nspans, 0);
    std::vector<uint32_t> span_steps(nspans, 0);
    std::vector<uint32_t> span_nwin(nspans, 0);
    std::vector<uint32_t> win_steps;
    std::vector<gr_window> first_win(nspans);
    // window tables of the spans that cross window edges: (back, front, pair) per window and the
    // step at which each window ends; placed in the staging block, addresses patched below
    std::vector<const void *> wptr;
    std::vector<uint32_t> wedge;
    std::vector<size_t> wofs(nspans, 0);
    int rc;
    for (uint32_t i = 0; i < nspans; ++i)
    {
        const gr_span & sp = spans[i];
        gr_job * j = sp.job;
        if (!j || !sp.wind) return set_err(GR_E_INVALID, "span %u: null job or field", i);
        if (j->ctx != ctx) return set_err(GR_E_INVALID, "span %u: job of another context", i);
        if (j->full) return set_err(GR_E_INVALID, "span %u: the job uses features outside the batched kernel", i);
        if ((rc = alive_resolve(j, false)) != GR_OK) return rc;
        const gr_job_desc & d = j->d;
        const uint32_t nslices = sp.wind->d.nslots;
        gr_span_end end_i{};
        if (d.windfield_mode == 1)
            plan_span(d, sp, nslices, win_steps, &end_i);
        else
        {
            win_steps.clear(); // constant wind: one window of slot 0, bounded by max_sim_time
            gr_span one = sp;
            one.first_window = 0;
            gr_job_desc dd = d;
            dd.dtd = sp.max_sim_time > 0.f ? sp.max_sim_time : 3.4e38f;
            plan_span(dd, one, 2, win_steps, &end_i);
            if (win_steps.size() > 1) win_steps.resize(1);
            end_i.window = 0;
        }
        if (ends) ends[i] = end_i;
        uint32_t total = 0;
        for (uint32_t k : win_steps) total += k;
        span_steps[i] = total;
        span_nwin[i] = uint32_t(win_steps.size());
        gr_window & w = first_win[i];
        std::memset(&w, 0, sizeof(w));
        const uint32_t c0 = d.windfield_mode == 1 ? sp.first_window : 0;
        w.wind = sp.wind;
        w.wind_back = c0;
        w.wind_front = d.windfield_mode == 1 ? c0 + 1 : c0;
        w.time_back = float(c0) * d.dtd;
        w.time_front = float(c0 + 1) * d.dtd;
        w.time_back64 = double(c0) * double(d.dtd);
        w.time_front64 = double(c0 + 1) * double(d.dtd);
        w.time_start = sp.time_start;
        w.time_start64 = sp.time_start64;
        w.counter_start = sp.counter_start;
        w.nsteps = total;
        gr::StepParams & P = params[i];
        uint32_t mode = 0;
        if (total == 0)
        {
            std::memset(&P, 0, sizeof(P));
            continue;
        }
        if ((rc = prepare_advance(j, &w, P, mode, saves[i])) != GR_OK) return rc;
        if ((rc = wait_ring_reads(j, saves[i], s)) != GR_OK) return rc;
        // the remaining windows of the span: slices c, c + 1 for c = c0 + 1 ...
        const uint32_t nwin = uint32_t(win_steps.size());
        P.nwin = nwin;
        P.win_first = c0;
        P.dtd = d.dtd;
        if (nwin > 1)
        {
            wofs[i] = wedge.size();
            uint32_t edge = 0;
            for (uint32_t k = 0; k < nwin; ++k)
            {
                const uint32_t c = c0 + k;
                edge += win_steps[k];
                wedge.push_back(edge);
                if (f_wait(sp.wind, c, s) != GR_OK || f_wait(sp.wind, c + 1, s) != GR_OK) return GR_E_HIP;
                const void * pr = nullptr;
                if (mode == gr::MODE_PAIRED && (rc = field_pair(sp.wind, c, c + 1, s, &pr)) != GR_OK) return rc;
                wptr.push_back(sp.wind->slots[c]);
                wptr.push_back(sp.wind->slots[c + 1]);
                wptr.push_back(pr);
            }
        }
    }

    // Work items of one wavefront of particles, covering each job's particles [0, run_n) (the
    // rest are stopped). Items go to the queue of a group of compute units by their position in
    // that range (Morton-ordered after gr_job_reorder, so the alive particles of every job spread
    // over all queues), and inside a queue they are ordered by span length class (longest
    // first), field, then position: the wavefronts in flight on one group then sample one
    // compact region of one field, and long spans start early. The list depends only on the
    // spans' jobs, ranges, variants, fields and length classes, so a launch with the same ones
    // reuses the previous list on the device.
    const uint32_t group = std::max(1u, ctx->batch_group);
    const uint32_t nq = std::max<uint32_t>(1, std::min<uint32_t>((ctx->ncu + group - 1) / group, gr::MAX_QUEUES));
    std::vector<uint64_t> key;
    key.push_back(nq);
    std::vector<uint32_t> cls_of(nspans, 0), var_of(nspans, 0), isz(nspans, gr::BATCH_ITEM);
    // Launches of few live particles (in all their spans together, at most a job's quad limit)
    // run a quad of lanes per particle for the jobs with a specialised feature set in the
    // time-paired layout (as single-job launches do), in items of QUAD_ITEM particles.
    uint64_t live_total = 0;
    for (uint32_t i = 0; i < nspans; ++i)
        if (span_steps[i]) live_total += spans[i].job->run_n;
    bool any_quad = false;
    for (uint32_t i = 0; i < nspans; ++i)
    {
        if (!span_steps[i]) continue;
        gr_job * j = spans[i].job;
        for (uint32_t v = span_steps[i]; v; v >>= 1) ++cls_of[i];
        // the item variant (dimension, integrator, storage type, feature set, quad) and the job's
        // serial
        const int spec = j->specialize ? gr::lean_spec(params[i], gr::MODE_PAIRED) : -1;
        var_of[i] = uint32_t(gr::batch_variant(j->d.dim, j->d.integrator, spans[i].wind->d.dtype, spec));
        if (spec >= 0 && params[i].wind.pair && params[i].wind.pair32 && j->quad_max && live_total <= j->quad_max)
        {
            var_of[i] += gr::BATCH_QUAD;
            isz[i] = gr::QUAD_ITEM;
            any_quad = true;
        }
        key.insert(key.end(), {i, j->serial, j->run_n, uint64_t(uintptr_t(spans[i].wind)), cls_of[i], var_of[i]});
    }
    const bool items_dirty = !(ctx->items_valid && key == ctx->items_key);
    if (items_dirty)
    {
        // One counting sort over a composite bucket key: queue, length class (descending),
        // field, then the item's position inside the queue's range in 256 steps -- so the
        // ordering costs O(items) however many jobs share the launch.
        std::vector<uint32_t> cls_vals, fld_rank(nspans, 0);
        std::vector<uintptr_t> fld_vals;
        for (uint32_t i = 0; i < nspans; ++i)
            if (span_steps[i])
            {
                cls_vals.push_back(cls_of[i]);
                fld_vals.push_back(uintptr_t(spans[i].wind));
            }
        std::sort(cls_vals.begin(), cls_vals.end(), std::greater<uint32_t>());
        cls_vals.erase(std::unique(cls_vals.begin(), cls_vals.end()), cls_vals.end());
        std::sort(fld_vals.begin(), fld_vals.end());
        fld_vals.erase(std::unique(fld_vals.begin(), fld_vals.end()), fld_vals.end());
        const uint64_t ncls = std::max<size_t>(1, cls_vals.size()), nfld = std::max<size_t>(1, fld_vals.size());
        // position buckets per queue (fewer when many classes / fields would make the count
        // array large)
        const uint64_t PB = std::max<uint64_t>(1, std::min<uint64_t>(256, (uint64_t(1) << 24) / (uint64_t(nq) * ncls * nfld)));
        const uint64_t nbuckets = uint64_t(nq) * ncls * nfld * PB;
        std::vector<uint32_t> span_base(nspans, 0); // cls/field part of the key, per span
        for (uint32_t i = 0; i < nspans; ++i)
            if (span_steps[i])
            {
                const uint64_t cr = uint64_t(std::find(cls_vals.begin(), cls_vals.end(), cls_of[i]) - cls_vals.begin());
                const uint64_t fr = uint64_t(std::lower_bound(fld_vals.begin(), fld_vals.end(),
                                                              uintptr_t(spans[i].wind)) - fld_vals.begin());
                span_base[i] = uint32_t((cr * nfld + fr) * PB);
            }
        auto bucket_of = [&](uint32_t i, uint64_t f, uint64_t n) -> uint64_t {
            const uint32_t cnt = uint32_t(std::min<uint64_t>(isz[i], n - f));
            const double pos = (double(f) + 0.5 * cnt) / double(n); // in [0, 1)
            const double qp = pos * double(nq);
            const uint32_t q = std::min<uint32_t>(nq - 1, uint32_t(qp));
            const uint64_t pb = std::min<uint64_t>(PB - 1, uint64_t((qp - double(q)) * double(PB)));
            return uint64_t(q) * ncls * nfld * PB + span_base[i] + pb;
        };
        std::vector<uint32_t> count(nbuckets + 1, 0);
        uint64_t total = 0;
        for (uint32_t i = 0; i < nspans; ++i)
        {
            if (!span_steps[i]) continue;
            const uint64_t n = spans[i].job->run_n;
            for (uint64_t f = 0; f < n; f += isz[i]) ++count[bucket_of(i, f, n) + 1];
            total += (n + isz[i] - 1) / isz[i];
        }
        for (uint64_t b = 0; b < nbuckets; ++b) count[b + 1] += count[b];
        ctx->h_items.resize(total);
        ctx->items_quad = 0;
        ctx->qstart.assign(nq + 1, 0u);
        for (uint32_t q = 0; q < nq; ++q) ctx->qstart[q + 1] = count[uint64_t(q + 1) * ncls * nfld * PB];
        for (uint32_t i = 0; i < nspans; ++i)
        {
            if (!span_steps[i]) continue;
            const uint64_t n = spans[i].job->run_n;
            for (uint64_t f = 0; f < n; f += isz[i])
            {
                const uint32_t cnt = uint32_t(std::min<uint64_t>(isz[i], n - f));
                ctx->h_items[count[bucket_of(i, f, n)]++] = gr::BatchItem{i, uint32_t(f), cnt, var_of[i]};
            }
            if (var_of[i] >= gr::BATCH_QUAD) ctx->items_quad += (n + isz[i] - 1) / isz[i];
        }
        ctx->items_key = key;
        ctx->items_valid = false; // until the copy below is queued
    }
    const uint32_t nitems = uint32_t(ctx->h_items.size());

    // staging block: queue data (zeroed counters, queue starts, masks of the queues holding
    // items) | params | window pointers | window edges | items (only when rebuilt); one copy to
    // the device block, one to the item table
    const uint32_t ng = (nq + 31) / 32;
    const size_t qwords = size_t(nq) * gr::QUEUE_STRIDE + nq + 1 + ng;
    const size_t off_p = align256(qwords * sizeof(uint32_t));
    const size_t off_w = align256(off_p + size_t(nspans) * sizeof(gr::StepParams));
    const size_t off_e = off_w + wptr.size() * sizeof(void *);
    // per-step time tables of the spans (four floats per step)
    std::vector<size_t> tofs(nspans, 0);
    size_t tfloats = 0;
    for (uint32_t i = 0; i < nspans; ++i)
    {
        tofs[i] = tfloats;
        tfloats += 4 * size_t(span_steps[i]);
    }
    const size_t off_t = align256(off_e + wedge.size() * sizeof(uint32_t));
    const size_t off_i = align256(off_t + tfloats * sizeof(float));
    const size_t dev_bytes = off_i;
    const size_t host_bytes = off_i + (items_dirty ? size_t(nitems) * sizeof(gr::BatchItem) : 0);
    const int k = ctx->stage_i;
    if ((rc = stage_reserve(ctx, k, host_bytes)) != GR_OK) return rc;
    if ((rc = grow(ctx->d_stage, ctx->d_stage_cap, dev_bytes)) != GR_OK) return rc;
    if (items_dirty && (rc = grow(ctx->d_items, ctx->items_cap, std::max<size_t>(1, nitems))) != GR_OK) return rc;
    char * h = static_cast<char *>(ctx->stage[k].h);
    char * dbase = reinterpret_cast<char *>(ctx->d_stage);
    uint32_t * hq = reinterpret_cast<uint32_t *>(h);
    std::memset(hq, 0, qwords * sizeof(uint32_t));
    std::memcpy(hq + size_t(nq) * gr::QUEUE_STRIDE, ctx->qstart.data(), (nq + 1) * sizeof(uint32_t));
    uint32_t * hmask = hq + size_t(nq) * gr::QUEUE_STRIDE + nq + 1;
    for (uint32_t q = 0; q < nq; ++q)
        if (ctx->qstart[q + 1] > ctx->qstart[q]) hmask[q >> 5] |= 1u << (q & 31u);
    for (uint32_t i = 0; i < nspans; ++i)
        if (span_nwin[i] > 1 && span_steps[i])
        {
            params[i].win_ptr = reinterpret_cast<const void * const *>(dbase + off_w + 3 * wofs[i] * sizeof(void *));
            params[i].win_edge = reinterpret_cast<const uint32_t *>(dbase + off_e + wofs[i] * sizeof(uint32_t));
        }
    for (uint32_t i = 0; i < nspans; ++i)
        if (span_steps[i])
        {
            time_table(params[i], wedge.data() + wofs[i], reinterpret_cast<float *>(h + off_t) + tofs[i]);
            params[i].ttab = reinterpret_cast<const float *>(dbase + off_t) + tofs[i];
        }
    if (nspans) std::memcpy(h + off_p, params.data(), size_t(nspans) * sizeof(gr::StepParams));
    if (!wptr.empty()) std::memcpy(h + off_w, wptr.data(), wptr.size() * sizeof(void *));
    if (!wedge.empty()) std::memcpy(h + off_e, wedge.data(), wedge.size() * sizeof(uint32_t));
    if (items_dirty && nitems) std::memcpy(h + off_i, ctx->h_items.data(), size_t(nitems) * sizeof(gr::BatchItem));
    if (items_dirty && nitems &&
        gr::launch_stage_copy(static_cast<char *>(ctx->stage[k].dh) + off_i, ctx->d_items,
                              size_t(nitems) * sizeof(gr::BatchItem), s) != 0)
        return set_err(GR_E_HIP, "staging copy of the batch items failed");
    if ((rc = stage_commit(ctx, k, ctx->d_stage, dev_bytes, s)) != GR_OK) return rc;
    ctx->items_valid = true;

    // interpreter LDS of the launch: the deepest stack of its jobs' programs (batched jobs use the
    // lean kernel, which stages no reference slots)
    uint32_t lds_cols = 0;
    for (uint32_t i = 0; i < nspans; ++i)
        if (span_steps[i] && spans[i].job->d.nprog) lds_cols = std::max(lds_cols, spans[i].job->expr_cols);
    // narrow: every item's wavefront fits on the device at two per SIMD (and quad items, which
    // only the narrow builds run)
    const bool narrow = ctx->batch_reuse || any_quad || uint64_t(nitems) <= uint64_t(std::max(1, ctx->cus)) * 4 * 2;
    // a narrow launch whose spans all have one specialised variant runs that variant's kernel
    uint32_t uni = 0;
    if (narrow && ctx->batch_uniform)
        for (uint32_t i = 0; i < nspans; ++i)
        {
            if (!span_steps[i]) continue;
            const uint32_t base = var_of[i] >= gr::BATCH_QUAD ? var_of[i] - gr::BATCH_QUAD : var_of[i];
            if (base < 12 || (uni && uni != base))
            {
                uni = 0;
                break;
            }
            uni = base;
        }
    int per_cu = 0;
    const uint32_t lds = gr::launch_lds_bytes(block, lds_cols, ctx->batch_cu_waves, ctx->lds_per_cu);
    if ((uni ? gr::batch_uni_occupancy(uni, block, lds, &per_cu)
             : gr::batch_occupancy(block, ctx->batch_waves, lds, &per_cu)) != 0 ||
        per_cu <= 0)
        per_cu = 1;
    // enough workgroups to fill the device, no more than one wavefront per item
    const uint32_t wpb = block / 64u;
    const uint32_t grid = std::max<uint32_t>(
        1, std::min<uint32_t>((nitems + wpb - 1) / wpb, uint32_t(per_cu) * uint32_t(std::max(1, ctx->cus))));
    const gr::StepParams * dparams = reinterpret_cast<const gr::StepParams *>(dbase + off_p);
    uint32_t * dq = reinterpret_cast<uint32_t *>(dbase);
    // the launch's own elapsed time only when the caller asks for it (each record is one more
    // packet between consecutive launches)
    if (ms) GR_HIP(hipEventRecord(ctx->bev0, s));
    if (nitems && (rc = kt_mark(ctx, s, false)) != GR_OK) return rc;
    if (nitems && (uni ? gr::launch_batch_uni(uni, dparams, ctx->d_items, dq, nq, ctx->d_culut, group, ctx->dwork,
                                              grid, block, lds, s)
                       : gr::launch_batch(dparams, ctx->d_items, dq, nq, ctx->d_culut, group, ctx->dwork, grid,
                                          block, ctx->batch_waves, narrow, lds, s)) != 0)
        return set_err(GR_E_HIP, "batch kernel launch failed: %s", hipGetErrorString(hipGetLastError()));
    ctx->batch_last_kernel = !nitems ? 0u : uni ? 4u : narrow ? 3u : ctx->batch_waves >= 4 ? 2u : 1u;
    hipEvent_t done = nullptr;
    if (nitems && (rc = kt_mark(ctx, s, true, &done)) != GR_OK) return rc;
    if (ms) GR_HIP(hipEventRecord(ctx->bev1, s));
    std::vector<std::pair<gr_field *, uint32_t>> used;
    for (uint32_t i = 0; i < nspans; ++i)
    {
        if (!span_steps[i]) continue;
        gr_job * j = spans[i].job;
        // every slice the span read is in use until the launch ends (one event record per slot)
        const uint32_t c_end = std::min<uint32_t>(first_win[i].wind_back + span_nwin[i] + 1,
                                                  spans[i].wind->d.nslots);
        for (uint32_t c = first_win[i].wind_back; c < c_end; ++c)
        {
            const std::pair<gr_field *, uint32_t> u(spans[i].wind, c);
            if (std::find(used.begin(), used.end(), u) == used.end()) used.push_back(u);
        }
        if ((rc = finish_advance(j, &first_win[i], saves[i], nullptr)) != GR_OK) return rc;
        j->timed = false;
    }
    for (const auto & u : used) mark_inuse(u.first, u.second, done);
    if (ms)
    {
        GR_HIP(hipEventSynchronize(ctx->bev1));
        GR_HIP(hipEventElapsedTime(ms, ctx->bev0, ctx->bev1));
    }
    return GR_OK;
}

extern "C" int gr_job_read(gr_job * j, uint32_t kind, void * dst, uint64_t bytes)
{
    if (!j || !dst) return set_err(GR_E_INVALID, "gr_job_read: null argument");
    GR_HIP(hipSetDevice(j->ctx->device));
    hipStream_t s = j->ctx->compute;
    const uint64_t n = j->n;
    const void * src = nullptr;
    uint64_t need = n * 4;
    if (kind == GR_OUT_POSITIONS)
    {
        need = n * j->d.dim * 4;
        if (bytes < need) return set_err(GR_E_INVALID, "gr_job_read: buffer too small");
        if (!n) return GR_OK;
        float * soa[3] = {j->p[0], j->p[1], j->d.dim == 3 ? j->p[2] : j->p[1]};
        if (gr::launch_gather_positions(soa, j->scratch_aos, n, j->d.dim, j->d.offset, j->d.local_coords, s) != 0)
            return set_err(GR_E_HIP, "gather kernel failed");
        src = j->scratch_aos;
    }
    else if (kind == GR_OUT_STATUS)
        src = j->status;
    else if (kind == GR_OUT_ABORT_STEP)
        src = j->abort_step;
    else if (kind == GR_OUT_TOTAL_CURVATURE)
        src = j->total;
    else if (kind == GR_OUT_LENGTHS)
        src = j->len;
    else if (kind == GR_OUT_CURVATURE)
        src = j->curv;
    else if (kind >= GR_OUT_EXPR0 && kind < GR_OUT_EXPR0 + GR_MAX_PROGRAMS)
        src = j->expr[kind - GR_OUT_EXPR0];
    else if (kind >= GR_OUT_VOLUME0 && kind < GR_OUT_VOLUME0 + GR_MAX_VOLUMES)
        src = j->vol[kind - GR_OUT_VOLUME0];
    else
        return set_err(GR_E_INVALID, "gr_job_read: unknown kind %u", kind);
    if (!src) return set_err(GR_E_INVALID, "gr_job_read: output %u is not produced by this job", kind);
    if (bytes < need) return set_err(GR_E_INVALID, "gr_job_read: buffer too small");
    if (!n) return GR_OK;
    if (j->reordered)
    {
        // back to the caller's particle order
        const uint32_t w = kind == GR_OUT_POSITIONS ? j->d.dim : 1;
        if (gr::launch_scatter32(src, j->scratch2, j->orig, n, w, s) != 0)
            return set_err(GR_E_HIP, "scatter kernel failed");
        src = j->scratch2;
    }
    GR_HIP(hipMemcpyAsync(dst, src, need, hipMemcpyDeviceToHost, s));
    GR_HIP(hipStreamSynchronize(s));
    return GR_OK;
}

// The ring and column size of saved output `kind`, after checking that columns [first, first +
// count) are written and not yet overwritten.
static int saved_source(gr_job * j, uint32_t kind, uint64_t first, uint64_t count, const float ** base_out,
                        uint64_t * colsz_out)
{
    if (first + count > j->saved) return set_err(GR_E_STATE, "columns [%llu, %llu) not written yet",
                                                 (unsigned long long)first, (unsigned long long)(first + count));
    if (j->saved - first > j->cap) return set_err(GR_E_STATE, "column %llu was overwritten", (unsigned long long)first);
    if (j->reordered && first < j->reorder_col0)
        return set_err(GR_E_STATE, "column %llu was written before the last reorder", (unsigned long long)first);
    const uint64_t n = j->n;
    const float * base;
    uint64_t colsz;
    if (kind == GR_OUT_POSITIONS)
    {
        base = j->ring_pos;
        colsz = n * j->d.dim;
    }
    else if (kind == GR_OUT_CURVATURE)
    {
        base = j->ring_curv;
        colsz = n;
    }
    else if (kind >= GR_OUT_EXPR0 && kind < GR_OUT_EXPR0 + j->d.nprog)
    {
        base = j->ring_expr ? j->ring_expr + uint64_t(kind - GR_OUT_EXPR0) * j->cap * n : nullptr;
        colsz = n;
    }
    else if (kind >= GR_OUT_VOLUME0 && kind < GR_OUT_VOLUME0 + j->d.nvol)
    {
        base = j->ring_vol ? j->ring_vol + uint64_t(kind - GR_OUT_VOLUME0) * j->cap * n : nullptr;
        colsz = n;
    }
    else
        return set_err(GR_E_INVALID, "gr_job_read_saved: kind %u has no saved columns", kind);
    if (!base) return set_err(GR_E_INVALID, "gr_job_read_saved: output %u is not saved by this job", kind);
    *base_out = base;
    *colsz_out = colsz;
    return GR_OK;
}

// The step kernels save only particles alive at the saving step: complete columns [first,
// first + count) with the final state of the particles stopped before them (compute stream).
static int complete_frozen(gr_job * j, uint32_t kind, const float * base, uint64_t first, uint64_t count)
{
    const float * src[3] = {nullptr, nullptr, nullptr};
    uint32_t w = 1, local = 0;
    if (kind == GR_OUT_POSITIONS)
    {
        w = j->d.dim;
        for (uint32_t d = 0; d < w; ++d) src[d] = j->p[d];
        local = j->d.local_coords ? 1u : 0u;
    }
    else if (kind == GR_OUT_CURVATURE)
        src[0] = j->curv;
    else if (kind >= GR_OUT_VOLUME0)
        src[0] = j->vol[kind - GR_OUT_VOLUME0];
    else
        src[0] = j->expr[kind - GR_OUT_EXPR0];
    for (uint64_t c = 0; c < count; c += 65536)
        if (gr::launch_fill_frozen(const_cast<float *>(base), j->cap, first + c,
                                   uint32_t(std::min<uint64_t>(65536, count - c)), j->n, w, j->status,
                                   j->abort_step, src, j->d.save_interval, local, j->d.offset, j->ctx->compute) != 0)
            return set_err(GR_E_HIP, "fill kernel failed");
    return GR_OK;
}

extern "C" int gr_job_read_saved(gr_job * j, uint32_t kind, uint64_t first, uint64_t count, void * dst)
{
    if (!j || (count && !dst)) return set_err(GR_E_INVALID, "gr_job_read_saved: null argument");
    if (count == 0) return GR_OK;
    GR_HIP(hipSetDevice(j->ctx->device));
    const float * base = nullptr;
    uint64_t colsz = 0;
    int rc = saved_source(j, kind, first, count, &base, &colsz);
    if (rc != GR_OK || (rc = complete_frozen(j, kind, base, first, count)) != GR_OK) return rc;
    hipStream_t s = j->ctx->compute;
    float * out = static_cast<float *>(dst);
    if (j->reordered)
    {
        const uint32_t w = kind == GR_OUT_POSITIONS ? j->d.dim : 1;
        for (uint64_t c = 0; c < count; ++c)
        {
            const uint64_t slot = (first + c) % j->cap;
            if (gr::launch_scatter32(base + slot * colsz, j->scratch2, j->orig, j->n, w, s) != 0)
                return set_err(GR_E_HIP, "scatter kernel failed");
            GR_HIP(hipMemcpyAsync(out + c * colsz, j->scratch2, colsz * sizeof(float),
                                  hipMemcpyDeviceToHost, s));
        }
        GR_HIP(hipStreamSynchronize(s));
        return GR_OK;
    }
    uint64_t done = 0;
    while (done < count)
    {
        const uint64_t slot = (first + done) % j->cap;
        const uint64_t run = std::min<uint64_t>(count - done, j->cap - slot);
        GR_HIP(hipMemcpyAsync(out + done * colsz, base + slot * colsz, run * colsz * sizeof(float),
                              hipMemcpyDeviceToHost, s));
        done += run;
    }
    GR_HIP(hipStreamSynchronize(s));
    return GR_OK;
}

// A read record with buffers of at least `host` pinned and `dev` device bytes, from the pool
// when one fits.
static int read_record(gr_job * j, size_t host, size_t dev, gr_job::Read & out)
{
    size_t best = j->read_pool.size();
    for (size_t i = 0; i < j->read_pool.size(); ++i)
    {
        const gr_job::Read & r = j->read_pool[i];
        if (r.host_cap >= host && r.dstage_cap >= dev &&
            (best == j->read_pool.size() || r.host_cap < j->read_pool[best].host_cap))
            best = i;
    }
    if (best == j->read_pool.size() && !j->read_pool.empty()) best = 0; // regrow one
    gr_job::Read r;
    if (best < j->read_pool.size())
    {
        r = std::move(j->read_pool[best]);
        j->read_pool.erase(j->read_pool.begin() + long(best));
    }
    int rc = GR_OK;
    if (r.host_cap < host)
    {
        if (r.host) (void)hipHostFree(r.host);
        r.host = nullptr;
        r.host_cap = 0;
        if (hipHostMalloc(&r.host, host, hipHostMallocDefault) != hipSuccess)
        {
            (void)hipGetLastError();
            r.host = nullptr;
            rc = set_err(GR_E_NOMEM, "pinned read buffer of %zu bytes failed", host);
        }
        else
            r.host_cap = host;
    }
    if (rc == GR_OK && r.dstage_cap < dev)
    {
        dfree(r.dstage);
        r.dstage = nullptr;
        r.dstage_cap = 0;
        if ((rc = dalloc(&r.dstage, dev)) == GR_OK) r.dstage_cap = dev;
    }
    if (rc == GR_OK && !r.ready && hipEventCreateWithFlags(&r.ready, hipEventDisableTiming | hipEventDisableSystemFence) != hipSuccess)
        rc = set_err(GR_E_HIP, "event creation failed");
    if (rc == GR_OK && !r.done && hipEventCreateWithFlags(&r.done, hipEventDisableTiming) != hipSuccess)
        rc = set_err(GR_E_HIP, "event creation failed");
    if (rc != GR_OK)
    {
        j->read_pool.push_back(std::move(r));
        return rc;
    }
    out = std::move(r);
    return GR_OK;
}

extern "C" int gr_job_read_saved_async(gr_job * j, uint32_t kind, uint64_t first, uint64_t count,
                                       void * const * dst_cols, uint64_t * ticket)
{
    if (!j || !ticket || (count && !dst_cols)) return set_err(GR_E_INVALID, "gr_job_read_saved_async: null argument");
    *ticket = 0;
    if (count == 0) return GR_OK;
    for (uint64_t c = 0; c < count; ++c)
        if (!dst_cols[c]) return set_err(GR_E_INVALID, "gr_job_read_saved_async: column %llu has no destination",
                                         (unsigned long long)c);
    gr_ctx * ctx = j->ctx;
    GR_HIP(hipSetDevice(ctx->device));
    const float * base = nullptr;
    uint64_t colsz = 0;
    int rc = saved_source(j, kind, first, count, &base, &colsz);
    if (rc != GR_OK) return rc;
    const size_t bytes = size_t(count) * colsz * sizeof(float);
    gr_job::Read r;
    if ((rc = read_record(j, bytes, j->reordered ? bytes : 0, r)) != GR_OK) return rc;
    auto fail = [&](int code) {
        j->read_pool.push_back(std::move(r));
        return code;
    };
    if ((rc = complete_frozen(j, kind, base, first, count)) != GR_OK) return fail(rc);
    hipStream_t s = ctx->compute;
    r.from_ring = !j->reordered;
    r.first = first;
    r.count = count;
    r.col_bytes = colsz * sizeof(float);
    r.dst.assign(dst_cols, dst_cols + count);
    if (j->reordered)
    {
        // back to the caller's particle order, column by column into the staging copy
        const uint32_t w = kind == GR_OUT_POSITIONS ? j->d.dim : 1;
        float * st = static_cast<float *>(r.dstage);
        for (uint64_t c = 0; c < count; ++c)
            if (gr::launch_scatter32(base + ((first + c) % j->cap) * colsz, st + c * colsz, j->orig, j->n, w, s) != 0)
                return fail(set_err(GR_E_HIP, "scatter kernel failed"));
    }
    if (hipEventRecord(r.ready, s) != hipSuccess || hipStreamWaitEvent(ctx->copy, r.ready, 0) != hipSuccess)
        return fail(set_err(GR_E_HIP, "read event failed"));
    char * host = static_cast<char *>(r.host);
    if (j->reordered)
    {
        if (hipMemcpyAsync(host, r.dstage, bytes, hipMemcpyDeviceToHost, ctx->copy) != hipSuccess)
            return fail(set_err(GR_E_HIP, "device-to-host copy failed"));
    }
    else
    {
        uint64_t done = 0;
        while (done < count)
        {
            const uint64_t slot = (first + done) % j->cap;
            const uint64_t run = std::min<uint64_t>(count - done, j->cap - slot);
            if (hipMemcpyAsync(host + done * r.col_bytes, base + slot * colsz, run * r.col_bytes,
                               hipMemcpyDeviceToHost, ctx->copy) != hipSuccess)
                return fail(set_err(GR_E_HIP, "device-to-host copy failed"));
            done += run;
        }
    }
    if (hipEventRecord(r.done, ctx->copy) != hipSuccess) return fail(set_err(GR_E_HIP, "read event failed"));
    r.ticket = j->next_ticket++;
    *ticket = r.ticket;
    j->reads.push_back(std::move(r));
    return GR_OK;
}

// Copies `count` columns of `col_bytes` from contiguous src to the column pointers, on several
// threads when the amount is large.
static void copy_columns(const char * src, const std::vector<void *> & dst, size_t col_bytes)
{
    const size_t count = dst.size(), total = count * col_bytes;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = total < (64u << 20) ? 1 : std::min<size_t>(std::min<size_t>(16, hw), total / (16u << 20));
    auto part = [&](size_t t) {
        // byte range [b, e) of the concatenated columns
        const size_t b = total * t / nt, e = total * (t + 1) / nt;
        size_t c = b / col_bytes, off = b % col_bytes, at = b;
        while (at < e)
        {
            const size_t len = std::min(col_bytes - off, e - at);
            std::memcpy(static_cast<char *>(dst[c]) + off, src + at, len);
            at += len;
            ++c;
            off = 0;
        }
    };
    if (nt <= 1)
    {
        part(0);
        return;
    }
    std::vector<std::thread> th;
    for (size_t t = 1; t < nt; ++t) th.emplace_back(part, t);
    part(0);
    for (auto & x : th) x.join();
}

extern "C" int gr_job_read_wait(gr_job * j, uint64_t ticket)
{
    if (!j) return set_err(GR_E_INVALID, "gr_job_read_wait: null job");
    if (ticket == 0) return GR_OK;
    size_t i = 0;
    while (i < j->reads.size() && j->reads[i].ticket != ticket) ++i;
    if (i == j->reads.size()) return set_err(GR_E_INVALID, "gr_job_read_wait: unknown ticket %llu", (unsigned long long)ticket);
    GR_HIP(hipSetDevice(j->ctx->device));
    gr_job::Read r = std::move(j->reads[i]);
    j->reads.erase(j->reads.begin() + long(i));
    const hipError_t e = hipEventSynchronize(r.done);
    if (e == hipSuccess) copy_columns(static_cast<const char *>(r.host), r.dst, r.col_bytes);
    r.dst.clear();
    j->read_pool.push_back(std::move(r));
    if (e != hipSuccess) return set_err(GR_E_HIP, "gr_job_read_wait: %s", hipGetErrorString(e));
    return GR_OK;
}

// Before a launch that saves `nsaves` columns from column j->saved on: the compute stream waits
// for the in-flight copies of ring columns that those saves would overwrite (the same ring slot,
// a multiple of the capacity later).
static int wait_ring_reads(gr_job * j, uint64_t nsaves, hipStream_t s)
{
    if (!nsaves) return GR_OK;
    const uint64_t s0 = j->saved, cap = j->cap;
    for (const gr_job::Read & r : j->reads)
    {
        if (!r.from_ring || !r.count) continue;
        // writes g in [s0, s0 + nsaves), reads c in [first, first + count), c < s0 <= g: g hits
        // the slot of c when g - c is a positive multiple of cap
        const uint64_t lo = s0 - (r.first + r.count - 1), hi = s0 + nsaves - 1 - r.first;
        if (hi / cap >= (lo + cap - 1) / cap && hi >= cap) GR_HIP(hipStreamWaitEvent(s, r.done, 0));
    }
    return GR_OK;
}

extern "C" int gr_job_count_alive(gr_job * j, uint64_t * alive)
{
    if (!j || !alive) return set_err(GR_E_INVALID, "gr_job_count_alive: null argument");
    GR_HIP(hipSetDevice(j->ctx->device));
    hipStream_t s = j->ctx->compute;
    GR_HIP(hipMemsetAsync(j->ctx->dcount, 0, sizeof(unsigned long long), s));
    if (gr::launch_count_alive(j->status, j->n, j->ctx->dcount, s) != 0)
        return set_err(GR_E_HIP, "count kernel failed");
    unsigned long long h = 0;
    GR_HIP(hipMemcpyAsync(&h, j->ctx->dcount, sizeof(h), hipMemcpyDeviceToHost, s));
    GR_HIP(hipStreamSynchronize(s));
    *alive = h;
    return GR_OK;
}

extern "C" int gr_job_live_range(gr_job * j, uint64_t * range)
{
    if (!j || !range) return set_err(GR_E_INVALID, "gr_job_live_range: null argument");
    GR_HIP(hipSetDevice(j->ctx->device));
    const int rc = alive_resolve(j, false);
    if (rc != GR_OK) return rc;
    *range = j->run_n;
    return GR_OK;
}

extern "C" int gr_job_saved_columns(gr_job * j, uint64_t * cols)
{
    if (!j || !cols) return set_err(GR_E_INVALID, "null argument");
    *cols = j->saved;
    return GR_OK;
}

extern "C" int gr_job_last_advance_ms(gr_job * j, float * ms)
{
    if (!j || !ms) return set_err(GR_E_INVALID, "null argument");
    if (!j->timed) return set_err(GR_E_STATE, "no advance recorded");
    GR_HIP(hipSetDevice(j->ctx->device));
    GR_HIP(hipEventSynchronize(j->ev1));
    GR_HIP(hipEventElapsedTime(ms, j->ev0, j->ev1));
    return GR_OK;
}

// --------------------------------------------------------------------------------------------
// autotuner of the batched kernel
// --------------------------------------------------------------------------------------------

// Runs every candidate launch configuration of the batched kernel (workgroup size x compute
// units per queue x wavefronts per CU) on the same probe workload -- jobs of the caller's
// description with n seeds each spread uniformly over the domain, Morton-ordered, advanced
// `steps` steps from window 0 -- and keeps the fastest as the context's batch options. Every
// candidate starts from identical state and computes identical results, so device times compare
// directly. The context's work counter and kernel-time pool are left as they were.
extern "C" int gr_batch_autotune(gr_ctx * ctx, gr_field * wind, const gr_job_desc * desc, uint32_t n,
                                 uint32_t njobs, uint32_t steps, uint32_t trials, gr_tune_entry * results,
                                 uint32_t max_results, uint32_t * nresults)
{
    if (!ctx || !wind || !desc) return set_err(GR_E_INVALID, "gr_batch_autotune: null argument");
    if (n == 0 || steps == 0) return set_err(GR_E_INVALID, "gr_batch_autotune: n and steps must be > 0");
    if (njobs == 0) njobs = 1;
    if (njobs > 256) return set_err(GR_E_INVALID, "gr_batch_autotune: at most 256 probe jobs");
    if (trials == 0) trials = 2;
    GR_HIP(hipSetDevice(ctx->device));
    int rc;
    // park the context's measurement state
    if ((rc = kt_fold(ctx)) != GR_OK) return rc;
    const double kms_saved = ctx->kms_acc;
    const uint64_t kl_saved = ctx->klaunch_acc;
    std::vector<unsigned long long> work_saved(size_t(gr::WORK_SLOTS) * gr::WORK_STRIDE);
    GR_HIP(hipMemcpyAsync(work_saved.data(), ctx->dwork, work_saved.size() * 8, hipMemcpyDeviceToHost, ctx->compute));
    GR_HIP(hipStreamSynchronize(ctx->compute));
    const uint32_t block0 = ctx->batch_block, waves0 = ctx->batch_waves, group0 = ctx->batch_group;
    const uint32_t cuw0 = ctx->batch_cu_waves;

    // probe seeds of each job: deterministic low-discrepancy spreads over the domain (inset by
    // 1e-3), a different phase per job
    const uint32_t dim = desc->dim;
    std::vector<std::vector<float>> seeds(njobs, std::vector<float>(size_t(n) * dim));
    const double alpha[3] = {0.7548776662466927, 0.5698402909980532, 0.4142135623730950};
    for (uint32_t jb = 0; jb < njobs; ++jb)
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t d = 0; d < dim; ++d)
            {
                double u = std::fmod(0.5 + alpha[d] * (double(i + 1) + 0.3819660112501051 * jb * double(n)), 1.0);
                u = 1e-3 + u * (1.0 - 2e-3);
                seeds[jb][size_t(i) * dim + d] = float(desc->offset[d] + u * double(desc->domain[d]));
            }

    // The register budget stays the compiler's allocation (the 128-VGPR build spills and was never
    // faster); the occupancy cap is the wavefronts-per-CU dimension: 0 = as many as registers
    // allow (12 at block 64), or 4 / 6 / 8 (fewer concurrent stencil sets per CU cache).
    struct Cand { uint32_t block, waves, group, cu_waves; };
    std::vector<Cand> cands;
    const uint32_t gx = std::max(1u, ctx->ncu / std::max(1u, ctx->nxcd));
    std::vector<uint32_t> groups{gx};
    if (gx > 8) groups.push_back(8);
    for (uint32_t g : groups)
        for (uint32_t cw : {0u, 4u, 6u, 8u})
            for (uint32_t b : {64u, 128u, 256u})
            {
                if (cw && cw < b / 64) continue; // the cap admits one workgroup at least
                cands.push_back(Cand{b, 0u, g, cw});
            }
    float best = 3.4e38f;
    size_t best_i = 0;
    uint32_t nres = 0;
    std::vector<gr_job *> jobs(njobs, nullptr);
    std::vector<gr_span> sp(njobs);
    for (size_t ci = 0; ci < cands.size() && rc == GR_OK; ++ci)
    {
        float tmin = 3.4e38f;
        for (uint32_t t = 0; t < trials && rc == GR_OK; ++t)
        {
            for (uint32_t jb = 0; jb < njobs && rc == GR_OK; ++jb)
                rc = gr_job_create(ctx, desc, seeds[jb].data(), n, steps, &jobs[jb]);
            if (rc == GR_OK && n >= 4096) rc = gr_jobs_reorder(jobs.data(), njobs);
            if (rc == GR_OK)
            {
                ctx->batch_block = cands[ci].block;
                ctx->batch_waves = cands[ci].waves;
                ctx->batch_group = cands[ci].group;
                ctx->batch_cu_waves = cands[ci].cu_waves;
                for (uint32_t jb = 0; jb < njobs; ++jb)
                {
                    std::memset(&sp[jb], 0, sizeof(gr_span));
                    sp[jb].job = jobs[jb];
                    sp[jb].wind = wind;
                    sp[jb].max_steps = steps;
                }
                float ms = 0.f;
                rc = gr_batch_run(ctx, sp.data(), njobs, 0, nullptr, &ms);
                if (rc == GR_OK) tmin = std::min(tmin, ms);
            }
            for (gr_job *& j : jobs)
            {
                if (j) gr_job_destroy(j);
                j = nullptr;
            }
        }
        if (rc != GR_OK) break;
        if (results && nres < max_results)
            results[nres++] = gr_tune_entry{cands[ci].block, cands[ci].waves, cands[ci].group, tmin,
                                            cands[ci].cu_waves};
        if (tmin < best)
        {
            best = tmin;
            best_i = ci;
        }
    }
    // restore the measurement state; keep the winner (or the old options on failure)
    int rc2 = kt_fold(ctx);
    ctx->kms_acc = kms_saved;
    ctx->klaunch_acc = kl_saved;
    if (hipMemcpyAsync(ctx->dwork, work_saved.data(), work_saved.size() * 8, hipMemcpyHostToDevice, ctx->compute) != hipSuccess ||
        hipStreamSynchronize(ctx->compute) != hipSuccess)
        rc2 = set_err(GR_E_HIP, "gr_batch_autotune: restoring the work counter failed");
    if (rc != GR_OK)
    {
        ctx->batch_block = block0;
        ctx->batch_waves = waves0;
        ctx->batch_group = group0;
        ctx->batch_cu_waves = cuw0;
        return rc;
    }
    ctx->batch_block = cands[best_i].block;
    ctx->batch_waves = cands[best_i].waves;
    ctx->batch_group = cands[best_i].group;
    ctx->batch_cu_waves = cands[best_i].cu_waves;
    if (nresults) *nresults = nres;
    return rc2;
}

// --------------------------------------------------------------------------------------------
// autotuner of single-job launches
// --------------------------------------------------------------------------------------------

// Runs every candidate launch configuration of gr_job_advance (workgroup size x lanes per
// particle x LDS field tile) on one probe job with the caller's description, from identical
// state each time, and keeps the fastest as the job's options.
extern "C" int gr_job_autotune(gr_job * job, gr_field * wind, uint32_t steps, uint32_t trials,
                               gr_job_tune_entry * results, uint32_t max_results, uint32_t * nresults)
{
    if (!job || !wind) return set_err(GR_E_INVALID, "gr_job_autotune: null argument");
    if (steps == 0) return set_err(GR_E_INVALID, "gr_job_autotune: steps must be > 0");
    const gr_job_desc & d = job->d;
    if (d.uplift_mode || d.constants_mode || d.nconst || d.nvol || d.volume_mode)
        return set_err(GR_E_INVALID, "gr_job_autotune: jobs with uplift, constants or volumes are not tuned");
    if (d.windfield_mode == 1 && wind->d.nslots < 2)
        return set_err(GR_E_INVALID, "gr_job_autotune: a dynamic wind field needs two slots");
    if (trials == 0) trials = 2;
    gr_ctx * ctx = job->ctx;
    GR_HIP(hipSetDevice(ctx->device));
    int rc;
    // park the context's measurement state (kernel time, work and tile counters)
    if ((rc = kt_fold(ctx)) != GR_OK) return rc;
    const double kms_saved = ctx->kms_acc;
    const uint64_t kl_saved = ctx->klaunch_acc;
    const size_t nctr = size_t(gr::WORK_SLOTS) * gr::WORK_STRIDE;
    std::vector<unsigned long long> work_saved(nctr), tile_saved(nctr);
    GR_HIP(hipMemcpyAsync(work_saved.data(), ctx->dwork, nctr * 8, hipMemcpyDeviceToHost, ctx->compute));
    GR_HIP(hipMemcpyAsync(tile_saved.data(), ctx->dtile, nctr * 8, hipMemcpyDeviceToHost, ctx->compute));
    GR_HIP(hipStreamSynchronize(ctx->compute));

    // probe seeds: a deterministic low-discrepancy spread over the domain (inset by 1e-3)
    const uint32_t dim = d.dim;
    const uint32_t n = uint32_t(std::min<uint64_t>(job->n, 1ull << 20));
    std::vector<float> seeds(size_t(n) * dim);
    const double alpha[3] = {0.7548776662466927, 0.5698402909980532, 0.4142135623730950};
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t k = 0; k < dim; ++k)
        {
            double u = std::fmod(0.5 + alpha[k] * double(i + 1), 1.0);
            u = 1e-3 + u * (1.0 - 2e-3);
            seeds[size_t(i) * dim + k] = float(d.offset[k] + u * double(d.domain[k]));
        }

    struct Cand { uint32_t block, lanes, tile; };
    std::vector<Cand> cands;
    const bool tile_ok = dim == 3 && d.windfield_mode == 1 && job->paired;
    for (uint32_t b : {64u, 128u, 256u})
    {
        cands.push_back(Cand{b, 4u, 0u});
        for (uint32_t t : {0u, 16384u, 32768u})
            if (t == 0 || tile_ok) cands.push_back(Cand{b, 1u, t});
    }
    gr_window w;
    std::memset(&w, 0, sizeof(w));
    w.wind = wind;
    w.wind_back = 0;
    w.wind_front = d.windfield_mode == 1 ? 1 : 0;
    w.time_back = 0.f;
    w.time_front = d.dtd;
    w.time_back64 = 0.0;
    w.time_front64 = double(d.dtd);
    w.counter_start = 0;
    w.nsteps = steps;
    float best = 3.4e38f;
    size_t best_i = 0;
    uint32_t nres = 0;
    for (size_t ci = 0; ci < cands.size() && rc == GR_OK; ++ci)
    {
        float tmin = 3.4e38f;
        for (uint32_t t = 0; t < trials && rc == GR_OK; ++t)
        {
            gr_job * pj = nullptr;
            rc = gr_job_create(ctx, &d, seeds.data(), n, steps, &pj);
            if (rc != GR_OK) break;
            pj->paired = job->paired;
            pj->specialize = job->specialize;
            pj->reuse = job->reuse;
            pj->cu_waves = job->cu_waves;
            pj->tile_halo = job->tile_halo;
            pj->block = cands[ci].block;
            pj->quad_max = cands[ci].lanes == 4 ? uint64_t(n) : 0u;
            pj->tile_bytes = cands[ci].tile;
            if (n >= 4096) rc = gr_job_reorder(pj);
            float ms = 0.f;
            if (rc == GR_OK) rc = gr_job_advance(pj, &w);
            if (rc == GR_OK) rc = gr_job_last_advance_ms(pj, &ms);
            if (rc == GR_OK) tmin = std::min(tmin, ms);
            const int rd = gr_job_destroy(pj);
            if (rc == GR_OK) rc = rd;
        }
        if (rc != GR_OK) break;
        if (results && nres < max_results)
            results[nres++] = gr_job_tune_entry{cands[ci].block, cands[ci].lanes, cands[ci].tile, tmin};
        if (tmin < best)
        {
            best = tmin;
            best_i = ci;
        }
    }
    // restore the measurement state; set the winner (nothing on failure)
    int rc2 = kt_fold(ctx);
    ctx->kms_acc = kms_saved;
    ctx->klaunch_acc = kl_saved;
    if (hipMemcpyAsync(ctx->dwork, work_saved.data(), nctr * 8, hipMemcpyHostToDevice, ctx->compute) != hipSuccess ||
        hipMemcpyAsync(ctx->dtile, tile_saved.data(), nctr * 8, hipMemcpyHostToDevice, ctx->compute) != hipSuccess ||
        hipStreamSynchronize(ctx->compute) != hipSuccess)
        rc2 = set_err(GR_E_HIP, "gr_job_autotune: restoring the counters failed");
    if (rc != GR_OK) return rc;
    job->block = cands[best_i].block;
    job->tile_bytes = cands[best_i].tile;
    if (cands[best_i].lanes == 4) job->quad_max = std::max<uint64_t>(job->quad_max, job->n);
    if (nresults) *nresults = nres;
    return rc2;
}
