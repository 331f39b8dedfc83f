// Support kernels: slice packing and time pairing, seeds, gathers, alive counts, frozen
// saved columns, and the counting sort that regroups particles by cell.
#include "gr_step.h"

#include <hipcub/device/device_scan.hpp>

namespace gr
{

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

// Launch tables (kernel parameters, queues, time weights, items, sort and gather tables) move from
// the host's pinned staging block to the device with this kernel rather than with a runtime copy:
// it runs in the compute stream's own queue, between the kernels that precede and follow it, so
// the step kernel that reads the tables sees every byte of them.
__global__ void k_stage_copy(const uint4 * __restrict__ src, uint4 * __restrict__ dst, uint64_t n16)
{
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n16;
         i += uint64_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

// Streaming copy for measuring the device's copy bandwidth (gr_ctx_copy_bandwidth): 16 bytes
// per lane per iteration, a grid of many workgroups per CU.
__global__ void __launch_bounds__(256) k_stream_copy(const uint4 * __restrict__ src, uint4 * __restrict__ dst,
                                                     uint64_t n16)
{
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n16;
         i += uint64_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

int launch_stream_copy(const void * src, void * dst, size_t bytes, uint32_t grid, void * stream)
{
    if (bytes == 0 || (bytes & 15)) return -1;
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, hipStream_t(stream),
                       static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), uint64_t(bytes / 16));
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_stage_copy(const void * src, void * dst, size_t bytes, void * stream)
{
    if (bytes == 0) return 0;
    if ((bytes & 15) || (reinterpret_cast<uintptr_t>(src) & 15) || (reinterpret_cast<uintptr_t>(dst) & 15))
        return -1;
    const uint64_t n16 = bytes / 16;
    const uint32_t grid = uint32_t(std::min<uint64_t>((n16 + 255) / 256, 256));
    hipLaunchKernelGGL(k_stage_copy, dim3(grid), dim3(256), 0, hipStream_t(stream),
                       static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), n16);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

__global__ void k_fill_u32(uint32_t * __restrict__ p, uint64_t n, uint32_t v)
{
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_iota_u32(uint32_t * __restrict__ p, uint64_t n)
{
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = uint32_t(i);
}

int launch_iota(uint32_t * p, uint64_t n, void * stream)
{
    if (n == 0) return 0;
    const uint32_t block = 256;
    hipLaunchKernelGGL(k_iota_u32, dim3(uint32_t((n + block - 1) / block)), dim3(block), 0, hipStream_t(stream), p, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
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
