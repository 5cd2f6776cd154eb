// Batch traversal kernels behind dl_hip_trace_closest / dl_hip_trace_shadow.
// Rays are 8 planes of n floats (32 B/ray in), hits 4 planes (16 B/ray out),
// occlusion 1 B/ray out.
#include <algorithm>
#include <mutex>

#include "backend.h"
#include "trace.h"

namespace dl {

// Rays of a batch: 8 planes of n floats; results in planes of n.
struct BatchIO {
    const float* __restrict__ rays;
    size_t N;
    uint32_t* __restrict__ prim;
    float *__restrict__ th, *__restrict__ uh, *__restrict__ vh;
    uint32_t *__restrict__ cn, *__restrict__ ct;
    uint8_t* __restrict__ occ;
    // rays load in a few coalesced reads (measured on 4M diffuse-bounce and
    // 4M shadow rays of the 2M-triangle displaced-sphere scene: closest hit
    // 9% faster refilling 8 lanes at a time than at every finished ray,
    // shadow rays 13% faster at every finished ray)
    static constexpr uint32_t kRefillMin = 8, kRefillMinAny = 1;
    __device__ __forceinline__ uint32_t entry(uint32_t slot) const { return slot; }
    __device__ __forceinline__ void load(uint32_t i, vec3& o, vec3& d, float& tmin, float& tmax) const {
        o = v3(rays[i], rays[N + i], rays[2 * N + i]);
        d = v3(rays[3 * N + i], rays[4 * N + i], rays[5 * N + i]);
        tmin = rays[6 * N + i];
        tmax = rays[7 * N + i];
    }
    __device__ __forceinline__ int root(uint32_t) const { return 0; }
    __device__ __forceinline__ void store_hit(uint32_t i, uint32_t p, float t, float u, float v, uint32_t nn,
                                              uint32_t nt) const {
        prim[i] = p;
        th[i] = t;
        uh[i] = u;
        vh[i] = v;
        if (cn) {
            cn[i] = nn;
            ct[i] = nt;
        }
    }
    __device__ __forceinline__ void store_occ(uint32_t i, bool o, uint32_t nn, uint32_t nt) const {
        occ[i] = o ? 1 : 0;
        if (cn) {
            cn[i] = nn;
            ct[i] = nt;
        }
    }
};

// Persistent grids (Ctx::trace_grid / trace_grid_count blocks) with per-lane refill.
template <bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void k_trace_closest(BvhView bvh, BatchIO io, uint32_t n, uint32_t* work,
                                                               int* __restrict__ ovf) {
    __shared__ int lds[kLdsStack * kTraceBlock];
    const uint32_t g = blockIdx.x * kTraceBlock + threadIdx.x, G = gridDim.x * kTraceBlock;
    LaneStack stk{lds, (int)threadIdx.x, kTraceBlock, ovf + g, G};
    TraceCounters tot;
    traverse_persistent<false, COUNT>(bvh, n, work, stk, io, tot);
}

template <bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void k_trace_shadow(BvhView bvh, BatchIO io, uint32_t n, uint32_t* work,
                                                              int* __restrict__ ovf) {
    __shared__ int lds[kLdsStack * kTraceBlock];
    const uint32_t g = blockIdx.x * kTraceBlock + threadIdx.x, G = gridDim.x * kTraceBlock;
    LaneStack stk{lds, (int)threadIdx.x, kTraceBlock, ovf + g, G};
    TraceCounters tot;
    traverse_persistent<true, COUNT>(bvh, n, work, stk, io, tot);
}

// Resident blocks per CU of the batch traversal kernels (the fewer of the
// closest-hit and any-hit kernel), without or with traversal counters.
int batch_trace_blocks_per_cu(bool counting) {
    int a = 0, b = 0;
    hipError_t e = counting ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_trace_closest<true>, kTraceBlock, 0)
                            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_trace_closest<false>, kTraceBlock, 0);
    if (e == hipSuccess)
        e = counting ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_trace_shadow<true>, kTraceBlock, 0)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_trace_shadow<false>, kTraceBlock, 0);
    if (e != hipSuccess) return kTraceBlocksPerCu;
    return std::max(1, std::min(kTraceBlocksPerCu, std::min(a, b)));
}

__global__ void k_zero_work(uint32_t* w) {
    if (threadIdx.x < 4) w[threadIdx.x] = 0;
}

// Lanes of a persistent grid step through ray indices by the grid size; a
// 32-bit index must not wrap past n.
constexpr uint32_t kMaxBatchRays = 1u << 31;

dl_status launch_trace_closest(Ctx* c, const float* rays, uint32_t n, uint32_t* prim, float* t, float* u, float* v,
                               uint32_t* cn, uint32_t* ct) {
    if (n == 0) return DL_OK;
    if (n > kMaxBatchRays) return set_error(c, DL_ERR_INVALID, "more than 2^31 rays in one batch");
    BvhView bvh{c->ds.nodes, c->ds.tris};
    const uint32_t blocks = cn ? c->trace_grid_count : c->trace_grid;
    DL_HIP_CHECK(c, hipEventRecord(c->ev0, c->stream));
    BatchIO io{rays, (size_t)n, prim, t, u, v, cn, ct, nullptr};
    k_zero_work<<<1, 64, 0, c->stream>>>(c->work);
    if (cn)
        k_trace_closest<true><<<blocks, kTraceBlock, 0, c->stream>>>(bvh, io, n, c->work, c->stack_ovf);
    else
        k_trace_closest<false><<<blocks, kTraceBlock, 0, c->stream>>>(bvh, io, n, c->work, c->stack_ovf);
    DL_HIP_CHECK(c, hipGetLastError());
    DL_HIP_CHECK(c, hipEventRecord(c->ev1, c->stream));
    DL_HIP_CHECK(c, hipEventSynchronize(c->ev1));
    float ms = 0.0f;
    DL_HIP_CHECK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_kernel_ms = ms;
    return DL_OK;
}

dl_status launch_trace_shadow(Ctx* c, const float* rays, uint32_t n, uint8_t* occ, uint32_t* cn, uint32_t* ct) {
    if (n == 0) return DL_OK;
    if (n > kMaxBatchRays) return set_error(c, DL_ERR_INVALID, "more than 2^31 rays in one batch");
    BvhView bvh{c->ds.nodes, c->ds.tris};
    const uint32_t blocks = cn ? c->trace_grid_count : c->trace_grid;
    DL_HIP_CHECK(c, hipEventRecord(c->ev0, c->stream));
    BatchIO io{rays, (size_t)n, nullptr, nullptr, nullptr, nullptr, cn, ct, occ};
    k_zero_work<<<1, 64, 0, c->stream>>>(c->work);
    if (cn)
        k_trace_shadow<true><<<blocks, kTraceBlock, 0, c->stream>>>(bvh, io, n, c->work, c->stack_ovf);
    else
        k_trace_shadow<false><<<blocks, kTraceBlock, 0, c->stream>>>(bvh, io, n, c->work, c->stack_ovf);
    DL_HIP_CHECK(c, hipGetLastError());
    DL_HIP_CHECK(c, hipEventRecord(c->ev1, c->stream));
    DL_HIP_CHECK(c, hipEventSynchronize(c->ev1));
    float ms = 0.0f;
    DL_HIP_CHECK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_kernel_ms = ms;
    return DL_OK;
}

// ---- allocation-free closest hits (dl_hip_trace_closest_noalloc) ---------
// The closest-hit traversal's scratch -- the per-lane stack overflow of a
// persistent grid of kNoAllocBlocks blocks, then its work counter -- in a
// static array that each device holds until the process ends (92.3 MB)
// instead of the context's buffers.  Host code never passes the array's
// name to a kernel: it stands for a host-side shadow there.  Its device
// address is looked up once per device at context creation
// (noalloc_scratch_init) and kept in g_scratch.
constexpr uint32_t kNoAllocBlocks = 1024;
constexpr size_t kNoAllocOvfInts = (size_t)kNoAllocBlocks * kTraceBlock * (kMaxStack - kLdsStack);
constexpr size_t kNoAllocInts = kNoAllocOvfInts + 64;
static __device__ __align__(256) int scratch[kNoAllocInts];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int* g_scratch[kMaxDevices] = {};
}  // namespace

// With c->device current: the device address of this device's scratch.
dl_status noalloc_scratch_init(Ctx* c) {
    if (c->device < 0 || c->device >= kMaxDevices) return set_error(c, DL_ERR_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[c->device]) return DL_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(c, e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    g_scratch[c->device] = (int*)p;
    return DL_OK;
}

dl_status launch_trace_closest_noalloc(Ctx* c, const float* rays, uint32_t n, uint32_t* prim, float* t, float* u,
                                       float* v) {
    if (n == 0) return DL_OK;
    if (n > kMaxBatchRays) return set_error(c, DL_ERR_INVALID, "more than 2^31 rays in one batch");
    int* s = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (c->device >= 0 && c->device < kMaxDevices) s = g_scratch[c->device];
    }
    if (!s) return set_error(c, DL_ERR_STATE, "no scratch address for this device");
    BvhView bvh{c->ds.nodes, c->ds.tris};
    BatchIO io{rays, (size_t)n, prim, t, u, v, nullptr, nullptr, nullptr};
    uint32_t* work = (uint32_t*)(s + kNoAllocOvfInts);
    DL_HIP_CHECK(c, hipEventRecord(c->ev0, c->stream));
    k_zero_work<<<1, 64, 0, c->stream>>>(work);
    k_trace_closest<false><<<kNoAllocBlocks, kTraceBlock, 0, c->stream>>>(bvh, io, n, work, s);
    DL_HIP_CHECK(c, hipGetLastError());
    DL_HIP_CHECK(c, hipEventRecord(c->ev1, c->stream));
    DL_HIP_CHECK(c, hipEventSynchronize(c->ev1));
    float ms = 0.0f;
    DL_HIP_CHECK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_kernel_ms = ms;
    return DL_OK;
}

}  // namespace dl
