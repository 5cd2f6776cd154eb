// Batch traversal kernels behind dl_hip_trace_closest / dl_hip_trace_shadow.
// Rays are 8 planes of n floats (32 B/ray in), hits 4 planes (16 B/ray out),
// occlusion 1 B/ray out.
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
    __device__ __forceinline__ void load(uint32_t i, vec3& o, vec3& d, float& tmin, float& tmax) const {
        o = v3(rays[i], rays[N + i], rays[2 * N + i]);
        d = v3(rays[3 * N + i], rays[4 * N + i], rays[5 * N + i]);
        tmin = rays[6 * N + i];
        tmax = rays[7 * N + i];
    }
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
    __device__ __forceinline__ void store_occ(uint32_t i, bool o) const { occ[i] = o ? 1 : 0; }
};

// Persistent grids (gridDim.x == Ctx::trace_blocks) with per-lane refill.
template <bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void k_trace_closest(BvhView bvh, BatchIO io, uint32_t n, uint32_t* work,
                                                               int* __restrict__ ovf) {
    __shared__ int lds[kLdsStack * kTraceBlock];
    const uint32_t g = blockIdx.x * kTraceBlock + threadIdx.x, G = gridDim.x * kTraceBlock;
    LaneStack stk{lds, (int)threadIdx.x, kTraceBlock, ovf + g, G};
    TraceCounters tot;
    traverse_persistent<false, COUNT>(bvh, n, work, stk, io, tot);
}

__global__ __launch_bounds__(kTraceBlock) void k_trace_shadow(BvhView bvh, BatchIO io, uint32_t n, uint32_t* work,
                                                              int* __restrict__ ovf) {
    __shared__ int lds[kLdsStack * kTraceBlock];
    const uint32_t g = blockIdx.x * kTraceBlock + threadIdx.x, G = gridDim.x * kTraceBlock;
    LaneStack stk{lds, (int)threadIdx.x, kTraceBlock, ovf + g, G};
    TraceCounters tot;
    traverse_persistent<true, false>(bvh, n, work, stk, io, tot);
}

__global__ void k_zero_work(uint32_t* w) {
    if (threadIdx.x < 4) w[threadIdx.x] = 0;
}

dl_status launch_trace_closest(Ctx* c, const float* rays, uint32_t n, uint32_t* prim, float* t, float* u, float* v,
                               uint32_t* cn, uint32_t* ct) {
    if (n == 0) return DL_OK;
    BvhView bvh{c->ds.nodes, c->ds.tris};
    const uint32_t blocks = c->trace_blocks;
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

dl_status launch_trace_shadow(Ctx* c, const float* rays, uint32_t n, uint8_t* occ) {
    if (n == 0) return DL_OK;
    BvhView bvh{c->ds.nodes, c->ds.tris};
    const uint32_t blocks = c->trace_blocks;
    DL_HIP_CHECK(c, hipEventRecord(c->ev0, c->stream));
    BatchIO io{rays, (size_t)n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, occ};
    k_zero_work<<<1, 64, 0, c->stream>>>(c->work);
    k_trace_shadow<<<blocks, kTraceBlock, 0, c->stream>>>(bvh, io, n, c->work, c->stack_ovf);
    DL_HIP_CHECK(c, hipGetLastError());
    DL_HIP_CHECK(c, hipEventRecord(c->ev1, c->stream));
    DL_HIP_CHECK(c, hipEventSynchronize(c->ev1));
    float ms = 0.0f;
    DL_HIP_CHECK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_kernel_ms = ms;
    return DL_OK;
}

}  // namespace dl
