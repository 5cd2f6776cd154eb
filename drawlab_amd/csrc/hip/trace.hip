// Batch traversal kernels behind dl_hip_trace_closest / dl_hip_trace_shadow.
// Rays are 8 planes of n floats (32 B/ray in), hits 4 planes (16 B/ray out),
// occlusion 1 B/ray out.
#include "backend.h"
#include "trace.h"

namespace dl {

// Persistent grids (gridDim.x == Ctx::trace_blocks): each lane loops over
// rays with a grid stride, reusing its stack.
template <bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void k_trace_closest(BvhView bvh, const float* __restrict__ rays, uint32_t n,
                                                               uint32_t* __restrict__ prim, float* __restrict__ th,
                                                               float* __restrict__ uh, float* __restrict__ vh,
                                                               uint32_t* __restrict__ cn, uint32_t* __restrict__ ct,
                                                               int* __restrict__ ovf) {
    __shared__ int lds[kLdsStack * kTraceBlock];
    const uint32_t g = blockIdx.x * kTraceBlock + threadIdx.x, G = gridDim.x * kTraceBlock;
    LaneStack stk{lds, (int)threadIdx.x, kTraceBlock, ovf + g, G};
    const size_t N = n;
    for (uint32_t i = g; i < n; i += G) {
        const vec3 o = v3(rays[i], rays[N + i], rays[2 * N + i]);
        const vec3 d = v3(rays[3 * N + i], rays[4 * N + i], rays[5 * N + i]);
        const float tmin = rays[6 * N + i], tmax = rays[7 * N + i];
        uint32_t p;
        float t, u, v;
        TraceCounters cnt;
        trace_closest<COUNT>(bvh, o, d, tmin, tmax, stk, p, t, u, v, cnt);
        prim[i] = p;
        th[i] = t;
        uh[i] = u;
        vh[i] = v;
        if (COUNT) {
            cn[i] = cnt.nodes;
            ct[i] = cnt.tris;
        }
    }
}

__global__ __launch_bounds__(kTraceBlock) void k_trace_shadow(BvhView bvh, const float* __restrict__ rays, uint32_t n,
                                                              uint8_t* __restrict__ occ, int* __restrict__ ovf) {
    __shared__ int lds[kLdsStack * kTraceBlock];
    const uint32_t g = blockIdx.x * kTraceBlock + threadIdx.x, G = gridDim.x * kTraceBlock;
    LaneStack stk{lds, (int)threadIdx.x, kTraceBlock, ovf + g, G};
    const size_t N = n;
    for (uint32_t i = g; i < n; i += G) {
        const vec3 o = v3(rays[i], rays[N + i], rays[2 * N + i]);
        const vec3 d = v3(rays[3 * N + i], rays[4 * N + i], rays[5 * N + i]);
        const float tmin = rays[6 * N + i], tmax = rays[7 * N + i];
        TraceCounters cnt;
        occ[i] = trace_any<false>(bvh, o, d, tmin, tmax, stk, cnt) ? 1 : 0;
    }
}

dl_status launch_trace_closest(Ctx* c, const float* rays, uint32_t n, uint32_t* prim, float* t, float* u, float* v,
                               uint32_t* cn, uint32_t* ct) {
    if (n == 0) return DL_OK;
    BvhView bvh{c->ds.nodes, c->ds.tris};
    const uint32_t blocks = c->trace_blocks;
    DL_HIP_CHECK(c, hipEventRecord(c->ev0, c->stream));
    if (cn)
        k_trace_closest<true><<<blocks, kTraceBlock, 0, c->stream>>>(bvh, rays, n, prim, t, u, v, cn, ct, c->stack_ovf);
    else
        k_trace_closest<false><<<blocks, kTraceBlock, 0, c->stream>>>(bvh, rays, n, prim, t, u, v, nullptr, nullptr,
                                                                      c->stack_ovf);
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
    k_trace_shadow<<<blocks, kTraceBlock, 0, c->stream>>>(bvh, rays, n, occ, c->stack_ovf);
    DL_HIP_CHECK(c, hipGetLastError());
    DL_HIP_CHECK(c, hipEventRecord(c->ev1, c->stream));
    DL_HIP_CHECK(c, hipEventSynchronize(c->ev1));
    float ms = 0.0f;
    DL_HIP_CHECK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_kernel_ms = ms;
    return DL_OK;
}

}  // namespace dl
