// Wavefront unidirectional MIS path integrator for gfx950.
//
// Restates PathIntegrator::Li (src/tracer/integrator/path.cpp:13-100) as a
// sequence of kernels over compacted queues of path indices, one bounce per
// iteration:
//
//   raygen          camera rays for every (pixel, sample) of the pass
//                   (CPURenderer::renderBlock, cpu_renderer.cpp:36-48;
//                   VirtualCamera::sampleRay, virtual.cpp:52-65)
//   trace_closest   closest hits of the queued rays
//   post_camera     first-hit emission / environment on a camera-ray miss
//   loop:
//     shade         per path, material-sorted: emitter sampling (NEE) with
//                   a shadow ray, BSDF sampling, continuation ray
//     trace_closest + trace_shadow
//     post          NEE resolution, emitter hit with MIS, miss with the
//                   environment, Russian roulette, depth cap; survivors are
//                   appended to the next material queues
//   film            deterministic gather of the pass's samples through the
//                   reconstruction filter (ImageBlock::put, block.cpp:81-113)
//
// Queues are appended with block-aggregated atomics (one atomic per block and
// queue).  Stage kernels are persistent grid-stride loops whose trip counts
// come from device-side counters, so no launch depends on a host readback
// of its size.
#include <algorithm>
#include <cstring>

#include "../render_sizes.h"
#include "backend.h"
#include "guide.h"
#include "shading.h"
#include "trace.h"

namespace dl {

constexpr int kStageBlock = 256;
constexpr int kTypes = 8;  // material queue segments (DL_BSDF_COUNT rounded up)

// kFlagCamera: the continuation is still the camera ray (it passed through
// a hidden light), so its next hit or miss is treated as the first one.
enum : uint32_t { kFlagDiffuse = 1u, kFlagNee = 2u, kFlagCont = 4u, kFlagCamera = 8u };
// Material-queue segment of paths that pass straight through a hidden
// light's surface (after the BSDF types).
constexpr int kPassQ = DL_BSDF_COUNT;
static_assert(kPassQ < kTypes, "pass-through queue needs a segment");

struct Counters {
    uint32_t mat[2][kTypes];
    uint32_t ray;     // continuation rays queued (block_append index 0)
    uint32_t shadow;  // shadow rays queued (block_append index 1)
    uint32_t work_ray;     // rays claimed by the closest-hit traversal
    uint32_t work_shadow;  // rays claimed by the shadow traversal
    uint32_t emit;         // continuations that hit an emitter (k_post -> k_post_emit)
    uint32_t pad[11];
    unsigned long long nodes;
    unsigned long long tris;
    unsigned long long invalid;
    unsigned long long nodes_sh;  // shadow rays: node visits, triangle tests (counting renders)
    unsigned long long tris_sh;
};

// Device buffers of a render (plain pointers: passed by value to kernels).
struct RenderBufs {
    float4 *ray_o, *ray_d, *hit, *sh_d;  // a shadow ray starts at ray_o
    float4 *thr, *rad, *nee;
    int4* aux;
    uint64_t* rng;
    uint32_t* seq;  // sample index of the path (its pcg32 stream selector)
    float2* pos;
    uint8_t* occ;
    uint32_t* matq[2];
    uint32_t* rayq;
    uint32_t* shq;
    uint32_t* emq;  // paths whose continuation hit an emitter (MIS-weighted emission pending)
    Counters* cnt;
    float4* film;
    float* rgb;
    float* gmix;  // guided rendering: mixture (40 floats) per entry of the queue being shaded
    float* grec;  // guided rendering: training records of the pass
    uint8_t* sid;  // scene of the path (a pass over several scenes)
    const int32_t* roots;  // BVH root node per scene (nullptr: one scene, root node 0)
};

struct RenderState {
    uint32_t capacity = 0;
    uint32_t film_pixels = 0;
    RenderBufs b{};
    bool film_valid = false;       // film holds the accumulation of earlier calls
    bool types_present[8] = {false, false, false, false, false, false, false, false};
    uint32_t gmix_cap = 0, grec_cap = 0;  // guided rendering buffers (entries)
    Counters* cnt_host = nullptr;  // pinned
    std::vector<void*> allocs;
    std::vector<hipEvent_t> evs;
};

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
constexpr int kStageWaves = kStageBlock / 64;
constexpr int kStageItems = 8;                          // items per thread per block-wide append

// Items per thread per loop trip: up to kStageItems, but no more than spread
// the queue over the whole grid (short queues keep every block busy).
__device__ __forceinline__ uint32_t stage_items(uint32_t n) {
    const uint32_t per_trip = gridDim.x * kStageBlock;
    return min((uint32_t)kStageItems, max(1u, (n + per_trip - 1) / per_trip));
}

// Block-aggregated queue append.  Every thread of the block calls it (the
// call sits in block-uniform control flow) with N items; bit q of bits[u]
// asks for a slot in queue q for item u, whose fill counter is counters[q].
// One atomic per block and non-empty queue: the counters share one cache
// line that every block of the grid updates, so finer-grained atomics
// serialise there.  Slots of a queue are dense; emit(q, u, slot) runs once
// per set bit.
template <int NQ, int N, typename Emit>
__device__ __forceinline__ void block_append(uint32_t* counters, const uint32_t (&bits)[N], Emit emit) {
    __shared__ uint32_t s_cnt[kStageWaves][NQ];
    __shared__ uint32_t s_base[kStageWaves][NQ];
    const int lane = __lane_id();
    const int wave = threadIdx.x >> 6;
    const unsigned long long lt = (1ull << lane) - 1ull;
    // the ballots are taken twice (for the counts, then for the slots)
    // rather than kept: NQ * N masks and prefixes would not fit the scalar
    // registers
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        uint32_t tot = 0;
#pragma unroll
        for (int u = 0; u < N; u++) tot += (uint32_t)__popcll(__ballot((bits[u] >> q) & 1u));
        if (lane == 0) s_cnt[wave][q] = tot;
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        const int q = threadIdx.x;
        uint32_t tot = 0;
        for (int w = 0; w < kStageWaves; w++) {
            const uint32_t c = s_cnt[w][q];
            s_cnt[w][q] = tot;
            tot += c;
        }
        const uint32_t base = tot ? atomicAdd(&counters[q], tot) : 0u;
        for (int w = 0; w < kStageWaves; w++) s_base[w][q] = base + s_cnt[w][q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        uint32_t pre = s_base[wave][q];
#pragma unroll
        for (int u = 0; u < N; u++) {
            const bool want = (bits[u] >> q) & 1u;
            const unsigned long long m = __ballot(want);
            if (want) emit(q, u, pre + (uint32_t)__popcll(m & lt));
            pre += (uint32_t)__popcll(m);
        }
    }
    __syncthreads();
}

__device__ __forceinline__ vec3 f3(const float4& a) { return v3(a.x, a.y, a.z); }
__device__ __forceinline__ float4 f4(vec3 a, float w) { return make_float4(a.x, a.y, a.z, w); }

// Locate entry i of the concatenated per-type queues.
__device__ __forceinline__ bool queue_entry(const uint32_t* counts, const uint32_t* q, uint32_t cap, uint32_t i,
                                            uint32_t& path, int& type) {
    uint32_t base = 0;
#pragma unroll
    for (int t = 0; t < kTypes; t++) {
        const uint32_t c = counts[t];
        if (i < base + c) {
            path = q[(size_t)t * cap + (i - base)];
            type = t;
            return true;
        }
        base += c;
    }
    return false;
}

__device__ __forceinline__ uint32_t queue_total(const uint32_t* counts) {
    uint32_t s = 0;
#pragma unroll
    for (int t = 0; t < kTypes; t++) s += counts[t];
    return s;
}

// Learned proposal during a pass: mixtures of the queue being shaded,
// training records (nullptr: not training) and the scene bounds.
struct GuideView {
    const float* mix;
    float frac;           // probability of sampling the proposal (0: BSDF only)
    float* rec;
    uint32_t rec_stride;  // paths p with p % rec_stride == 0 are recorded
    uint32_t rec_cap;     // records
    float lo[3], inv[3];
};

struct PassParams {
    uint32_t P;           // paths in this pass
    uint32_t W, H;
    uint32_t sample_base;
    uint32_t seed;
    int max_depth;
    int rr_depth;
    uint32_t cap;         // queue segment capacity
    float eps;            // ray offset (tmin; shadow rays end this short of the light)
    uint32_t scene;       // scene record of the paths k_raygen starts (a pass over several scenes)
};

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kStageBlock) void k_reset(Counters* cnt, int next) {
    const int t = threadIdx.x;
    if (t < kTypes) cnt->mat[next][t] = 0;
    if (t == 0) {
        cnt->ray = 0;
        cnt->shadow = 0;
        cnt->work_ray = 0;
        cnt->work_shadow = 0;
        cnt->emit = 0;
    }
}

__global__ __launch_bounds__(kStageBlock) void k_reset_all(Counters* cnt) {
    uint32_t* w = reinterpret_cast<uint32_t*>(cnt);
    for (uint32_t i = threadIdx.x; i < sizeof(Counters) / 4; i += blockDim.x) w[i] = 0;
}

__global__ __launch_bounds__(kStageBlock) void k_clear_film(float4* film, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        film[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Camera sample: pixel 2D, aperture 2D, VirtualCamera::sampleRay.
// out = M in for a row-major 4x4 M, each row accumulated from zero in index
// order (the reference's row(j).dot(v), include/core/math/matrix.h:255-262).
__device__ __forceinline__ void mat4_apply(const float* m, const float in[4], float out[4]) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; k++) s += m[4 * j + k] * in[k];
        out[j] = s;
    }
}

// Camera paths pp.P of one image (pp.W x pp.H, samples pp.sample_base ...)
// stored from path index `base` on; a pass of several images (one launch
// each) queues `total` paths.
__global__ __launch_bounds__(kStageBlock) void k_raygen(PassParams pp, dl_camera_desc cam, RenderBufs rs, uint32_t base,
                                                        uint32_t total) {
    const uint32_t WH = pp.W * pp.H;
    const float invx = 1.f / (float)cam.width, invy = 1.f / (float)cam.height;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < pp.P; i += gridDim.x * blockDim.x) {
        const uint32_t p = base + i;
        const uint32_t pix = i % WH, j = i / WH;
        const uint32_t x = pix % pp.W, y = pix / pp.W;
        Pcg32 rng;
        rng.seed((uint64_t)pix + ((uint64_t)pp.seed << 32), (uint64_t)(pp.sample_base + j));
        const float ax = rng.next_float();
        const float ay = rng.next_float();
        const float px = (float)x + ax, py = (float)y + ay;
        rng.next_float();  // aperture sample (unused by every camera)
        rng.next_float();
        if (cam.type == DL_CAMERA_PERSPECTIVE) {
            // PerspectiveCamera::sampleRay  perspective.cpp:75-95
            const float in[4] = {px * cam.inv_size[0], py * cam.inv_size[1], 0.0f, 1.0f};
            float q[4];
            mat4_apply(cam.sample_to_camera, in, q);
            const vec3 d = normalized(v3(q[0] / q[3], q[1] / q[3], q[2] / q[3]));
            const float inv_z = 1.0f / d.z;
            const float o_in[4] = {0.0f, 0.0f, 0.0f, 1.0f}, d_in[4] = {d.x, d.y, d.z, 0.0f};
            float o4[4], d4[4];
            mat4_apply(cam.camera_to_world, o_in, o4);
            mat4_apply(cam.camera_to_world, d_in, d4);
            rs.ray_o[p] = make_float4(o4[0] / o4[3], o4[1] / o4[3], o4[2] / o4[3], cam.near_clip * inv_z);
            rs.ray_d[p] = make_float4(d4[0], d4[1], d4[2], cam.far_clip * inv_z);
        } else if (cam.type == DL_CAMERA_OPENCV) {
            // OpencvCamera (include/optix/camera/opencv.h:28-46): the film
            // position relative to the principal point over the focal
            // lengths, at z = 1 (w = 1) in camera space, mapped to the world
            // by the inverse extrinsic matrix; the ray leaves the eye toward
            // that point
            const float in[4] = {(px - cam.intrinsics[0]) / cam.intrinsics[2],
                                 (py - cam.intrinsics[1]) / cam.intrinsics[3], 1.0f, 1.0f};
            float q[4];
            mat4_apply(cam.camera_to_world, in, q);
            const vec3 d = normalized(sub(v3(q[0], q[1], q[2]), pcol(cam.eye)));
            rs.ray_o[p] = make_float4(cam.eye[0], cam.eye[1], cam.eye[2], pp.eps);
            rs.ray_d[p] = f4(d, INFINITY);
        } else {
            // VirtualCamera::sampleRay  virtual.cpp:52-65
            const float lx = px, ly = ((float)cam.height - py) - 1;
            const float dx = lx * invx * 2 - 1, dy = ly * invy * 2 - 1;
            const vec3 d = normalized(add(add(scale(pcol(cam.U), dx), scale(pcol(cam.V), dy)), pcol(cam.W)));
            rs.ray_o[p] = make_float4(cam.eye[0], cam.eye[1], cam.eye[2], pp.eps);
            rs.ray_d[p] = f4(d, INFINITY);
        }
        // (the radiance starts at the first hit: k_post_camera / k_normals
        // write it for every camera ray; throughput 1, eta 1 and aux 0 are
        // implied at the first shading vertex, and written by k_post_camera
        // for the paths that pass through a hidden light first)
        rs.rng[p] = rng.state;
        rs.seq[p] = pp.sample_base + j;
        rs.pos[p] = make_float2(px, py);
        rs.sid[p] = (uint8_t)pp.scene;
        rs.rayq[p] = p;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) rs.cnt->ray = total;
}

// Queued path rays: work slot i holds path q[i], whose ray lives in
// ray_o/ray_d (shadow: ray_o/sh_d -- a shading vertex's shadow ray and
// continuation start at the same point); the traversal's entries are path
// indices.
struct QueueIO {
    const uint32_t* __restrict__ q;
    const float4* __restrict__ ro;
    const float4* __restrict__ rd;
    float4* __restrict__ hit;
    uint8_t* __restrict__ occ;
    const uint8_t* __restrict__ sid;
    const int32_t* __restrict__ roots;
    // a refill loads the queue entry and then the ray it names (8 measured
    // best of 1, 8 and 16 on the 2M-triangle displaced-sphere render; for
    // shadow rays 1 is 2% faster there but 15% slower on the MIS plates)
    static constexpr uint32_t kRefillMin = 8, kRefillMinAny = 8;
    __device__ __forceinline__ uint32_t entry(uint32_t i) const { return q[i]; }
    // first node of the ray's scene
    __device__ __forceinline__ int root(uint32_t p) const { return roots ? roots[sid[p]] : 0; }
    __device__ __forceinline__ void load(uint32_t p, vec3& o, vec3& d, float& tmin, float& tmax) const {
        const float4 a = ro[p], b = rd[p];
        o = v3(a.x, a.y, a.z);
        d = v3(b.x, b.y, b.z);
        tmin = a.w;
        tmax = b.w;
    }
    __device__ __forceinline__ void store_hit(uint32_t p, uint32_t prim, float t, float u, float v, uint32_t,
                                              uint32_t) const {
        hit[p] = make_float4(__uint_as_float(prim), t, u, v);
    }
    __device__ __forceinline__ void store_occ(uint32_t p, bool o, uint32_t, uint32_t) const { occ[p] = o ? 1 : 0; }
};

template <bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void k_trace_queue(BvhView bvh, RenderBufs rs, int* __restrict__ ovf) {
    __shared__ int lds[kLdsStack * kTraceBlock];
    const uint32_t n = rs.cnt->ray;
    const uint32_t g = blockIdx.x * kTraceBlock + threadIdx.x, G = gridDim.x * kTraceBlock;
    LaneStack stk{lds, (int)threadIdx.x, kTraceBlock, ovf + g, G};
    QueueIO io{rs.rayq, rs.ray_o, rs.ray_d, rs.hit, nullptr, rs.sid, rs.roots};
    TraceCounters cnt;
    traverse_persistent<false, COUNT>(bvh, n, &rs.cnt->work_ray, stk, io, cnt);
    if (COUNT) {
        wave_add(&rs.cnt->nodes, cnt.nodes);
        wave_add(&rs.cnt->tris, cnt.tris);
    }
}

template <bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void k_shadow_queue(BvhView bvh, RenderBufs rs, int* __restrict__ ovf) {
    __shared__ int lds[kLdsStack * kTraceBlock];
    const uint32_t n = rs.cnt->shadow;
    const uint32_t g = blockIdx.x * kTraceBlock + threadIdx.x, G = gridDim.x * kTraceBlock;
    LaneStack stk{lds, (int)threadIdx.x, kTraceBlock, ovf + g, G};
    QueueIO io{rs.shq, rs.ray_o, rs.sh_d, nullptr, rs.occ, rs.sid, rs.roots};
    TraceCounters cnt;
    traverse_persistent<true, COUNT>(bvh, n, &rs.cnt->work_shadow, stk, io, cnt);
    if (COUNT) {
        wave_add(&rs.cnt->nodes_sh, cnt.nodes);
        wave_add(&rs.cnt->tris_sh, cnt.tris);
    }
}

// Resident blocks per CU of the render's traversal kernels (the fewer of
// the closest-hit and shadow kernel), without or with traversal counters:
// the persistent grids are sized to it, so no block waits for a second round.
int render_trace_blocks_per_cu(bool counting) {
    int a = 0, b = 0;
    hipError_t e = counting ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_trace_queue<true>, kTraceBlock, 0)
                            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_trace_queue<false>, kTraceBlock, 0);
    if (e == hipSuccess)
        e = counting ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_shadow_queue<true>, kTraceBlock, 0)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_shadow_queue<false>, kTraceBlock, 0);
    if (e != hipSuccess) return kTraceBlocksPerCu;
    return std::max(1, std::min(kTraceBlocksPerCu, std::min(a, b)));
}

// Material type and area-light flag of a primitive (one byte).
__device__ __forceinline__ uint32_t prim_info(const SceneView& s, uint32_t prim) { return s.prim_info[prim]; }
__device__ __forceinline__ int info_type(uint32_t info) { return (int)(info & (kPrimLight - 1u)); }
__device__ __forceinline__ bool info_light(uint32_t info) { return (info & kPrimLight) != 0; }
// the area light of a primitive whose info has kPrimLight
__device__ __forceinline__ int light_of(const SceneView& s, uint32_t prim) { return s.meshes[s.prim_mesh[prim]].light; }

// First intersection (path.cpp:21-29) and a camera-ray miss (miss.cu:23-29).
// HIDE: light surfaces are hidden (hitgroup_pgs.cu:55-64): a camera ray that
// hits one adds no emission and passes through it (queue kPassQ), still as
// the camera ray.
template <bool HIDE>
__global__ __launch_bounds__(kStageBlock) void k_post_camera(PassParams pp, SceneView s, RenderBufs rs, int next) {
    const uint32_t n = rs.cnt->ray;
    __shared__ uint32_t s_p[kStageItems][kStageBlock];
    __shared__ uint32_t s_b[kStageItems][kStageBlock];
    const uint32_t items = stage_items(n);
    for (uint32_t i0 = blockIdx.x * items * kStageBlock; i0 < n; i0 += gridDim.x * items * kStageBlock) {
        // items are staged in LDS (each thread reads back only its own) so the
        // per-item body is not unrolled
#pragma unroll 1
        for (int u = 0; u < kStageItems; u++) {
            const uint32_t i = (uint32_t)u < items ? i0 + (uint32_t)u * kStageBlock + threadIdx.x : n;
            bool alive = false;
            int type = 0;
            uint32_t p = 0;
            if (i < n) {
                p = rs.rayq[i];
                const float4 h = rs.hit[p];
                const uint32_t prim = __float_as_uint(h.x);
                const vec3 d = f3(rs.ray_d[p]);
                vec3 rad = v3(0, 0, 0);
                if (prim == 0xffffffffu) {
                    const SceneRec sc = path_scene(s, p);
                    if (sc.env_light >= 0) {
                        const LightRec& e = s.lights[sc.env_light];
                        rad = e.env_visual ? env_eval(s, e, d) : v3(sc.bg[0], sc.bg[1], sc.bg[2]);
                    }
                } else {
                    const uint32_t info = prim_info(s, prim);
                    if (HIDE && info_light(info)) {
                        rs.thr[p] = make_float4(1.f, 1.f, 1.f, 1.f);
                        rs.aux[p] = make_int4(0, (int)(kFlagCont | kFlagCamera), 0, 0);
                        alive = true;
                        type = kPassQ;
                    } else {
                        if (info_light(info)) {
                            vec3 hp, hn;
                            hit_point_normal(s, prim, h.z, h.w, hp, hn);
                            const vec3 le =
                                dot(hn, neg(d)) > 0.f ? pcol(s.lights[light_of(s, prim)].radiance) : v3(0, 0, 0);
                            rad = add(rad, mul(v3(1.f, 1.f, 1.f), le));
                        }
                        alive = !(pp.max_depth >= 0 && 0 + 2 > pp.max_depth);
                        type = info_type(info);
                    }
                }
                rs.rad[p] = f4(rad, 0.f);
            }
            s_p[u][threadIdx.x] = p;
            s_b[u][threadIdx.x] = alive ? 1u << type : 0u;
        }
        uint32_t pv[kStageItems], bv[kStageItems];
#pragma unroll
        for (int u = 0; u < kStageItems; u++) {
            pv[u] = s_p[u][threadIdx.x];
            bv[u] = s_b[u][threadIdx.x];
        }
        uint32_t* const mq = rs.matq[next];
        const uint32_t cap = pp.cap;
        block_append<HIDE ? kPassQ + 1 : DL_BSDF_COUNT, kStageItems>(rs.cnt->mat[next], bv,
                                                                     [&](int t, int u, uint32_t slot) {
            mq[(size_t)t * cap + slot] = pv[u];
        });
    }
}

// Paths whose ray hit a hidden light surface continue from the hit point in
// the same direction (hitgroup_pgs.cu:55-64: fr = 1, eta = 1, p = its.p,
// wo = ray direction): no random numbers, no bounce, the previous vertex's
// BSDF pdf and diffuse flag kept for the MIS of the next emitter hit.
__global__ __launch_bounds__(kStageBlock) void k_pass(PassParams pp, SceneView s, RenderBufs rs, int cur) {
    const uint32_t n = rs.cnt->mat[cur][kPassQ];
    const uint32_t* __restrict__ queue = rs.matq[cur] + (size_t)kPassQ * pp.cap;
    __shared__ uint32_t s_p[kStageItems][kStageBlock];
    const uint32_t items = stage_items(n);
    for (uint32_t i0 = blockIdx.x * items * kStageBlock; i0 < n; i0 += gridDim.x * items * kStageBlock) {
#pragma unroll 1
        for (int u = 0; u < kStageItems; u++) {
            const uint32_t i = (uint32_t)u < items ? i0 + (uint32_t)u * kStageBlock + threadIdx.x : n;
            uint32_t p = 0xffffffffu;
            if (i < n) {
                p = queue[i];
                const float4 h = rs.hit[p];
                Its its;
                make_its(s, __float_as_uint(h.x), h.z, h.w, false, its);
                int4 aux = rs.aux[p];
                aux.y = (int)(((uint32_t)aux.y & (kFlagDiffuse | kFlagCamera)) | kFlagCont);
                rs.aux[p] = aux;
                rs.ray_o[p] = f4(its.p, pp.eps);
                rs.ray_d[p] = f4(f3(rs.ray_d[p]), INFINITY);
            }
            s_p[u][threadIdx.x] = p;
        }
        uint32_t pv[kStageItems], bv[kStageItems];
#pragma unroll
        for (int u = 0; u < kStageItems; u++) {
            pv[u] = s_p[u][threadIdx.x];
            bv[u] = pv[u] != 0xffffffffu ? 1u : 0u;
        }
        block_append<1, kStageItems>(&rs.cnt->ray, bv, [&](int, int u, uint32_t slot) { rs.rayq[slot] = pv[u]; });
    }
}

// The "normals" integrator (src/tracer/integrator/normal.cpp:9-19): the
// component-wise absolute value of the shading normal of the first hit
// (the record of OCTree::rayIntersect), zero on a miss.
__global__ __launch_bounds__(kStageBlock) void k_normals(SceneView s, RenderBufs rs) {
    const uint32_t n = rs.cnt->ray;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t p = rs.rayq[i];
        const float4 h = rs.hit[p];
        const uint32_t prim = __float_as_uint(h.x);
        vec3 rad = v3(0.0f, 0.0f, 0.0f);
        if (prim != 0xffffffffu) {
            Its its;
            make_its(s, prim, h.z, h.w, false, its);
            rad = v3(fabsf(its.sh.n.x), fabsf(its.sh.n.y), fabsf(its.sh.n.z));
        }
        rs.rad[p] = f4(rad, 0.f);
    }
}

// One bounce at the current vertex: emitter sampling + BSDF sampling
// (path.cpp:32-69).
// One kernel instantiation per BSDF type (the material-sorted queue of that
// type), so each carries only its own BSDF's code and registers.
template <int TYPE, bool GUIDED>
__global__ __launch_bounds__(kStageBlock) void k_shade(PassParams pp, SceneView s, RenderBufs rs, int cur,
                                                       GuideView gv, bool first) {
    const uint32_t n = rs.cnt->mat[cur][TYPE];
    const uint32_t* __restrict__ queue = rs.matq[cur] + (size_t)TYPE * pp.cap;
    constexpr int type = TYPE;
    __shared__ uint32_t s_p[kStageItems][kStageBlock];
    __shared__ uint32_t s_b[kStageItems][kStageBlock];
    const uint32_t items = stage_items(n);
    for (uint32_t i0 = blockIdx.x * items * kStageBlock; i0 < n; i0 += gridDim.x * items * kStageBlock) {
        // items are staged in LDS (each thread reads back only its own) so the
        // per-item body is not unrolled
#pragma unroll 1
        for (int u = 0; u < kStageItems; u++) {
            const uint32_t i = (uint32_t)u < items ? i0 + (uint32_t)u * kStageBlock + threadIdx.x : n;
            bool want_shadow = false, want_cont = false;
            uint32_t p = 0;
            if (i < n) {
                p = queue[i];
                const float4 h = rs.hit[p];
                const uint32_t prim = __float_as_uint(h.x);
                const vec3 d = f3(rs.ray_d[p]);
                // the first shading vertex of a camera path: throughput 1,
                // eta 1, aux 0 (not stored by k_raygen)
                float4 th4 = make_float4(1.f, 1.f, 1.f, 1.f);
                int4 aux = make_int4(0, 0, 0, 0);
                if (!first) {
                    th4 = rs.thr[p];
                    aux = rs.aux[p];
                }
                vec3 thr = f3(th4);
                float eta = th4.w;
                Pcg32 rng;
                rng.state = rs.rng[p];
                rng.inc = ((uint64_t)rs.seq[p] << 1u) | 1u;
                Its its;
                int mat;
                float prm[8];
                shading_point(s, prim, h.z, h.w, type, its, mat, prm);
                const float* gm = GUIDED ? gv.mix + (size_t)i * kGMix : nullptr;
                uint32_t flags = 0;
                // ---- emitter sampling (Scene::sampleEmitterDirection, scene.cpp:139-170)
                float lx = rng.next_float();
                const float ly = rng.next_float();
                DirSample ds;
                ds.pdf = 0.0f;
                ds.delta = false;
                vec3 spectrum = v3(0, 0, 0);
                const SceneRec sc = path_scene(s, p);
                const int L = sc.num_lights;
                if (L == 1) {
                    spectrum = emitter_sample(s, s.lights[sc.light0], lx, ly, its, ds, rng);
                } else if (L > 1) {
                    const float epdf = 1.f / (float)L;
                    uint32_t idx = (uint32_t)(lx * (float)L);
                    idx = idx < (uint32_t)(L - 1) ? idx : (uint32_t)(L - 1);
                    lx = lx * (float)L - (float)idx;
                    spectrum = emitter_sample(s, s.lights[sc.light0 + (int)idx], lx, ly, its, ds, rng);
                    spectrum = scale(spectrum, (float)L);
                    ds.pdf *= epdf;
                }
                const vec3 wi = to_local(its.sh, neg(normalized(d)));
                if (ds.pdf > 0) {
                    BRec q;
                    q.wi = wi;
                    q.wo = to_local(its.sh, ds.d);
                    q.eta = 1.f;
                    q.measure = kMeasureSolidAngle;
                    const vec3 bv = bsdf_eval(type, prm, q);
                    float bp = bsdf_pdf(type, prm, q);
                    // with the proposal, directions come from the 50/50 mixture
                    // of BSDF and proposal sampling: its pdf enters the MIS
                    if (GUIDED && gv.frac > 0.0f && guide_material(type, prm))
                        bp = (1.0f - gv.frac) * bp + gv.frac * guide_pdf(gm, ds.d);
                    const float mis = ds.delta ? 1.0f : mis_weight(ds.pdf, bp);
                    const vec3 contrib = mul(mul(scale(thr, mis), bv), spectrum);
                    if (csum(spectrum) > 0) {
                        rs.nee[p] = f4(contrib, 0.f);
                        rs.sh_d[p] = f4(ds.d, ds.dist - pp.eps);  // from ray_o, written below
                        flags |= kFlagNee;
                        want_shadow = true;
                    } else {
                        // (the path's radiance is read and written only here:
                        // post reads it next)
                        rs.rad[p] = f4(add(f3(rs.rad[p]), contrib), 0.f);
                    }
                }
                // ---- BSDF sampling
                BRec br;
                br.wi = wi;
                br.wo = v3(0, 0, 0);
                br.eta = 1.f;
                br.measure = kMeasureUnknown;
                vec3 fr;
                float prev_pdf;
                vec3 wd = v3(0.0f, 0.0f, 0.0f);
                if (GUIDED) {
                    // one-sample MIS: proposal with probability frac, else
                    // the BSDF; the weight divides by the mixture density
                    const float fg = guide_material(type, prm) ? gv.frac : 0.0f;
                    const float us = rng.next_float();
                    const float bx = rng.next_float();
                    const float by = rng.next_float();
                    if (us < fg) {
                        wd = guide_sample(gm, bx, by);
                        br.wo = to_local(its.sh, wd);
                        br.measure = kMeasureSolidAngle;
                        br.eta = 1.0f;
                        const float pb = bsdf_pdf(type, prm, br);
                        const float pm = (1.0f - fg) * pb + fg * guide_pdf(gm, wd);
                        fr = (br.wo.z > 0.0f && pm > 0.0f) ? scale(bsdf_eval(type, prm, br), br.wo.z / pm)
                                                           : v3(0.0f, 0.0f, 0.0f);
                        prev_pdf = pm;
                    } else {
                        const vec3 f0 = bsdf_sample(type, prm, br, bx, by);
                        wd = to_world(its.sh, br.wo);
                        const float pb = bsdf_pdf(type, prm, br);
                        const float pm = fg > 0.0f ? (1.0f - fg) * pb + fg * guide_pdf(gm, wd) : pb;
                        fr = pm > 0.0f ? scale(f0, pb / pm) : v3(0.0f, 0.0f, 0.0f);
                        prev_pdf = pm;
                    }
                } else {
                    const float bx = rng.next_float();
                    const float by = rng.next_float();
                    fr = bsdf_sample(type, prm, br, bx, by);
                    prev_pdf = bsdf_pdf(type, prm, br);
                }
                thr = mul(thr, fr);
                eta *= br.eta;
                if (bsdf_is_diffuse(type)) flags |= kFlagDiffuse;
                // the origin of the continuation and of the shadow ray
                if (want_shadow || !is_zero(thr)) rs.ray_o[p] = f4(its.p, pp.eps);
                if (!is_zero(thr)) {
                    rs.ray_d[p] = f4(to_world(its.sh, br.wo), INFINITY);
                    flags |= kFlagCont;
                    want_cont = true;
                }
                aux.w = 0;
                if (GUIDED && gv.rec && want_cont && guide_material(type, prm)) {
                    // training record of this vertex (its radiance before the
                    // continuation is completed by k_post)
                    const uint32_t depth = (uint32_t)aux.z;
                    const uint32_t slot = (p / gv.rec_stride) * kGRecDepth + depth;
                    if (depth < (uint32_t)kGRecDepth && p % gv.rec_stride == 0 && slot < gv.rec_cap) {
                        float* r = gv.rec + (size_t)slot * kGRecFloats;
                        const vec3 wo = neg(normalized(d));
                        r[kRecPos] = (its.p.x - gv.lo[0]) * gv.inv[0];
                        r[kRecPos + 1] = (its.p.y - gv.lo[1]) * gv.inv[1];
                        r[kRecPos + 2] = (its.p.z - gv.lo[2]) * gv.inv[2];
                        r[kRecWo] = wo.x;
                        r[kRecWo + 1] = wo.y;
                        r[kRecWo + 2] = wo.z;
                        r[kRecN] = its.sh.n.x;
                        r[kRecN + 1] = its.sh.n.y;
                        r[kRecN + 2] = its.sh.n.z;
                        r[kRecRough] = guide_roughness(type, prm);
                        r[kRecDir] = wd.x;
                        r[kRecDir + 1] = wd.y;
                        r[kRecDir + 2] = wd.z;
                        r[kRecPdf] = prev_pdf;
                        r[kRecThr] = thr.x;
                        r[kRecThr + 1] = thr.y;
                        r[kRecThr + 2] = thr.z;
                        r[kRecPath] = __uint_as_float(p);
                        r[kRecValid] = 1.0f;
                        aux.w = (int)slot + 1;
                    }
                }
                rs.thr[p] = f4(thr, eta);
                rs.rng[p] = rng.state;
                aux.x = __float_as_int(prev_pdf);
                aux.y = (int)flags;
                rs.aux[p] = aux;
            }
            s_p[u][threadIdx.x] = p;
            s_b[u][threadIdx.x] = (want_cont ? 1u : 0u) | (want_shadow ? 2u : 0u);
        }
        uint32_t pv[kStageItems], bv[kStageItems];
#pragma unroll
        for (int u = 0; u < kStageItems; u++) {
            pv[u] = s_p[u][threadIdx.x];
            bv[u] = s_b[u][threadIdx.x];
        }
        // counters: ray (continuations), then shadow -- adjacent fields
        block_append<2, kStageItems>(&rs.cnt->ray, bv, [&](int q, int u, uint32_t slot) {
            if (q == 0)
                rs.rayq[slot] = pv[u];
            else
                rs.shq[slot] = pv[u];
        });
    }
}

// NEE resolution, emitter hit / miss with MIS, Russian roulette
// (path.cpp:46-56, 71-97; miss.cu:31-42).
// A continuation that hit an area light: its MIS-weighted emission
// (path.cpp:72-84) is added by k_post_emit, after the rest of k_post, from
// the throughput before Russian roulette, kept in the path's NEE slot (the
// NEE contribution has been consumed) -- this keeps the rarely needed
// emitter pdf out of k_post's registers.
// The lanes of a wave that call it together take consecutive queue slots
// with one atomic (the queue's order does not matter: each entry is one
// path's own emission).
__device__ __forceinline__ void defer_emission(const RenderBufs& rs, uint32_t p, vec3 thr, int light) {
    rs.nee[p] = f4(thr, __int_as_float(light));
    const unsigned long long mask = __ballot(1);
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&rs.cnt->emit, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    rs.emq[base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = p;
}

// Emission of the continuations k_post deferred: rad += thr * mis * Le, in
// the order k_post would have added it (after the NEE contribution).
__global__ __launch_bounds__(kStageBlock) void k_post_emit(SceneView s, RenderBufs rs) {
    const uint32_t n = rs.cnt->emit;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t p = rs.emq[i];
        const float4 h = rs.hit[p];
        const uint32_t prim = __float_as_uint(h.x);
        const float4 t4 = rs.nee[p];
        const int light = __float_as_int(t4.w);
        const int4 aux = rs.aux[p];
        const float prev_pdf = __int_as_float(aux.x);
        const bool diffuse = ((uint32_t)aux.y & kFlagDiffuse) != 0;
        const vec3 o = f3(rs.ray_o[p]);
        const vec3 d = f3(rs.ray_d[p]);
        vec3 hp, hn;
        hit_point_normal(s, prim, h.z, h.w, hp, hn);
        const vec3 ev = dot(hn, neg(d)) > 0.f ? pcol(s.lights[light].radiance) : v3(0, 0, 0);
        const DirSample hs = dir_sample(o, hp, hn);
        const float epdf = pdf_emitter_direction(s, light, path_scene(s, p).num_lights, hs);
        const float mis = diffuse ? mis_weight(prev_pdf, epdf) : 1.f;
        const vec3 rad = add(f3(rs.rad[p]), mul(scale(f3(t4), mis), ev));
        rs.rad[p] = f4(rad, 0.f);
    }
}

// HIDE: light surfaces are hidden (hitgroup_pgs.cu:55-64): a path ray that
// hits one collects its emission as usual, unless it is still the camera
// ray, and passes through (queue kPassQ) without a bounce; a camera ray that
// passed through is treated as the camera ray at its next hit or miss.
template <bool HIDE>
__global__ __launch_bounds__(kStageBlock) void k_post(PassParams pp, SceneView s, RenderBufs rs, int cur,
                                                      float* __restrict__ grec) {
    // the lengths of the queues being read, loaded once: the counters'
    // cache line takes every block's appends to the next queues meanwhile
    uint32_t counts[kTypes];
#pragma unroll
    for (int t = 0; t < kTypes; t++) counts[t] = rs.cnt->mat[cur][t];
    const uint32_t n = queue_total(counts);
    const int next = cur ^ 1;
    __shared__ uint32_t s_p[kStageItems][kStageBlock];
    __shared__ uint32_t s_b[kStageItems][kStageBlock];
    const uint32_t items = stage_items(n);
    for (uint32_t i0 = blockIdx.x * items * kStageBlock; i0 < n; i0 += gridDim.x * items * kStageBlock) {
        // items are staged in LDS (each thread reads back only its own) so the
        // per-item body is not unrolled
#pragma unroll 1
        for (int u = 0; u < kStageItems; u++) {
            const uint32_t i = (uint32_t)u < items ? i0 + (uint32_t)u * kStageBlock + threadIdx.x : n;
            bool alive = false;
            int ntype = 0;
            uint32_t p = 0;
            if (i < n) {
                int type;
                queue_entry(counts, rs.matq[cur], pp.cap, i, p, type);
                int4 aux = rs.aux[p];
                const uint32_t flags = (uint32_t)aux.y;
                vec3 rad = f3(rs.rad[p]);
                if ((flags & kFlagNee) && rs.occ[p] == 0) rad = add(rad, f3(rs.nee[p]));
                if (grec && aux.w > 0) {
                    // radiance before the continuation of a recorded vertex
                    float* r = grec + (size_t)(aux.w - 1) * kGRecFloats;
                    r[kRecRad] = rad.x;
                    r[kRecRad + 1] = rad.y;
                    r[kRecRad + 2] = rad.z;
                }
                if (flags & kFlagCont) {
                    const float4 th4 = rs.thr[p];
                    vec3 thr = f3(th4);
                    const float eta = th4.w;
                    const float prev_pdf = __int_as_float(aux.x);
                    const bool diffuse = (flags & kFlagDiffuse) != 0;
                    const float4 h = rs.hit[p];
                    const uint32_t prim = __float_as_uint(h.x);
                    const vec3 d = f3(rs.ray_d[p]);
                    const bool cam = HIDE && (flags & kFlagCamera) != 0;
                    if (prim == 0xffffffffu) {
                        const SceneRec sc = path_scene(s, p);
                        if (sc.env_light >= 0) {
                            const LightRec& e = s.lights[sc.env_light];
                            if (cam) {
                                rad = add(rad,
                                          mul(thr, e.env_visual ? env_eval(s, e, d) : v3(sc.bg[0], sc.bg[1], sc.bg[2])));
                            } else {
                                float epdf;
                                const vec3 ev = env_eval_pdf(s, e, sc.num_lights, d, epdf);
                                const float mis = diffuse ? mis_weight(prev_pdf, epdf) : 1.f;
                                rad = add(rad, mul(scale(thr, mis), ev));
                            }
                        }
                    } else {
                        const uint32_t info = prim_info(s, prim);
                        if (HIDE && (cam || info_light(info))) {
                            if (info_light(info)) {
                                if (!cam) defer_emission(rs, p, thr, light_of(s, prim));
                                aux.w = 0;  // the record's radiance before its continuation is already kept
                                rs.aux[p] = aux;
                                alive = true;
                                ntype = kPassQ;
                            } else {
                                // the camera ray's first visible surface (as k_post_camera)
                                alive = !(pp.max_depth >= 0 && 0 + 2 > pp.max_depth);
                                ntype = info_type(info);
                            }
                        } else {
                            if (info_light(info)) defer_emission(rs, p, thr, light_of(s, prim));
                            alive = true;
                            const int depth = aux.z;
                            if (depth >= pp.rr_depth) {
                                Pcg32 rng;
                                rng.state = rs.rng[p];
                                rng.inc = ((uint64_t)rs.seq[p] << 1u) | 1u;
                                const float pr = fmin0(cmax(scale(scale(thr, eta), eta)), 0.99f);
                                if (rng.next_float() > pr) {
                                    alive = false;
                                } else {
                                    thr = divs(thr, pr);
                                    rs.thr[p] = f4(thr, eta);
                                }
                                rs.rng[p] = rng.state;
                            }
                            if (alive) {
                                const int nd = depth + 1;
                                if (pp.max_depth >= 0 && nd + 2 > pp.max_depth) alive = false;
                                aux.z = nd;
                                rs.aux[p] = aux;
                                ntype = info_type(info);
                            }
                        }
                    }
                }
                rs.rad[p] = f4(rad, 0.f);
            }
            s_p[u][threadIdx.x] = p;
            s_b[u][threadIdx.x] = alive ? 1u << ntype : 0u;
        }
        uint32_t pv[kStageItems], bv[kStageItems];
#pragma unroll
        for (int u = 0; u < kStageItems; u++) {
            pv[u] = s_p[u][threadIdx.x];
            bv[u] = s_b[u][threadIdx.x];
        }
        uint32_t* const mq = rs.matq[next];
        const uint32_t cap = pp.cap;
        block_append<HIDE ? kPassQ + 1 : DL_BSDF_COUNT, kStageItems>(rs.cnt->mat[next], bv,
                                                                     [&](int t, int u, uint32_t slot) {
            mq[(size_t)t * cap + slot] = pv[u];
        });
    }
}

// Proposal mixtures for the queue of material TYPE: features of each queued
// shading point (position in the scene box, wo, shading normal, roughness),
// the network on the matrix cores (64 points per wave), decoded mixtures
// written per queue entry.
template <int TYPE>
__global__ __launch_bounds__(kGWaves * 64) void k_guide_infer(PassParams pp, SceneView s, RenderBufs rs, int cur,
                                                              GuideView gv, Net net) {
    __shared__ uint16_t s_act[kGWaves][kGRowsPerWave * kGRowStride];
    const uint32_t n = rs.cnt->mat[cur][TYPE];
    const uint32_t* __restrict__ queue = rs.matq[cur] + (size_t)TYPE * pp.cap;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint16_t* sm = s_act[wave];
    float* smf = reinterpret_cast<float*>(sm);
    for (uint32_t tile = blockIdx.x * kGWaves + wave; (size_t)tile * kGRowsPerWave < n; tile += gridDim.x * kGWaves) {
        const uint32_t row0 = tile * kGRowsPerWave, i = row0 + lane;
        vec3 pn = v3(0.f, 0.f, 0.f), wo = v3(0.f, 0.f, 0.f), nn = v3(0.f, 0.f, 0.f);
        float rough = 0.f;
        if (i < n) {
            const uint32_t p = queue[i];
            const float4 h = rs.hit[p];
            Its its;
            int mat;
            float prm[8];
            shading_point(s, __float_as_uint(h.x), h.z, h.w, TYPE, its, mat, prm);
            pn = v3((its.p.x - gv.lo[0]) * gv.inv[0], (its.p.y - gv.lo[1]) * gv.inv[1], (its.p.z - gv.lo[2]) * gv.inv[2]);
            wo = neg(normalized(f3(rs.ray_d[p])));
            nn = its.sh.n;
            rough = guide_roughness(TYPE, prm);
        }
        for (int q = 0; q < kGIn / 8; q++) {
            float f[8];
            if (i < n)
                guide_feature_chunk(pn, wo, nn, rough, q, f);
            else
                for (int j = 0; j < 8; j++) f[j] = 0.0f;
            *reinterpret_cast<bf16x8*>(sm + lane * kGRowStride + q * 8) = pack8(f);
        }
        // raw outputs overwrite the activation rows of their own 32-row block
        // after that block's last MFMA (byte ranges checked: 192 B per output
        // row vs 208 B per activation row)
        forward_rows(net, sm, row0, n, false, nullptr, [&](int r, int c, float v) { smf[r * kGOut + c] = v; });
        if (i < n) {
            float o[kGOut], mix[kGMix];
            for (int j = 0; j < kGOut; j++) o[j] = smf[lane * kGOut + j];
            guide_decode(o, mix);
            float4* dst = reinterpret_cast<float4*>(rs.gmix + (size_t)i * kGMix);
            for (int j = 0; j < kGMix / 4; j++) dst[j] = make_float4(mix[4 * j], mix[4 * j + 1], mix[4 * j + 2], mix[4 * j + 3]);
        }
        __builtin_amdgcn_wave_barrier();  // the next tile overwrites these LDS rows
    }
}

__global__ void k_clear_records(float* rec, uint32_t R) {
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < R; r += gridDim.x * blockDim.x)
        rec[(size_t)r * kGRecFloats + kRecValid] = 0.0f;
}


__device__ __forceinline__ bool sample_valid(const float4& c) {  // Color3f::isValid  color.h:50-58
    return !(c.x < 0 || !isfinite(c.x) || c.y < 0 || !isfinite(c.y) || c.z < 0 || !isfinite(c.z));
}

// Deterministic gather of one pass's samples into the film: pixel q receives
// every sample whose filter footprint, computed exactly as ImageBlock::put
// computes it inside the sample's 32x32 block, covers q.

// Per-sample contribution of sample (px, py) of one layer to pixel (qx, qy):
// the block-local footprint test and weights of ImageBlock::put.
__device__ __forceinline__ void film_add(const dl_film_desc& film, uint32_t W, uint32_t H, int qx, int qy, int px,
                                         int py, const float4& L, const float2& P, float4& acc) {
    const int b = film.border;
    const float r = film.radius;
    const int oy = (py / 32) * 32, sy = min(32, (int)H - oy);
    const int ox = (px / 32) * 32, sx = min(32, (int)W - ox);
    const int yl = qy - (oy - b), xl = qx - (ox - b);
    if (yl < 0 || yl > sy + 2 * b - 1 || xl < 0 || xl > sx + 2 * b - 1) return;
    const float plx = P.x - 0.5f - (float)(ox - b);
    const float ply = P.y - 0.5f - (float)(oy - b);
    const int x0 = max((int)ceilf(plx - r), 0), x1 = min((int)floorf(plx + r), sx + 2 * b - 1);
    const int y0 = max((int)ceilf(ply - r), 0), y1 = min((int)floorf(ply + r), sy + 2 * b - 1);
    if (xl < x0 || xl > x1 || yl < y0 || yl > y1) return;
    const float wx = film.table[(int)(fabsf((float)xl - plx) * film.lookup)];
    const float wy = film.table[(int)(fabsf((float)yl - ply) * film.lookup)];
    acc.x += (L.x * wx) * wy;
    acc.y += (L.y * wx) * wy;
    acc.z += (L.z * wx) * wy;
    acc.w += (1.0f * wx) * wy;
}

// Generic gather (any filter radius): sample layers outermost, then rows and
// columns of the neighbourhood -- the order of the tiled kernel below.
__global__ __launch_bounds__(kStageBlock) void k_film(PassParams pp, uint32_t k, dl_film_desc film, RenderBufs rs) {
    const uint32_t W = pp.W, H = pp.H, WH = W * H;
    const int R = (int)ceilf(film.radius) + 1;
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < WH; q += gridDim.x * blockDim.x) {
        const int qx = (int)(q % W), qy = (int)(q / W);
        float4 acc = rs.film[q];
        uint32_t invalid = 0;
        for (uint32_t j = 0; j < k; ++j)
            for (int py = max(0, qy - R); py <= min((int)H - 1, qy + R); ++py)
                for (int px = max(0, qx - R); px <= min((int)W - 1, qx + R); ++px) {
                    const uint32_t sp = j * WH + (uint32_t)py * W + (uint32_t)px;
                    const float4 L = rs.rad[sp];
                    if (!sample_valid(L)) {
                        if (px == qx && py == qy) invalid++;
                        continue;
                    }
                    film_add(film, W, H, qx, qy, px, py, L, rs.pos[sp], acc);
                }
        rs.film[q] = acc;
        if (invalid) atomicAdd(&rs.cnt->invalid, (unsigned long long)invalid);
    }
}

// Tiled gather for the filters whose footprint reaches at most kFilmHalo - 1
// pixels from a sample's own pixel: one 16x16 pixel tile per block; each
// sample layer's tile plus halo is staged in LDS once, with film_add's
// weights already computed -- for each sample, the x weight of each of the
// 2R + 1 film columns and the y weight of each of the 2R + 1 film rows it
// can reach (0 where its block-local footprint does not cover the pixel, and
// radiance 0 for an invalid sample or a position outside the image).  A
// (pixel, sample) pair then costs three LDS reads and film_add's products
// and sums, without a test: a zero weight adds +0 or -0, which leaves the
// accumulator unchanged (it is never -0).  Samples are added in k_film's
// order (layers, then rows, then columns), so the pixel sums are those of
// film_add bit for bit.
constexpr int kFilmHalo = 5;
constexpr int kFilmTile = 16;
template <int R>
__global__ __launch_bounds__(kFilmTile* kFilmTile) void k_film_tiled(PassParams pp, uint32_t k, dl_film_desc film,
                                                                     RenderBufs rs) {
    constexpr int S = kFilmTile + 2 * R;  // staged span
    constexpr int D = 2 * R + 1;          // film columns (rows) a sample can reach
    __shared__ float4 s_rad[S * S];
    // s_wx[d][e]: weight of staged sample e for film column px - R + d
    // (s_wy: row py - R + d)
    __shared__ float s_wx[D][S * S];
    __shared__ float s_wy[D][S * S];
    __shared__ float s_table[33];  // the filter table, indexed per lane
    if (threadIdx.x < 33) s_table[threadIdx.x] = film.table[threadIdx.x];
    __syncthreads();
    const uint32_t W = pp.W, H = pp.H, WH = W * H;
    const int b = film.border;
    const float r = film.radius;
    const int tx0 = (int)blockIdx.x * kFilmTile, ty0 = (int)blockIdx.y * kFilmTile;
    const int lx = (int)(threadIdx.x % kFilmTile), ly = (int)(threadIdx.x / kFilmTile);
    const int qx = tx0 + lx, qy = ty0 + ly;
    const bool inside = qx < (int)W && qy < (int)H;
    const uint32_t q = inside ? (uint32_t)qy * W + (uint32_t)qx : 0u;
    float4 acc = inside ? rs.film[q] : make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t invalid = 0;
    for (uint32_t j = 0; j < k; ++j) {
        for (int e = (int)threadIdx.x; e < S * S; e += (int)blockDim.x) {
            const int gx = tx0 - R + e % S, gy = ty0 - R + e / S;
            float4 L = make_float4(0.f, 0.f, 0.f, 0.f);
            float wx[D], wy[D];
#pragma unroll
            for (int d = 0; d < D; d++) wx[d] = wy[d] = 0.0f;
            if (gx >= 0 && gx < (int)W && gy >= 0 && gy < (int)H) {
                const uint32_t sp = j * WH + (uint32_t)gy * W + (uint32_t)gx;
                const float4 Ls = rs.rad[sp];
                const float2 P = rs.pos[sp];
                if (sample_valid(Ls)) {
                    L = Ls;
                    // film_add's terms (block.cpp:68-128)
                    const int oy = (gy / 32) * 32, sy = min(32, (int)H - oy);
                    const int ox = (gx / 32) * 32, sx = min(32, (int)W - ox);
                    const float plx = P.x - 0.5f - (float)(ox - b);
                    const float ply = P.y - 0.5f - (float)(oy - b);
                    const int x0 = max((int)ceilf(plx - r), 0), x1 = min((int)floorf(plx + r), sx + 2 * b - 1);
                    const int y0 = max((int)ceilf(ply - r), 0), y1 = min((int)floorf(ply + r), sy + 2 * b - 1);
#pragma unroll
                    for (int d = 0; d < D; d++) {
                        const int xl = gx - R + d - (ox - b), yl = gy - R + d - (oy - b);
                        if (xl >= x0 && xl <= x1) wx[d] = s_table[(int)(fabsf((float)xl - plx) * film.lookup)];
                        if (yl >= y0 && yl <= y1) wy[d] = s_table[(int)(fabsf((float)yl - ply) * film.lookup)];
                    }
                } else if (gx >= tx0 && gx < tx0 + kFilmTile && gy >= ty0 && gy < ty0 + kFilmTile) {
                    invalid++;
                }
            }
            s_rad[e] = L;
#pragma unroll
            for (int d = 0; d < D; d++) {
                s_wx[d][e] = wx[d];
                s_wy[d][e] = wy[d];
            }
        }
        __syncthreads();
        // the sample at staged column lx + cx lies R - cx columns left of
        // the pixel: the pixel is its film column px - R + (2R - cx)
#pragma unroll
        for (int cy = 0; cy < D; cy++) {
#pragma unroll
            for (int cx = 0; cx < D; cx++) {
                const int e = (ly + cy) * S + lx + cx;
                const float4 L = s_rad[e];
                const float wx = s_wx[2 * R - cx][e], wy = s_wy[2 * R - cy][e];
                acc.x += (L.x * wx) * wy;
                acc.y += (L.y * wx) * wy;
                acc.z += (L.z * wx) * wy;
                acc.w += (1.0f * wx) * wy;
            }
        }
        __syncthreads();
    }
    if (inside) rs.film[q] = acc;
    if (invalid) atomicAdd(&rs.cnt->invalid, (unsigned long long)invalid);
}

// The film gather of one pass's k sample layers: the tiled gather for
// footprints of up to kFilmHalo - 1 pixels, else the generic one.
static void film_gather(const PassParams& pp, uint32_t k, const dl_film_desc& film, const RenderBufs& bufs,
                        int stage_grid, hipStream_t stream) {
    // A sample at P in [px, px + 1] reaches the film pixels with
    // ceil(P - 1/2 - r) <= q <= floor(P - 1/2 + r), i.e. within
    // floor(r + 1/2) of its own pixel px.  P -/+ r is rounded once in
    // binary32 (block-relative, |P| < 64: by at most 2^-19), which can reach
    // one pixel further when r + 1/2 lies just below an integer.
    // (the radius is finite and positive: scene upload checks it)
    const float reach = film.radius + 0.5f;
    int R = kFilmHalo;
    if (reach < (float)kFilmHalo) {
        R = (int)floorf(reach);
        if (reach - (float)R > 0.9999f) R++;
    }
    const dim3 fgrid((pp.W + kFilmTile - 1) / kFilmTile, (pp.H + kFilmTile - 1) / kFilmTile);
    const int threads = kFilmTile * kFilmTile;
    switch (R) {
        case 0: k_film_tiled<0><<<fgrid, threads, 0, stream>>>(pp, k, film, bufs); break;
        case 1: k_film_tiled<1><<<fgrid, threads, 0, stream>>>(pp, k, film, bufs); break;
        case 2: k_film_tiled<2><<<fgrid, threads, 0, stream>>>(pp, k, film, bufs); break;
        case 3: k_film_tiled<3><<<fgrid, threads, 0, stream>>>(pp, k, film, bufs); break;
        case 4: k_film_tiled<4><<<fgrid, threads, 0, stream>>>(pp, k, film, bufs); break;
        default: k_film<<<stage_grid, kStageBlock, 0, stream>>>(pp, k, film, bufs); break;
    }
}

// ImageBlock::toBitmap / Color4f::divideByFilterWeight  block.cpp:58-66, color.h:120-125
__global__ __launch_bounds__(kStageBlock) void k_resolve(const float4* film, float* rgb, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 c = film[i];
        if (c.w != 0) {
            rgb[3 * i] = c.x / c.w;
            rgb[3 * i + 1] = c.y / c.w;
            rgb[3 * i + 2] = c.z / c.w;
        } else {
            rgb[3 * i] = rgb[3 * i + 1] = rgb[3 * i + 2] = 0.0f;
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
void free_render_state(Ctx* c) {
    RenderState* rs = (RenderState*)c->rs;
    if (!rs) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void* p : rs->allocs) (void)dfree(c, p);
    if (rs->cnt_host) (void)hipHostFree(rs->cnt_host);
    for (hipEvent_t e : rs->evs) (void)hipEventDestroy(e);
    delete rs;
    c->rs = nullptr;
}

void invalidate_film(Ctx* c) {
    if (c->rs) ((RenderState*)c->rs)->film_valid = false;
}

void free_film_slots(Ctx* c) {
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& f : c->film_slots) {
        (void)dfree(c, f.film);
        (void)dfree(c, f.rgb);
    }
    c->film_slots.clear();
}

namespace {

template <typename T>
dl_status alloc(Ctx* c, RenderState* rs, T** p, size_t count) {
    void* q = nullptr;
    DL_HIP_CHECK(c, dmalloc(c, (void**)&q, std::max<size_t>(1, count) * sizeof(T)));
    rs->allocs.push_back(q);
    *p = (T*)q;
    return DL_OK;
}

#define DL_TRY(expr)                \
    do {                            \
        dl_status s_ = (expr);      \
        if (s_ != DL_OK) return s_; \
    } while (0)

dl_status ensure_state(Ctx* c, uint32_t P, uint32_t pixels) {
    RenderState* rs = (RenderState*)c->rs;
    if (rs && rs->capacity >= P && rs->film_pixels >= pixels) return DL_OK;
    free_render_state(c);
    rs = new RenderState();
    c->rs = rs;
    rs->capacity = P;
    rs->film_pixels = pixels;
    DL_TRY(alloc(c, rs, &rs->b.ray_o, P));
    DL_TRY(alloc(c, rs, &rs->b.ray_d, P));
    DL_TRY(alloc(c, rs, &rs->b.hit, P));
    DL_TRY(alloc(c, rs, &rs->b.sh_d, P));
    DL_TRY(alloc(c, rs, &rs->b.thr, P));
    DL_TRY(alloc(c, rs, &rs->b.rad, P));
    DL_TRY(alloc(c, rs, &rs->b.nee, P));
    DL_TRY(alloc(c, rs, &rs->b.aux, P));
    DL_TRY(alloc(c, rs, &rs->b.rng, P));
    DL_TRY(alloc(c, rs, &rs->b.seq, P));
    DL_TRY(alloc(c, rs, &rs->b.pos, P));
    DL_TRY(alloc(c, rs, &rs->b.occ, P));
    DL_TRY(alloc(c, rs, &rs->b.sid, P));
    DL_TRY(alloc(c, rs, &rs->b.matq[0], (size_t)P * kTypes));
    DL_TRY(alloc(c, rs, &rs->b.matq[1], (size_t)P * kTypes));
    DL_TRY(alloc(c, rs, &rs->b.rayq, P));
    DL_TRY(alloc(c, rs, &rs->b.shq, P));
    DL_TRY(alloc(c, rs, &rs->b.emq, P));
    DL_TRY(alloc(c, rs, &rs->b.cnt, 1));
    DL_TRY(alloc(c, rs, &rs->b.film, pixels));
    DL_TRY(alloc(c, rs, &rs->b.rgb, (size_t)pixels * 3));
    DL_HIP_CHECK(c, hipHostMalloc((void**)&rs->cnt_host, sizeof(Counters), hipHostMallocDefault));
    return DL_OK;
}

constexpr uint32_t kGuideRecords = dl_sizes::kGuideRecordSlots;  // training record slots per pass (96 B each)
static_assert(kGRecFloats == dl_sizes::kGuideRecordFloats && kGMix == dl_sizes::kGuideMixFloats,
              "guided buffer sizes the render service reserves");
// the per-path bytes ensure_state allocates, as the render service reserves them
static_assert(8 * sizeof(float4) + sizeof(uint64_t) + sizeof(uint32_t) + sizeof(float2) + 2 * sizeof(uint8_t) +
                      2 * kTypes * sizeof(uint32_t) + 3 * sizeof(uint32_t) ==
                  dl_sizes::kStateBytesPerPath &&
              kTypes == dl_sizes::kMatQueueSegs && sizeof(float4) + 3 * sizeof(float) == dl_sizes::kFilmBytesPerPixel,
              "path state bytes the render service reserves");

dl_status ensure_guide_bufs(Ctx* c, RenderState* rs, uint32_t cap) {
    if (rs->gmix_cap < cap) {
        DL_TRY(alloc(c, rs, &rs->b.gmix, (size_t)cap * kGMix));
        rs->gmix_cap = cap;
    }
    if (rs->grec_cap < kGuideRecords) {
        DL_TRY(alloc(c, rs, &rs->b.grec, (size_t)kGuideRecords * kGRecFloats));
        rs->grec_cap = kGuideRecords;
    }
    return DL_OK;
}

// The kernels' view of a device layout; the light count, environment and
// background are those of a layout holding one scene (a layout holding
// several gives them per scene, in SceneView::scenes).
SceneView layout_view(const DevScene& d, int num_lights, int env_light, const float* bg) {
    SceneView s;
    s.pos = d.pos;
    s.nrm = d.nrm;
    s.tan = d.tan;
    s.idx = d.idx;
    s.prim_mesh = d.prim_mesh;
    s.prim_info = d.prim_info;
    s.meshes = (const MeshRec*)d.meshes;
    s.mats = (const dl_material_desc*)d.materials;
    s.lights = (const LightRec*)d.lights;
    s.light_cdf = d.light_cdf;
    s.light_guide = d.light_guide;
    s.env_rgb = d.env_rgb;
    s.env_q = d.env_q;
    s.env_alias = d.env_alias;
    s.env_pdf = d.env_pdf;
    s.uv = d.uv;
    s.texels = d.texels;
    s.tex_info = d.tex_info;
    s.num_lights = num_lights;
    s.env_light = env_light;
    for (int k = 0; k < 3; k++) s.bg[k] = bg[k];
    s.scenes = nullptr;
    s.sid = nullptr;
    return s;
}

SceneView scene_view(Ctx* c) {
    return layout_view(c->ds, (int)c->light_recs.size(), c->env_light, c->view_set ? c->view.bg_color : c->bg);
}

struct Timer {
    Ctx* c;
    RenderState* rs;
    std::vector<int>* tags;
    dl_status start(int tag) {
        const size_t k = tags->size();
        while (rs->evs.size() < 2 * (k + 1)) {
            hipEvent_t e;
            DL_HIP_CHECK(c, hipEventCreate(&e));
            rs->evs.push_back(e);
        }
        tags->push_back(tag);
        DL_HIP_CHECK(c, hipEventRecord(rs->evs[2 * k], c->stream));
        return DL_OK;
    }
    dl_status stop() {
        const size_t k = tags->size() - 1;
        DL_HIP_CHECK(c, hipEventRecord(rs->evs[2 * k + 1], c->stream));
        return DL_OK;
    }
};

enum { kTagClosest = 0, kTagShadow = 1, kTagShade = 2, kTagFilm = 3, kTagOther = 4, kTagGuide = 5, kTagTrain = 6 };

}  // namespace

// The closest-hit trace of the camera rays k_raygen queued, the first hits,
// then bounces (shading, continuation and shadow traces, post) until every
// path of the pass has ended.
dl_status run_bounces(Ctx* c, RenderState* rs, const PassParams& pp, const SceneView& sv, const BvhView& bvh,
                      bool count, bool normals, bool hide, bool guided, const bool* guide_types, const GuideView& gv,
                      const Net& net, int guide_grid, int stage_grid, int trace_grid, Timer& tm, double& closest,
                      double& shadow) {
    DL_TRY(tm.start(kTagClosest));
    if (count)
        k_trace_queue<true><<<trace_grid, kTraceBlock, 0, c->stream>>>(bvh, rs->b, c->stack_ovf);
    else
        k_trace_queue<false><<<trace_grid, kTraceBlock, 0, c->stream>>>(bvh, rs->b, c->stack_ovf);
    DL_TRY(tm.stop());
    DL_TRY(tm.start(kTagShade));
    if (normals)
        k_normals<<<stage_grid, kStageBlock, 0, c->stream>>>(sv, rs->b);
    else if (hide)
        k_post_camera<true><<<stage_grid, kStageBlock, 0, c->stream>>>(pp, sv, rs->b, 0);
    else
        k_post_camera<false><<<stage_grid, kStageBlock, 0, c->stream>>>(pp, sv, rs->b, 0);
    DL_TRY(tm.stop());
    DL_HIP_CHECK(c, hipGetLastError());
    if (normals) return DL_OK;  // no bounces
    // One bounce: shading of the paths in queue `cur` (whose lengths per
    // material type are `len`; `first`: the camera paths' first shading
    // vertices), their continuation and shadow traces, and post (which
    // queues the survivors in cur ^ 1).  Every kernel reads its
    // queue length from the device counters; the host skips the shading
    // kernels of empty queues.
    auto bounce = [&](int cur, const uint32_t* len, bool first) -> dl_status {
        k_reset<<<1, kStageBlock, 0, c->stream>>>(rs->b.cnt, cur ^ 1);
        DL_TRY(tm.start(kTagShade));
        if (hide && len[kPassQ] > 0) k_pass<<<stage_grid, kStageBlock, 0, c->stream>>>(pp, sv, rs->b, cur);
        for (int t = 0; t < DL_BSDF_COUNT; t++) {
            if (!rs->types_present[t] || len[t] == 0) continue;
            if (guided && guide_types[t]) {
                DL_TRY(tm.stop());
                DL_TRY(tm.start(kTagGuide));
                if (gv.frac > 0.0f) switch (t) {
                    case DL_BSDF_DIFFUSE: k_guide_infer<DL_BSDF_DIFFUSE><<<guide_grid, kGWaves * 64, 0, c->stream>>>(pp, sv, rs->b, cur, gv, net); break;
                    case DL_BSDF_MICROFACET: k_guide_infer<DL_BSDF_MICROFACET><<<guide_grid, kGWaves * 64, 0, c->stream>>>(pp, sv, rs->b, cur, gv, net); break;
                    case DL_BSDF_ANISO_GGX: k_guide_infer<DL_BSDF_ANISO_GGX><<<guide_grid, kGWaves * 64, 0, c->stream>>>(pp, sv, rs->b, cur, gv, net); break;
                    case DL_BSDF_CONDUCTOR: k_guide_infer<DL_BSDF_CONDUCTOR><<<guide_grid, kGWaves * 64, 0, c->stream>>>(pp, sv, rs->b, cur, gv, net); break;
                }
                DL_TRY(tm.stop());
                DL_TRY(tm.start(kTagShade));
                switch (t) {
                    case DL_BSDF_DIFFUSE: k_shade<DL_BSDF_DIFFUSE, true><<<stage_grid, kStageBlock, 0, c->stream>>>(pp, sv, rs->b, cur, gv, first); break;
                    case DL_BSDF_MICROFACET: k_shade<DL_BSDF_MICROFACET, true><<<stage_grid, kStageBlock, 0, c->stream>>>(pp, sv, rs->b, cur, gv, first); break;
                    case DL_BSDF_ANISO_GGX: k_shade<DL_BSDF_ANISO_GGX, true><<<stage_grid, kStageBlock, 0, c->stream>>>(pp, sv, rs->b, cur, gv, first); break;
                    case DL_BSDF_CONDUCTOR: k_shade<DL_BSDF_CONDUCTOR, true><<<stage_grid, kStageBlock, 0, c->stream>>>(pp, sv, rs->b, cur, gv, first); break;
                }
                continue;
            }
            switch (t) {
                case DL_BSDF_DIFFUSE: k_shade<DL_BSDF_DIFFUSE, false><<<stage_grid, kStageBlock, 0, c->stream>>>(pp, sv, rs->b, cur, gv, first); break;
                case DL_BSDF_MIRROR: k_shade<DL_BSDF_MIRROR, false><<<stage_grid, kStageBlock, 0, c->stream>>>(pp, sv, rs->b, cur, gv, first); break;
                case DL_BSDF_DIELECTRIC: k_shade<DL_BSDF_DIELECTRIC, false><<<stage_grid, kStageBlock, 0, c->stream>>>(pp, sv, rs->b, cur, gv, first); break;
                case DL_BSDF_MICROFACET: k_shade<DL_BSDF_MICROFACET, false><<<stage_grid, kStageBlock, 0, c->stream>>>(pp, sv, rs->b, cur, gv, first); break;
                case DL_BSDF_ANISO_GGX: k_shade<DL_BSDF_ANISO_GGX, false><<<stage_grid, kStageBlock, 0, c->stream>>>(pp, sv, rs->b, cur, gv, first); break;
                case DL_BSDF_CONDUCTOR: k_shade<DL_BSDF_CONDUCTOR, false><<<stage_grid, kStageBlock, 0, c->stream>>>(pp, sv, rs->b, cur, gv, first); break;
            }
        }
        DL_TRY(tm.stop());
        DL_TRY(tm.start(kTagClosest));
        if (count)
            k_trace_queue<true><<<trace_grid, kTraceBlock, 0, c->stream>>>(bvh, rs->b, c->stack_ovf);
        else
            k_trace_queue<false><<<trace_grid, kTraceBlock, 0, c->stream>>>(bvh, rs->b, c->stack_ovf);
        DL_TRY(tm.stop());
        DL_TRY(tm.start(kTagShadow));
        if (count)
            k_shadow_queue<true><<<trace_grid, kTraceBlock, 0, c->stream>>>(bvh, rs->b, c->stack_ovf);
        else
            k_shadow_queue<false><<<trace_grid, kTraceBlock, 0, c->stream>>>(bvh, rs->b, c->stack_ovf);
        DL_TRY(tm.stop());
        DL_TRY(tm.start(kTagShade));
        if (hide)
            k_post<true><<<stage_grid, kStageBlock, 0, c->stream>>>(pp, sv, rs->b, cur, gv.rec);
        else
            k_post<false><<<stage_grid, kStageBlock, 0, c->stream>>>(pp, sv, rs->b, cur, gv.rec);
        k_post_emit<<<stage_grid, kStageBlock, 0, c->stream>>>(sv, rs->b);
        DL_TRY(tm.stop());
        DL_HIP_CHECK(c, hipGetLastError());
        return DL_OK;
    };
    // The counters after bounce k - 1 (k = 0: after the first hits) give the
    // queue lengths bounce k shades and the rays bounce k - 1 traced; the
    // host reads them before it launches bounce k.  (Queuing bounce k + 1
    // before reading bounce k's counters -- every kernel exits on an empty
    // queue -- removes this wait but launches one bounce over empty queues
    // at the end and every material type each bounce: measured no faster on
    // the 1080p workload and 2-3% slower on small renders, so the host waits.)
    for (int k = 0;; ++k) {
        const int cur = k & 1;
        DL_HIP_CHECK(c, hipMemcpyAsync(rs->cnt_host, rs->b.cnt, sizeof(Counters), hipMemcpyDeviceToHost, c->stream));
        DL_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        const Counters& s = *rs->cnt_host;
        if (k > 0) {
            closest += s.ray;
            shadow += s.shadow;
        }
        uint32_t live = 0;
        for (int t = 0; t < kTypes; t++) live += s.mat[cur][t];
        if (live == 0) break;
        uint32_t len[kTypes];
        for (int t = 0; t < kTypes; t++) len[t] = s.mat[cur][t];
        DL_TRY(bounce(cur, len, k == 0));
    }
    return DL_OK;
}

dl_status render(Ctx* c, const dl_render_params* prm, float* rgb_host, dl_render_stats* stats, int paths_sample,
                 float* paths_out) {
    if (prm->spp <= 0) return set_error(c, DL_ERR_INVALID, "spp must be positive");
    if (prm->guide_steps < 0 || prm->guide_steps > 4096)
        return set_error(c, DL_ERR_INVALID, "guide_steps must be in [0, 4096]");
    if (c->light_recs.empty()) return set_error(c, DL_ERR_INVALID, "There is not emitter in the scene.");
    const bool normals = (prm->flags & DL_RENDER_NORMALS) != 0;
    if (normals && (prm->flags & DL_RENDER_GUIDED))
        return set_error(c, DL_ERR_INVALID, "DL_RENDER_NORMALS and DL_RENDER_GUIDED exclude each other");
    const bool hide = c->view_set && c->view.hide_lights != 0;
    const float eps = c->view_set && c->view.epsilon > 0.0f ? c->view.epsilon : kEpsilon;
    (void)hipSetDevice(c->device);
    const uint32_t W = (uint32_t)c->cam.width, H = (uint32_t)c->cam.height, WH = W * H;
    // Default pass: at most 2^27 paths (64 spp at 1920x1080, ~30 GB of path
    // state); a caller may ask for up to 2^29 (256 spp at 1080p, ~120 GB).
    // The last bounces of a pass hold few paths but still cost a traversal
    // of their slowest ray each, so larger passes amortise that tail
    // (measured at 1080p on the 2M-triangle displaced sphere: 32 and 64 spp
    // per pass give 1053 and 1107 Mpaths/s; with later kernels 64, 128 and
    // 256 give 1349, 1407 and 1448-1493; at 1280x720 on the MIS plates 128
    // spp per pass is 2% slower than 64).
    uint32_t k = prm->spp_per_pass > 0 ? (uint32_t)prm->spp_per_pass : std::max(1u, (128u << 20) / WH);
    k = std::min<uint32_t>(k, (uint32_t)prm->spp);
    if (paths_sample >= 0) k = 1;
    if ((uint64_t)WH * k > (1ull << 29)) return set_error(c, DL_ERR_INVALID, "pass too large");
    const uint32_t cap = WH * k;
    DL_TRY(ensure_state(c, cap, WH));
    RenderState* rs = (RenderState*)c->rs;
    rs->b.roots = nullptr;  // one scene: every ray starts at node 0
    const SceneView sv = scene_view(c);
    for (int t = 0; t < 8; t++) rs->types_present[t] = false;
    bool guide_types[8] = {false, false, false, false, false, false, false, false};
    for (const auto& m : c->mats) {
        rs->types_present[m.type] = true;
        if (guide_material(m.type, m.params)) guide_types[m.type] = true;
    }
    const BvhView bvh{c->ds.nodes, c->ds.tris};
    const bool count = stats && (prm->flags & DL_RENDER_COUNT_TRAVERSAL);
    const bool guided = (prm->flags & DL_RENDER_GUIDED) != 0;
    const bool train = guided && !(prm->flags & DL_RENDER_GUIDE_FROZEN) && paths_sample < 0;
    GuideView gv{};
    Net net{};
    if (guided) {
        if (!c->guide) DL_TRY(guide_init(c, 0x5EEDull + prm->seed, 1e-2f, nullptr));
        DL_TRY(ensure_guide_bufs(c, rs, cap));
        guide_net_ptrs(c, &net.w, &net.wt, &net.bias);
        gv.mix = rs->b.gmix;
        gv.frac = kGuideFraction;
        gv.rec = train ? rs->b.grec : nullptr;
        gv.rec_cap = kGuideRecords;
        for (int k = 0; k < 3; k++) {
            gv.lo[k] = c->bbox_lo[k];
            gv.inv[k] = c->bbox_inv[k];
        }
    }
    // 53 KB of LDS per block (4 waves x 64 rows): three blocks fit a CU
    const int guide_grid = std::max(1, c->num_cus) * 3;
    const int stage_grid = std::max(1, c->num_cus) * 8;
    const int trace_grid = (int)(count ? c->trace_grid_count : c->trace_grid);
    std::vector<int> tags;
    Timer tm{c, rs, &tags};
    double closest = 0, shadow = 0, paths = 0;

    DL_HIP_CHECK(c, hipEventRecord(c->ev0, c->stream));
    k_reset_all<<<1, kStageBlock, 0, c->stream>>>(rs->b.cnt);
    if (!(prm->flags & DL_RENDER_ACCUMULATE) || !rs->film_valid)
        k_clear_film<<<stage_grid, kStageBlock, 0, c->stream>>>(rs->b.film, WH);
    rs->film_valid = paths_sample < 0;
    const uint32_t total_spp = paths_sample >= 0 ? 1u : (uint32_t)prm->spp;
    // A training render without a fixed pass size starts with short passes
    // (1, 1, 2, 4, ... samples per pixel up to k) so the network, trained
    // after every pass, improves early; a pass of a network that has never
    // been trained samples the BSDF only (its records start the training).
    const bool schedule = train && prm->spp_per_pass <= 0;

    double train_steps = 0;
    uint32_t pass = 0;
    for (uint32_t done = 0; done < total_spp; ++pass) {
        uint32_t kk = std::min(k, total_spp - done);
        if (schedule) kk = std::min(kk, pass == 0 ? 1u : 1u << std::min<uint32_t>(pass - 1, 30));
        if (guided) gv.frac = (train && guide_step_count(c) == 0) ? 0.0f : kGuideFraction;
        PassParams pp{};
        pp.P = WH * kk;
        pp.W = W;
        pp.H = H;
        pp.sample_base = paths_sample >= 0 ? (uint32_t)paths_sample : (uint32_t)prm->first_sample + done;
        pp.seed = prm->seed;
        pp.max_depth = prm->max_depth;
        pp.rr_depth = prm->rr_depth;
        pp.cap = cap;
        pp.eps = eps;
        paths += pp.P;
        closest += pp.P;
        if (train) {
            // one record slot per (recorded path, depth < kGRecDepth)
            gv.rec_stride = std::max<uint32_t>(1u, (uint32_t)(((uint64_t)pp.P * kGRecDepth + kGuideRecords - 1) /
                                                              kGuideRecords));
            k_clear_records<<<std::max(1, c->num_cus) * 4, 256, 0, c->stream>>>(rs->b.grec, kGuideRecords);
        }
        k_reset<<<1, kStageBlock, 0, c->stream>>>(rs->b.cnt, 0);
        k_reset<<<1, kStageBlock, 0, c->stream>>>(rs->b.cnt, 1);
        DL_TRY(tm.start(kTagOther));
        k_raygen<<<stage_grid, kStageBlock, 0, c->stream>>>(pp, c->cam, rs->b, 0u, pp.P);
        DL_TRY(tm.stop());
        DL_TRY(run_bounces(c, rs, pp, sv, bvh, count, normals, hide, guided, guide_types, gv, net, guide_grid,
                           stage_grid, trace_grid, tm, closest, shadow));
        if (paths_sample >= 0) break;
        if (train) {
            DL_TRY(tm.start(kTagTrain));
            const int steps_per_pass = prm->guide_steps > 0                          ? prm->guide_steps
                                       : guide_step_count(c) < kGuideWarmSteps ? kGuideWarmStepsPerPass
                                                                                : kGuideStepsPerPass;
            DL_TRY(guide_train_records(c, rs->b.grec, kGuideRecords, rs->b.rad, steps_per_pass));
            DL_TRY(tm.stop());
            train_steps += steps_per_pass;
        }
        DL_TRY(tm.start(kTagFilm));
        film_gather(pp, kk, c->film, rs->b, stage_grid, c->stream);
        DL_TRY(tm.stop());
        done += kk;
    }
    if (paths_sample < 0 && rgb_host) k_resolve<<<stage_grid, kStageBlock, 0, c->stream>>>(rs->b.film, rs->b.rgb, WH);
    DL_HIP_CHECK(c, hipGetLastError());
    DL_HIP_CHECK(c, hipEventRecord(c->ev1, c->stream));
    DL_HIP_CHECK(c, hipEventSynchronize(c->ev1));
    if (paths_out) {
        std::vector<float2> pos(WH);
        std::vector<float4> rad(WH);
        DL_HIP_CHECK(c, hipMemcpy(pos.data(), rs->b.pos, WH * sizeof(float2), hipMemcpyDeviceToHost));
        DL_HIP_CHECK(c, hipMemcpy(rad.data(), rs->b.rad, WH * sizeof(float4), hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < WH; i++) {
            paths_out[5 * i] = pos[i].x;
            paths_out[5 * i + 1] = pos[i].y;
            paths_out[5 * i + 2] = rad[i].x;
            paths_out[5 * i + 3] = rad[i].y;
            paths_out[5 * i + 4] = rad[i].z;
        }
    }
    if (rgb_host) DL_HIP_CHECK(c, hipMemcpy(rgb_host, rs->b.rgb, (size_t)WH * 3 * sizeof(float), hipMemcpyDeviceToHost));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        float ms = 0.f;
        DL_HIP_CHECK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        stats->ms_total = ms;
        for (size_t i = 0; i < tags.size(); i++) {
            float e = 0.f;
            DL_HIP_CHECK(c, hipEventElapsedTime(&e, rs->evs[2 * i], rs->evs[2 * i + 1]));
            switch (tags[i]) {
                case kTagClosest: stats->ms_closest += e; stats->launches_closest += 1; break;
                case kTagShadow: stats->ms_shadow += e; stats->launches_shadow += 1; break;
                case kTagShade: stats->ms_shade += e; break;
                case kTagFilm: stats->ms_film += e; break;
                case kTagGuide: stats->ms_guide += e; break;
                case kTagTrain: stats->ms_train += e; break;
                default: stats->ms_other += e; break;
            }
        }
        DL_HIP_CHECK(c, hipMemcpy(rs->cnt_host, rs->b.cnt, sizeof(Counters), hipMemcpyDeviceToHost));
        stats->paths = paths;
        stats->closest_rays = closest;
        stats->shadow_rays = shadow;
        stats->invalid_samples = (double)rs->cnt_host->invalid;
        stats->nodes_closest = (double)rs->cnt_host->nodes;
        stats->tris_closest = (double)rs->cnt_host->tris;
        stats->nodes_shadow = (double)rs->cnt_host->nodes_sh;
        stats->tris_shadow = (double)rs->cnt_host->tris_sh;
        stats->train_steps = train_steps;
    }
    return DL_OK;
}

// One pass shared by several images of the current scene: each image's
// camera paths are generated into consecutive ranges of the path state (its
// pcg32 streams keyed by its own pixels, samples and seed), every path then
// goes through the same queues and kernels as in render(), and each image's
// film gathers its own range with the same kernel and order as a pass of
// render() -- so an image equals its dl_hip_render with the same pass sizes.
dl_status render_batch(Ctx* c, const dl_render_params* prm, const dl_batch_item* items, uint32_t n,
                       dl_render_stats* stats) {
    if (prm->flags & DL_RENDER_GUIDED) return set_error(c, DL_ERR_INVALID, "shared passes do not take DL_RENDER_GUIDED");
    (void)hipSetDevice(c->device);
    // the pass's scenes (an item's key 0: the current scene), in item order
    std::vector<uint64_t> keys;
    std::vector<uint32_t> item_key(n);
    for (uint32_t i = 0; i < n; i++) {
        const uint64_t k = items[i].scene != 0 ? items[i].scene : c->scene_key;
        const auto it = std::find(keys.begin(), keys.end(), k);
        item_key[i] = (uint32_t)(it - keys.begin());
        if (it == keys.end()) keys.push_back(k);
    }
    std::vector<SceneParts> parts(keys.size());
    for (size_t k = 0; k < keys.size(); k++) {
        if (!scene_parts(c, keys[k], parts[k]))
            return set_error(c, DL_ERR_STATE, keys[k] == c->scene_key ? "render_batch before build_accel"
                                                                     : "no resident scene with this key");
        if (parts[k].lights->empty()) return set_error(c, DL_ERR_INVALID, "There is not emitter in the scene.");
        if (parts[k].hide != parts[0].hide || parts[k].eps != parts[0].eps)
            return set_error(c, DL_ERR_INVALID, "scenes of a shared pass differ in hide_lights or epsilon");
    }
    uint64_t total = 0;
    for (uint32_t i = 0; i < n; i++)
        total += (uint64_t)items[i].camera.width * items[i].camera.height * (uint64_t)items[i].spp;
    if (total > (1ull << 27)) return set_error(c, DL_ERR_INVALID, "pass too large");
    // several scenes: one layout holding them all, a record per scene
    SceneUnion* uni = nullptr;
    if (keys.size() > 1) DL_TRY(scene_union_for(c, keys, &uni));
    const uint32_t P = (uint32_t)total;
    DL_TRY(ensure_state(c, P, 1));
    RenderState* rs = (RenderState*)c->rs;
    if (c->film_slots.size() < (size_t)DL_BATCH_MAX) c->film_slots.resize(DL_BATCH_MAX);
    for (uint32_t i = 0; i < n; i++) {
        Ctx::FilmSlot& f = c->film_slots[items[i].film];
        const uint32_t w = (uint32_t)items[i].camera.width, h = (uint32_t)items[i].camera.height;
        if (items[i].accumulate && (f.w != w || f.h != h))
            return set_error(c, DL_ERR_INVALID, "film slot accumulates an image of another size");
        if (f.cap < w * h) {
            (void)dfree(c, f.film);
            (void)dfree(c, f.rgb);
            f = Ctx::FilmSlot{};
            DL_HIP_CHECK(c, dmalloc(c, (void**)&f.film, (size_t)w * h * sizeof(float4)));
            DL_HIP_CHECK(c, dmalloc(c, (void**)&f.rgb, (size_t)w * h * 3 * sizeof(float)));
            f.cap = w * h;
        }
        f.w = w;
        f.h = h;
    }
    SceneView sv;
    BvhView bvh;
    std::vector<uint32_t> item_rec(n, 0);  // scene record of each item's paths
    if (!uni) {
        const SceneParts& s = parts[0];
        sv = layout_view(*s.ds, (int)s.lights->size(), s.env_light, s.bg);
        bvh = BvhView{s.ds->nodes, s.ds->tris};
        rs->b.roots = nullptr;
    } else {
        sv = layout_view(uni->ds, 0, -1, parts[0].bg);
        std::vector<SceneRec>& recs = uni->host_recs;
        recs.resize(uni->keys.size());
        for (size_t m = 0; m < uni->keys.size(); m++) {
            SceneParts mp;
            if (!scene_parts(c, uni->keys[m], mp)) return set_error(c, DL_ERR_STATE, "resident scene lost");
            SceneRec& r = recs[m];
            r.light0 = uni->light0[m];
            r.num_lights = (int32_t)mp.lights->size();
            r.env_light = mp.env_light >= 0 ? uni->light0[m] + mp.env_light : -1;
            r.root = uni->root[m];
            for (int k = 0; k < 3; k++) r.bg[k] = mp.bg[k];
            r.pad = 0;
        }
        DL_HIP_CHECK(c, hipMemcpyAsync(uni->recs, recs.data(), recs.size() * sizeof(SceneRec), hipMemcpyHostToDevice,
                                       c->stream));
        for (uint32_t i = 0; i < n; i++)
            item_rec[i] = (uint32_t)(std::find(uni->keys.begin(), uni->keys.end(), keys[item_key[i]]) -
                                     uni->keys.begin());
        sv.scenes = uni->recs;
        sv.sid = rs->b.sid;
        bvh = BvhView{uni->ds.nodes, uni->ds.tris};
        rs->b.roots = uni->roots;
    }
    for (int t = 0; t < 8; t++) rs->types_present[t] = false;
    for (const SceneParts& s : parts)
        for (const auto& m : *s.mats) rs->types_present[m.type] = true;
    const bool guide_types[8] = {false, false, false, false, false, false, false, false};
    const bool count = stats && (prm->flags & DL_RENDER_COUNT_TRAVERSAL);
    const bool normals = (prm->flags & DL_RENDER_NORMALS) != 0;
    const bool hide = parts[0].hide;
    const float eps = parts[0].eps;
    const int stage_grid = std::max(1, c->num_cus) * 8;
    const int trace_grid = (int)(count ? c->trace_grid_count : c->trace_grid);
    std::vector<int> tags;
    Timer tm{c, rs, &tags};
    double closest = 0, shadow = 0;
    PassParams pp{};
    pp.P = P;
    pp.W = 1;
    pp.H = 1;
    pp.max_depth = prm->max_depth;
    pp.rr_depth = prm->rr_depth;
    pp.cap = P;
    pp.eps = eps;
    closest += P;
    std::vector<PassParams> ipp(n);
    std::vector<uint32_t> base(n);
    DL_HIP_CHECK(c, hipEventRecord(c->ev0, c->stream));
    k_reset_all<<<1, kStageBlock, 0, c->stream>>>(rs->b.cnt);
    k_reset<<<1, kStageBlock, 0, c->stream>>>(rs->b.cnt, 0);
    k_reset<<<1, kStageBlock, 0, c->stream>>>(rs->b.cnt, 1);
    DL_TRY(tm.start(kTagOther));
    uint32_t off = 0;
    for (uint32_t i = 0; i < n; i++) {
        const dl_batch_item& it = items[i];
        PassParams& q = ipp[i];
        q = pp;
        q.W = (uint32_t)it.camera.width;
        q.H = (uint32_t)it.camera.height;
        q.P = q.W * q.H * (uint32_t)it.spp;
        q.sample_base = (uint32_t)it.first_sample;
        q.seed = it.seed;
        q.scene = item_rec[i];
        base[i] = off;
        if (!it.accumulate)
            k_clear_film<<<stage_grid, kStageBlock, 0, c->stream>>>((float4*)c->film_slots[it.film].film, q.W * q.H);
        k_raygen<<<stage_grid, kStageBlock, 0, c->stream>>>(q, it.camera, rs->b, off, P);
        off += q.P;
    }
    DL_TRY(tm.stop());
    const GuideView gv{};
    const Net net{};
    DL_TRY(run_bounces(c, rs, pp, sv, bvh, count, normals, hide, false, guide_types, gv, net, 1, stage_grid,
                       trace_grid, tm, closest, shadow));
    DL_TRY(tm.start(kTagFilm));
    for (uint32_t i = 0; i < n; i++) {
        const dl_batch_item& it = items[i];
        const PassParams& q = ipp[i];
        RenderBufs fb = rs->b;
        fb.rad += base[i];
        fb.pos += base[i];
        fb.film = (float4*)c->film_slots[it.film].film;
        const dl_film_desc& film = *parts[item_key[i]].film;  // the image's scene's reconstruction filter
        film_gather(q, (uint32_t)it.spp, film, fb, stage_grid, c->stream);
        if (it.rgb_out)
            k_resolve<<<stage_grid, kStageBlock, 0, c->stream>>>((const float4*)c->film_slots[it.film].film,
                                                                  c->film_slots[it.film].rgb, q.W * q.H);
    }
    DL_TRY(tm.stop());
    DL_HIP_CHECK(c, hipGetLastError());
    DL_HIP_CHECK(c, hipEventRecord(c->ev1, c->stream));
    DL_HIP_CHECK(c, hipEventSynchronize(c->ev1));
    for (uint32_t i = 0; i < n; i++)
        if (items[i].rgb_out)
            DL_HIP_CHECK(c, hipMemcpy(items[i].rgb_out, c->film_slots[items[i].film].rgb,
                                      (size_t)ipp[i].W * ipp[i].H * 3 * sizeof(float), hipMemcpyDeviceToHost));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        float ms = 0.f;
        DL_HIP_CHECK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        stats->ms_total = ms;
        for (size_t i = 0; i < tags.size(); i++) {
            float e = 0.f;
            DL_HIP_CHECK(c, hipEventElapsedTime(&e, rs->evs[2 * i], rs->evs[2 * i + 1]));
            switch (tags[i]) {
                case kTagClosest: stats->ms_closest += e; stats->launches_closest += 1; break;
                case kTagShadow: stats->ms_shadow += e; stats->launches_shadow += 1; break;
                case kTagShade: stats->ms_shade += e; break;
                case kTagFilm: stats->ms_film += e; break;
                default: stats->ms_other += e; break;
            }
        }
        DL_HIP_CHECK(c, hipMemcpy(rs->cnt_host, rs->b.cnt, sizeof(Counters), hipMemcpyDeviceToHost));
        stats->paths = P;
        stats->closest_rays = closest;
        stats->shadow_rays = shadow;
        stats->invalid_samples = (double)rs->cnt_host->invalid;
        stats->nodes_closest = (double)rs->cnt_host->nodes;
        stats->tris_closest = (double)rs->cnt_host->tris;
        stats->nodes_shadow = (double)rs->cnt_host->nodes_sh;
        stats->tris_shadow = (double)rs->cnt_host->tris_sh;
    }
    return DL_OK;
}

}  // namespace dl
