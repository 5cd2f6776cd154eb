// Device BVH builder (DL_BUILD_GPU): a linear BVH (Karras, "Maximizing
// parallelism in the construction of BVHs, octrees and k-d trees", HPG 2012)
// emitted directly in the child-pair layout of bvh_layout.h.
//
//   1. primitive boxes and the centroid bounds (block reduction + atomics)
//   2. 64-bit keys: 30-bit Morton code of the centroid << 32 | primitive id
//      (unique, so the tree and the traversal order are deterministic)
//   3. radix sort of the keys (hipCUB)
//   4. hierarchy: every internal node finds its key range and split in
//      parallel; children covering <= kMaxLeafSize primitives become leaves
//   5. bottom-up boxes with per-node arrival counters
//   6. node emission (padded child boxes) and triangle pre-gather
//
// Each level of a Karras tree strictly lengthens the common key prefix, so
// over 64-bit keys no internal node is deeper than 64: the traversal stack
// (kMaxStack) always suffices.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>

#include "../host/bvh_build.h"
#include "backend.h"
#include "dl_math.h"

namespace dl {

namespace {

constexpr int kB = 256;

__device__ __forceinline__ uint32_t ordered(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float unordered(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// bounds: [0..2] ordered min of centroids, [3..5] ordered max
__global__ __launch_bounds__(kB) void k_prim_boxes(const float* __restrict__ pos, const uint32_t* __restrict__ idx,
                                                   uint32_t n, float4* __restrict__ blo, float4* __restrict__ bhi,
                                                   uint32_t* __restrict__ bounds) {
    __shared__ uint32_t smin[3][kB], smax[3][kB];
    uint32_t lmin[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, lmax[3] = {0, 0, 0};
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
        float lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            const float x = pos[3 * idx[3 * i] + a], y = pos[3 * idx[3 * i + 1] + a], z = pos[3 * idx[3 * i + 2] + a];
            lo[a] = fminf(x, fminf(y, z));
            hi[a] = fmaxf(x, fmaxf(y, z));
            const uint32_t c = ordered(0.5f * (lo[a] + hi[a]));
            lmin[a] = min(lmin[a], c);
            lmax[a] = max(lmax[a], c);
        }
        blo[i] = make_float4(lo[0], lo[1], lo[2], 0.f);
        bhi[i] = make_float4(hi[0], hi[1], hi[2], 0.f);
    }
    for (int a = 0; a < 3; a++) {
        smin[a][threadIdx.x] = lmin[a];
        smax[a][threadIdx.x] = lmax[a];
    }
    __syncthreads();
    for (int s = kB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int a = 0; a < 3; a++) {
                smin[a][threadIdx.x] = min(smin[a][threadIdx.x], smin[a][threadIdx.x + s]);
                smax[a][threadIdx.x] = max(smax[a][threadIdx.x], smax[a][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int a = 0; a < 3; a++) {
            atomicMin(&bounds[a], smin[a][0]);
            atomicMax(&bounds[3 + a], smax[a][0]);
        }
}

__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__global__ __launch_bounds__(kB) void k_keys(const float4* __restrict__ blo, const float4* __restrict__ bhi, uint32_t n,
                                             const uint32_t* __restrict__ bounds, unsigned long long* __restrict__ keys) {
    float lo[3], ext[3];
    for (int a = 0; a < 3; a++) {
        lo[a] = unordered(bounds[a]);
        ext[a] = unordered(bounds[3 + a]) - lo[a];
    }
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
        const float4 l = blo[i], h = bhi[i];
        const float c[3] = {0.5f * (l.x + h.x), 0.5f * (l.y + h.y), 0.5f * (l.z + h.z)};
        uint32_t q[3];
        for (int a = 0; a < 3; a++) {
            const float t = ext[a] > 0.f ? (c[a] - lo[a]) / ext[a] : 0.5f;
            q[a] = (uint32_t)fminf(fmaxf(t * 1024.0f, 0.0f), 1023.0f);
        }
        const uint32_t m = (expand_bits(q[0]) << 2) | (expand_bits(q[1]) << 1) | expand_bits(q[2]);
        keys[i] = ((unsigned long long)m << 32) | (unsigned long long)i;
    }
}

__device__ __forceinline__ int delta(const unsigned long long* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    return __clzll(k[i] ^ k[j]);
}

// Internal node i: range [first, last], split gamma (left child ends at gamma).
__global__ __launch_bounds__(kB) void k_hierarchy(const unsigned long long* __restrict__ keys, int n,
                                                  int2* __restrict__ range, int* __restrict__ split,
                                                  int* __restrict__ parent_int, int* __restrict__ parent_leaf) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < n - 1; i += gridDim.x * kB) {
        const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
        const int dmin = delta(keys, n, i, i - d);
        int lmax = 2;
        while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
        const int j = i + l * d;
        const int dnode = delta(keys, n, i, j);
        int s = 0;
        int t = l;
        do {
            t = (t + 1) / 2;
            if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        } while (t > 1);
        const int gamma = i + s * d + min(d, 0);
        const int first = min(i, j), last = max(i, j);
        range[i] = make_int2(first, last);
        split[i] = gamma;
        if (first == gamma)
            parent_leaf[gamma] = i;
        else
            parent_int[gamma] = i;
        if (last == gamma + 1)
            parent_leaf[gamma + 1] = i;
        else
            parent_int[gamma + 1] = i;
    }
}

// ilo/ihi are written by other lanes of the same kernel (k_boxes): plain
// pointers, no __restrict__, so loads stay ordered after the acquire fence.
__device__ __forceinline__ void child_box(int first, int last, int node_or_leaf, bool leaf, const float4* blo,
                                          const float4* bhi, const uint32_t* sorted, const float4* ilo,
                                          const float4* ihi, float4& lo, float4& hi) {
    (void)first;
    (void)last;
    if (leaf) {
        const uint32_t p = sorted[node_or_leaf];
        lo = blo[p];
        hi = bhi[p];
    } else {
        lo = ilo[node_or_leaf];
        hi = ihi[node_or_leaf];
    }
}

__global__ __launch_bounds__(kB) void k_boxes(int n, const int2* __restrict__ range, const int* __restrict__ split,
                                              const int* __restrict__ parent_int, const int* __restrict__ parent_leaf,
                                              const uint32_t* __restrict__ sorted, const float4* __restrict__ blo,
                                              const float4* __restrict__ bhi, float4* ilo, float4* ihi,
                                              int* __restrict__ arrive) {
    for (int k = blockIdx.x * kB + threadIdx.x; k < n; k += gridDim.x * kB) {
        int node = parent_leaf[k];
        // A well-formed tree reaches the root within 64 levels; the bound
        // only guarantees termination.
        for (int level = 0; level < 2 * kMaxStack; ++level) {
            __threadfence();
            if (atomicAdd(&arrive[node], 1) == 0) break;
            __threadfence();
            const int2 r = range[node];
            const int g = split[node];
            float4 a0, a1, b0, b1;
            child_box(r.x, g, g, r.x == g, blo, bhi, sorted, ilo, ihi, a0, a1);
            child_box(g + 1, r.y, g + 1, r.y == g + 1, blo, bhi, sorted, ilo, ihi, b0, b1);
            ilo[node] = make_float4(fminf(a0.x, b0.x), fminf(a0.y, b0.y), fminf(a0.z, b0.z), 0.f);
            ihi[node] = make_float4(fmaxf(a1.x, b1.x), fmaxf(a1.y, b1.y), fmaxf(a1.z, b1.z), 0.f);
            if (node == 0) break;
            node = parent_int[node];
        }
    }
}

// Number of internal ancestors of every leaf (an upper bound on the
// traversal depth of the collapsed tree).
__global__ __launch_bounds__(kB) void k_depth(int n, const int* __restrict__ parent_int,
                                              const int* __restrict__ parent_leaf, uint32_t* __restrict__ max_depth) {
    uint32_t dmax = 0;
    for (int k = blockIdx.x * kB + threadIdx.x; k < n; k += gridDim.x * kB) {
        int node = parent_leaf[k];
        uint32_t d = 1;
        while (node != 0 && d < 2 * kMaxStack) {
            node = parent_int[node];
            ++d;
        }
        dmax = max(dmax, d);
    }
    atomicMax(max_depth, dmax);
}

__device__ __forceinline__ void pad(const float4& lo, const float4& hi, float* plo, float* phi) {
    float m = fmaxf(fmaxf(fmaxf(fabsf(lo.x), fabsf(hi.x)), fmaxf(fabsf(lo.y), fabsf(hi.y))),
                    fmaxf(fabsf(lo.z), fabsf(hi.z)));
    const float p = fmaxf(m, 1.17549435e-38f) * (1.0f / 65536.0f);
    plo[0] = lo.x - p;
    plo[1] = lo.y - p;
    plo[2] = lo.z - p;
    phi[0] = hi.x + p;
    phi[1] = hi.y + p;
    phi[2] = hi.z + p;
}

__global__ __launch_bounds__(kB) void k_emit(int n, const int2* __restrict__ range, const int* __restrict__ split,
                                             const uint32_t* __restrict__ sorted, const float4* __restrict__ blo,
                                             const float4* __restrict__ bhi, const float4* __restrict__ ilo,
                                             const float4* __restrict__ ihi, float4* __restrict__ nodes,
                                             uint32_t* __restrict__ leaves) {
    uint32_t nl = 0;
    for (int i = blockIdx.x * kB + threadIdx.x; i < n - 1; i += gridDim.x * kB) {
        const int2 r = range[i];
        const int g = split[i];
        const int cf[2] = {r.x, g + 1}, cl[2] = {g, r.y};
        float lo[2][3], hi[2][3];
        int ref[2], cnt[2];
        for (int c = 0; c < 2; c++) {
            const int size = cl[c] - cf[c] + 1;
            const int id = c == 0 ? g : g + 1;
            float4 a, b;
            child_box(cf[c], cl[c], id, size == 1, blo, bhi, sorted, ilo, ihi, a, b);
            pad(a, b, lo[c], hi[c]);
            if (size <= kMaxLeafSize) {
                ref[c] = ~cf[c];
                cnt[c] = size;
                nl++;
            } else {
                ref[c] = id;
                cnt[c] = 0;
            }
        }
        float4* np = nodes + 4 * (size_t)i;
        np[0] = make_float4(lo[0][0], hi[0][0], lo[1][0], hi[1][0]);
        np[1] = make_float4(lo[0][1], hi[0][1], lo[1][1], hi[1][1]);
        np[2] = make_float4(lo[0][2], hi[0][2], lo[1][2], hi[1][2]);
        reinterpret_cast<int4*>(np)[3] = make_int4(ref[0], ref[1], cnt[0], cnt[1]);
    }
    if (nl) atomicAdd(leaves, nl);
}

__global__ __launch_bounds__(kB) void k_gather(int n, const uint32_t* __restrict__ sorted, const float* __restrict__ pos,
                                               const uint32_t* __restrict__ idx, float4* __restrict__ tris) {
    for (int k = blockIdx.x * kB + threadIdx.x; k < n; k += gridDim.x * kB) {
        const uint32_t p = sorted[k];
        const uint32_t i0 = idx[3 * p], i1 = idx[3 * p + 1], i2 = idx[3 * p + 2];
        const float p0x = pos[3 * i0], p0y = pos[3 * i0 + 1], p0z = pos[3 * i0 + 2];
        const float e1x = pos[3 * i1] - p0x, e1y = pos[3 * i1 + 1] - p0y, e1z = pos[3 * i1 + 2] - p0z;
        const float e2x = pos[3 * i2] - p0x, e2y = pos[3 * i2 + 1] - p0y, e2z = pos[3 * i2 + 2] - p0z;
        float4* t = tris + 3 * (size_t)k;
        t[0] = make_float4(p0x, p0y, p0z, e1x);
        t[1] = make_float4(e1y, e1z, e2x, e2y);
        t[2] = make_float4(e2z, __uint_as_float(p), 0.f, 0.f);
    }
}

__global__ __launch_bounds__(kB) void k_low_ids(const unsigned long long* __restrict__ keys, int n,
                                                uint32_t* __restrict__ sorted) {
    for (int k = blockIdx.x * kB + threadIdx.x; k < n; k += gridDim.x * kB)
        sorted[k] = (uint32_t)(keys[k] & 0xffffffffull);
}

__global__ void k_fill_int(int* p, int v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_init_bounds(uint32_t* b) {
    if (threadIdx.x < 3) {
        b[threadIdx.x] = 0xffffffffu;
        b[3 + threadIdx.x] = 0u;
    }
    if (threadIdx.x == 6) b[6] = 0u;  // max depth
    if (threadIdx.x == 7) b[7] = 0u;  // leaves
}

struct Scratch {
    Ctx* c;
    std::vector<void*> ptrs;
    template <typename T>
    dl_status get(T** p, size_t count) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(1, count) * sizeof(T));
        if (e != hipSuccess) return hip_fail(c, e, "hipMalloc");
        ptrs.push_back(q);
        *p = (T*)q;
        return DL_OK;
    }
    ~Scratch() {
        if (c->stream) (void)hipStreamSynchronize(c->stream);
        for (void* p : ptrs) (void)hipFree(p);
    }
};

#define DL_TRY(expr)                \
    do {                            \
        dl_status s_ = (expr);      \
        if (s_ != DL_OK) return s_; \
    } while (0)

}  // namespace

dl_status build_bvh_gpu(Ctx* c, dl_accel_stats* stats) {
    const int n = (int)c->nprims;
    if (n <= kMaxLeafSize) {
        // Too small for a hierarchy: the host builder makes the single node.
        HostBvh bvh;
        build_bvh_host(c->tri_verts.data(), c->nprims, bvh);
        void *pn = nullptr, *pt = nullptr;
        DL_HIP_CHECK(c, hipMalloc(&pn, bvh.nodes.size() * sizeof(BvhNodeH)));
        DL_HIP_CHECK(c, hipMalloc(&pt, std::max<size_t>(1, bvh.tris.size()) * sizeof(BvhTriH)));
        DL_HIP_CHECK(c, hipMemcpy(pn, bvh.nodes.data(), bvh.nodes.size() * sizeof(BvhNodeH), hipMemcpyHostToDevice));
        if (!bvh.tris.empty())
            DL_HIP_CHECK(c, hipMemcpy(pt, bvh.tris.data(), bvh.tris.size() * sizeof(BvhTriH), hipMemcpyHostToDevice));
        c->ds.nodes = (float4*)pn;
        c->ds.tris = (float4*)pt;
        stats->num_nodes = (uint32_t)bvh.nodes.size();
        stats->num_leaves = bvh.num_leaves;
        stats->max_depth = bvh.max_depth;
        return DL_OK;
    }
    Scratch sc{c};
    float4 *blo, *bhi, *ilo, *ihi;
    unsigned long long *keys, *keys_sorted;
    uint32_t *bounds, *sorted;
    int2* range;
    int *split, *parent_int, *parent_leaf, *arrive;
    DL_TRY(sc.get(&blo, n));
    DL_TRY(sc.get(&bhi, n));
    DL_TRY(sc.get(&ilo, n));
    DL_TRY(sc.get(&ihi, n));
    DL_TRY(sc.get(&keys, n));
    DL_TRY(sc.get(&keys_sorted, n));
    DL_TRY(sc.get(&bounds, 8));
    DL_TRY(sc.get(&sorted, n));
    DL_TRY(sc.get(&range, n));
    DL_TRY(sc.get(&split, n));
    DL_TRY(sc.get(&parent_int, n));
    DL_TRY(sc.get(&parent_leaf, n));
    DL_TRY(sc.get(&arrive, n));
    size_t temp_bytes = 0;
    DL_HIP_CHECK(c, hipcub::DeviceRadixSort::SortKeys(nullptr, temp_bytes, keys, keys_sorted, n, 0, 64, c->stream));
    void* temp = nullptr;
    DL_TRY(sc.get((char**)&temp, temp_bytes));
    void *pn = nullptr, *pt = nullptr;
    DL_HIP_CHECK(c, hipMalloc(&pn, (size_t)(n - 1) * kNodeBytes));
    DL_HIP_CHECK(c, hipMalloc(&pt, (size_t)n * kTriBytes));
    c->ds.nodes = (float4*)pn;
    c->ds.tris = (float4*)pt;

    const int grid = std::max(1, c->num_cus) * 8;
    DL_HIP_CHECK(c, hipEventRecord(c->ev0, c->stream));
    k_init_bounds<<<1, 64, 0, c->stream>>>(bounds);
    k_fill_int<<<grid, kB, 0, c->stream>>>(arrive, 0, (size_t)n);
    k_prim_boxes<<<grid, kB, 0, c->stream>>>(c->ds.pos, c->ds.idx, (uint32_t)n, blo, bhi, bounds);
    k_keys<<<grid, kB, 0, c->stream>>>(blo, bhi, (uint32_t)n, bounds, keys);
    DL_HIP_CHECK(c, hipcub::DeviceRadixSort::SortKeys(temp, temp_bytes, keys, keys_sorted, n, 0, 64, c->stream));
    k_low_ids<<<grid, kB, 0, c->stream>>>(keys_sorted, n, sorted);
    k_hierarchy<<<grid, kB, 0, c->stream>>>(keys_sorted, n, range, split, parent_int, parent_leaf);
    k_boxes<<<grid, kB, 0, c->stream>>>(n, range, split, parent_int, parent_leaf, sorted, blo, bhi, ilo, ihi, arrive);
    k_depth<<<grid, kB, 0, c->stream>>>(n, parent_int, parent_leaf, bounds + 6);
    k_emit<<<grid, kB, 0, c->stream>>>(n, range, split, sorted, blo, bhi, ilo, ihi, (float4*)pn, bounds + 7);
    k_gather<<<grid, kB, 0, c->stream>>>(n, sorted, c->ds.pos, c->ds.idx, (float4*)pt);
    DL_HIP_CHECK(c, hipGetLastError());
    DL_HIP_CHECK(c, hipEventRecord(c->ev1, c->stream));
    DL_HIP_CHECK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    DL_HIP_CHECK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    uint32_t hb[8];
    DL_HIP_CHECK(c, hipMemcpy(hb, bounds, sizeof(hb), hipMemcpyDeviceToHost));
    stats->num_nodes = (uint32_t)(n - 1);
    stats->num_leaves = hb[7];
    stats->max_depth = hb[6];
    stats->build_ms = ms;
    if (hb[6] > (uint32_t)kMaxStack) return set_error(c, DL_ERR_STATE, "LBVH deeper than the traversal stack");
    return DL_OK;
}

}  // namespace dl
