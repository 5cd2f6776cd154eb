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
// (kMaxBinaryDepth) always suffices.
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

// 21-bit integer -> 63-bit word with two zero bits after every bit (the
// interleave of a 3-D Morton code), by shift-and-mask halving steps.
__device__ __forceinline__ unsigned long long spread3(uint32_t a) {
    unsigned long long x = a & 0x1fffffull;
    x = (x | (x << 32)) & 0x001f00000000ffffull;
    x = (x | (x << 16)) & 0x001f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

// 63-bit Morton codes (21 bits per axis) of the centroids, with the
// primitive ids as sort values (PLOC: finer than the 30-bit keys, so
// dense meshes keep a spatial order inside a cell of the coarse grid).
__global__ __launch_bounds__(kB) void k_keys63(const float4* __restrict__ blo, const float4* __restrict__ bhi,
                                               uint32_t n, const uint32_t* __restrict__ bounds,
                                               unsigned long long* __restrict__ keys, uint32_t* __restrict__ ids) {
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
            q[a] = (uint32_t)fminf(fmaxf(t * 2097152.0f, 0.0f), 2097151.0f);
        }
        keys[i] = (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
        ids[i] = i;
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
        for (int level = 0; level < 2 * kMaxBinaryDepth; ++level) {
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
        while (node != 0 && d < 2 * kMaxBinaryDepth) {
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
        hipError_t e = dmalloc(c, (void**)&q, std::max<size_t>(1, count) * sizeof(T));
        if (e != hipSuccess) return hip_fail(c, e, "hipMalloc");
        ptrs.push_back(q);
        *p = (T*)q;
        return DL_OK;
    }
    ~Scratch() {
        if (c->stream) (void)hipStreamSynchronize(c->stream);
        for (void* p : ptrs) (void)dfree(c, p);
    }
};

#define DL_TRY(expr)                \
    do {                            \
        dl_status s_ = (expr);      \
        if (s_ != DL_OK) return s_; \
    } while (0)

}  // namespace

// Scenes of at most kMaxLeafSize triangles: the host builder makes the
// single node.
static dl_status build_small(Ctx* c, dl_accel_stats* stats) {
    HostBvh bvh;
    build_bvh_host(c->tri_verts.data(), c->nprims, bvh);
    void *pn = nullptr, *pt = nullptr;
    DL_HIP_CHECK(c, dmalloc(c, &pn, bvh.nodes.size() * sizeof(BvhNodeH)));
    DL_HIP_CHECK(c, dmalloc(c, &pt, std::max<size_t>(1, bvh.tris.size()) * sizeof(BvhTriH)));
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


dl_status build_bvh_lbvh(Ctx* c, dl_accel_stats* stats) {
    const int n = (int)c->nprims;
    if (n <= kMaxLeafSize) return build_small(c, stats);
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
    DL_HIP_CHECK(c, dmalloc(c, &pn, (size_t)(n - 1) * kNodeBytes));
    DL_HIP_CHECK(c, dmalloc(c, &pt, (size_t)n * kTriBytes));
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
    if (hb[6] > (uint32_t)kMaxBinaryDepth) {
        c->bvh_too_deep = true;
        return set_error(c, DL_ERR_STATE, "LBVH deeper than the traversal stack");
    }
    return DL_OK;
}


// ===========================================================================
// PLOC (Meister and Bittner, "Parallel locally-ordered clustering for
// bounding volume hierarchy construction", TVCG 2018): starting from the
// Morton-ordered primitives, every cluster finds its nearest neighbour
// (smallest merged surface area) within kPlocRadius positions, mutual
// nearest neighbours merge, and the survivors are compacted in order, until
// one cluster remains.  Each merge also computes the node's SAH cost
// bottom-up (traversal kSahNode per node, kSahTri per triangle test):
// a subtree of at most kMaxLeafSize primitives becomes one leaf when that
// is cheaper than keeping it as internal nodes.  The triangles are laid out
// in depth-first leaf order, so every leaf's triangles are contiguous.
// ===========================================================================
namespace {

constexpr int kPlocRadius = 24;
// a visit of the 4-wide node that the binary node becomes costs about a
// triangle test, the traversal testing a leaf's triangles two at a time
// (measured on the 2M-triangle workload: 1.0 beats 0.35, 0.5, 0.7, 1.5 and
// 2.0 by 1-3%, with 8.48 node visits and 3.16 triangle tests per ray
// against 8.87 and 2.59 at 0.5)
constexpr float kSahNode = 1.0f, kSahTri = 1.0f;

__device__ __forceinline__ float half_area(const float4& lo, const float4& hi) {
    const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    return dx * dy + dy * dz + dz * dx;
}

// Cluster references: r >= 0 internal node, r < 0 leaf of sorted position ~r.
__global__ __launch_bounds__(kB) void k_ploc_init(int n, const uint32_t* __restrict__ sorted,
                                                  const float4* __restrict__ blo, const float4* __restrict__ bhi,
                                                  int* __restrict__ cref, float4* __restrict__ clo,
                                                  float4* __restrict__ chi) {
    for (int k = blockIdx.x * kB + threadIdx.x; k < n; k += gridDim.x * kB) {
        const uint32_t p = sorted[k];
        cref[k] = ~k;
        clo[k] = blo[p];
        chi[k] = bhi[p];
    }
}

__global__ __launch_bounds__(kB) void k_ploc_nn(int n, const float4* __restrict__ clo, const float4* __restrict__ chi,
                                                int* __restrict__ nn) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
        const float4 a = clo[i], b = chi[i];
        float best = INFINITY;
        int bj = i > 0 ? i - 1 : i + 1;
        const int j0 = max(0, i - kPlocRadius), j1 = min(n - 1, i + kPlocRadius);
        for (int j = j0; j <= j1; ++j) {
            if (j == i) continue;
            const float4 c = clo[j], d = chi[j];
            const float4 lo = make_float4(fminf(a.x, c.x), fminf(a.y, c.y), fminf(a.z, c.z), 0.f);
            const float4 hi = make_float4(fmaxf(b.x, d.x), fmaxf(b.y, d.y), fmaxf(b.z, d.z), 0.f);
            const float ar = half_area(lo, hi);
            if (ar < best) {  // j increases, so ties keep the lowest index
                best = ar;
                bj = j;
            }
        }
        nn[i] = bj;
    }
}

// flags: merge[i] = 1 when i starts a merge (mutual pair, i < partner);
// keep[i] = 0 for the partner that disappears.
__global__ __launch_bounds__(kB) void k_ploc_flags(int n, const int* __restrict__ nn, int* __restrict__ merge,
                                                   int* __restrict__ keep) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
        const int j = nn[i];
        const bool mutual = nn[j] == i;
        merge[i] = (mutual && i < j) ? 1 : 0;
        keep[i] = (mutual && i > j) ? 0 : 1;
    }
}

__global__ __launch_bounds__(kB) void k_ploc_merge(int n, int top, const int* __restrict__ nn,
                                                   const int* __restrict__ merge, const int* __restrict__ midx,
                                                   int* __restrict__ cref, float4* __restrict__ clo,
                                                   float4* __restrict__ chi, int2* __restrict__ child,
                                                   float4* __restrict__ nlo, float4* __restrict__ nhi,
                                                   int* __restrict__ ncount, int* __restrict__ ecount,
                                                   int* __restrict__ parent_int, int* __restrict__ parent_leaf,
                                                   float* __restrict__ ncost, int* __restrict__ nleaf) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
        if (!merge[i]) continue;
        const int j = nn[i];
        const int node = top - midx[i];
        const int ra = cref[i], rb = cref[j];
        const float4 a0 = clo[i], a1 = chi[i], b0 = clo[j], b1 = chi[j];
        const float4 lo = make_float4(fminf(a0.x, b0.x), fminf(a0.y, b0.y), fminf(a0.z, b0.z), 0.f);
        const float4 hi = make_float4(fmaxf(a1.x, b1.x), fmaxf(a1.y, b1.y), fmaxf(a1.z, b1.z), 0.f);
        child[node] = make_int2(ra, rb);
        nlo[node] = lo;
        nhi[node] = hi;
        const int cnt = (ra < 0 ? 1 : ncount[ra]) + (rb < 0 ? 1 : ncount[rb]);
        ncount[node] = cnt;
        // SAH: internal cost from the children's (already final) costs;
        // a small enough subtree becomes one leaf when that is not dearer
        const float area = half_area(lo, hi);
        const float ca = ra < 0 ? kSahTri : ncost[ra], cb = rb < 0 ? kSahTri : ncost[rb];
        const float cint = area > 0.0f ? kSahNode + (half_area(a0, a1) * ca + half_area(b0, b1) * cb) / area
                                       : kSahNode + ca + cb;
        const bool leaf = cnt <= kMaxLeafSize && (float)cnt * kSahTri <= cint;
        ncost[node] = leaf ? (float)cnt * kSahTri : cint;
        nleaf[node] = leaf ? 1 : 0;
        // nodes emitted in this subtree (none below a leaf)
        ecount[node] = leaf ? 0 : 1 + (ra < 0 ? 0 : ecount[ra]) + (rb < 0 ? 0 : ecount[rb]);
        if (ra < 0) parent_leaf[~ra] = node; else parent_int[ra] = node;
        if (rb < 0) parent_leaf[~rb] = node; else parent_int[rb] = node;
        cref[i] = node;
        clo[i] = lo;
        chi[i] = hi;
    }
}

__global__ __launch_bounds__(kB) void k_ploc_compact(int n, const int* __restrict__ keep, const int* __restrict__ kidx,
                                                     const int* __restrict__ cref, const float4* __restrict__ clo,
                                                     const float4* __restrict__ chi, int* __restrict__ oref,
                                                     float4* __restrict__ olo, float4* __restrict__ ohi) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
        if (!keep[i]) continue;
        const int d = kidx[i];
        oref[d] = cref[i];
        olo[d] = clo[i];
        ohi[d] = chi[i];
    }
}

// Depth-first position of every leaf (element k < n, sorted position k) and
// of every internal node (element n + i): the number of primitives to its
// left, summed over the ancestors it descends from as a right child.
// Also the depth-first pre-order rank of every surviving internal node among
// the surviving nodes (its ancestors, plus the surviving nodes of every left
// sibling subtree), which becomes its index in the emitted node array.
__global__ __launch_bounds__(kB) void k_ploc_offsets(int n, const int2* __restrict__ child,
                                                     const int* __restrict__ ncount, const int* __restrict__ ecount,
                                                     const int* __restrict__ parent_int,
                                                     const int* __restrict__ parent_leaf, int* __restrict__ leaf_pos,
                                                     int* __restrict__ node_pos, int* __restrict__ node_rank,
                                                     uint32_t* __restrict__ max_depth) {
    uint32_t dmax = 0;
    for (int e = blockIdx.x * kB + threadIdx.x; e < 2 * n - 1; e += gridDim.x * kB) {
        int ref, node;
        if (e < n) {
            ref = ~e;
            node = parent_leaf[e];
        } else {
            ref = e - n;
            if (ref == 0) {
                node_pos[0] = 0;
                node_rank[0] = 0;
                continue;
            }
            node = parent_int[ref];
        }
        int off = 0, rank = 0;
        uint32_t d = 1;
        for (;;) {
            const int2 ch = child[node];
            if (ch.y == ref && ch.x >= 0) rank += ecount[ch.x];
            if (ch.y == ref) off += ch.x < 0 ? 1 : ncount[ch.x];
            rank += 1;  // the ancestor itself precedes its subtree
            if (node == 0 || d >= 4 * kMaxBinaryDepth) break;
            ref = node;
            node = parent_int[node];
            ++d;
        }
        if (e < n) {
            leaf_pos[e] = off;
            dmax = max(dmax, d);
        } else {
            node_pos[e - n] = off;
            node_rank[e - n] = rank;
        }
    }
    atomicMax(max_depth, dmax);
}

// Node i lies inside a leaf: it or an ancestor was made a leaf (only
// subtrees of at most kMaxLeafSize primitives can be, so the walk is short).
__device__ __forceinline__ bool in_leaf(int i, const int* __restrict__ ncount, const int* __restrict__ nleaf,
                                        const int* __restrict__ parent_int) {
    for (;;) {
        if (nleaf[i]) return true;
        if (i == 0) return false;
        i = parent_int[i];
        if (ncount[i] > kMaxLeafSize) return false;
    }
}

__global__ __launch_bounds__(kB) void k_ploc_emit(int n, const int2* __restrict__ child,
                                                  const float4* __restrict__ nlo, const float4* __restrict__ nhi,
                                                  const int* __restrict__ ncount, const int* __restrict__ nleaf,
                                                  const int* __restrict__ parent_int, const int* __restrict__ leaf_pos,
                                                  const int* __restrict__ node_pos, const int* __restrict__ node_rank,
                                                  const uint32_t* __restrict__ sorted, const float4* __restrict__ blo,
                                                  const float4* __restrict__ bhi, float4* __restrict__ nodes,
                                                  uint32_t* __restrict__ leaves) {
    uint32_t nl = 0;
    for (int i = blockIdx.x * kB + threadIdx.x; i < n - 1; i += gridDim.x * kB) {
        if (in_leaf(i, ncount, nleaf, parent_int)) continue;  // part of a leaf of an ancestor
        const int2 ch = child[i];
        const int refs[2] = {ch.x, ch.y};
        float lo[2][3], hi[2][3];
        int ref[2], cnt[2];
        for (int c = 0; c < 2; c++) {
            const int r = refs[c];
            float4 a, b;
            int count, pos;
            bool leaf;
            if (r < 0) {
                const uint32_t p = sorted[~r];
                a = blo[p];
                b = bhi[p];
                count = 1;
                pos = leaf_pos[~r];
                leaf = true;
            } else {
                a = nlo[r];
                b = nhi[r];
                count = ncount[r];
                pos = node_pos[r];
                leaf = nleaf[r] != 0;
            }
            pad(a, b, lo[c], hi[c]);
            if (leaf) {
                ref[c] = ~pos;
                cnt[c] = count;
                nl++;
            } else {
                ref[c] = node_rank[r];
                cnt[c] = 0;
            }
        }
        float4* np = nodes + 4 * (size_t)node_rank[i];
        np[0] = make_float4(lo[0][0], hi[0][0], lo[1][0], hi[1][0]);
        np[1] = make_float4(lo[0][1], hi[0][1], lo[1][1], hi[1][1]);
        np[2] = make_float4(lo[0][2], hi[0][2], lo[1][2], hi[1][2]);
        reinterpret_cast<int4*>(np)[3] = make_int4(ref[0], ref[1], cnt[0], cnt[1]);
    }
    if (nl) atomicAdd(leaves, nl);
}

// Triangle of sorted position k goes to its depth-first position.
__global__ __launch_bounds__(kB) void k_ploc_gather(int n, const uint32_t* __restrict__ sorted,
                                                    const int* __restrict__ leaf_pos, const float* __restrict__ pos,
                                                    const uint32_t* __restrict__ idx, float4* __restrict__ tris) {
    for (int k = blockIdx.x * kB + threadIdx.x; k < n; k += gridDim.x * kB) {
        const uint32_t p = sorted[k];
        const uint32_t i0 = idx[3 * p], i1 = idx[3 * p + 1], i2 = idx[3 * p + 2];
        const float p0x = pos[3 * i0], p0y = pos[3 * i0 + 1], p0z = pos[3 * i0 + 2];
        const float e1x = pos[3 * i1] - p0x, e1y = pos[3 * i1 + 1] - p0y, e1z = pos[3 * i1 + 2] - p0z;
        const float e2x = pos[3 * i2] - p0x, e2y = pos[3 * i2 + 1] - p0y, e2z = pos[3 * i2 + 2] - p0z;
        float4* t = tris + 3 * (size_t)leaf_pos[k];
        t[0] = make_float4(p0x, p0y, p0z, e1x);
        t[1] = make_float4(e1y, e1z, e2x, e2y);
        t[2] = make_float4(e2z, __uint_as_float(p), 0.f, 0.f);
    }
}

}  // namespace

dl_status build_bvh_ploc(Ctx* c, dl_accel_stats* stats) {
    const int n = (int)c->nprims;
    if (n <= kMaxLeafSize) return build_small(c, stats);
    Scratch sc{c};
    float4 *blo, *bhi, *clo[2], *chi[2], *nlo, *nhi;
    unsigned long long *keys, *keys_sorted;
    uint32_t *bounds, *sorted;
    int *cref[2], *nn, *merge, *keep, *midx, *kidx, *ncount, *ecount, *parent_int, *parent_leaf, *leaf_pos, *node_pos,
        *node_rank, *nleaf;
    float* ncost;
    int2* child;
    DL_TRY(sc.get(&blo, n));
    DL_TRY(sc.get(&bhi, n));
    for (int b = 0; b < 2; b++) {
        DL_TRY(sc.get(&clo[b], n));
        DL_TRY(sc.get(&chi[b], n));
        DL_TRY(sc.get(&cref[b], n));
    }
    DL_TRY(sc.get(&nlo, n));
    DL_TRY(sc.get(&nhi, n));
    DL_TRY(sc.get(&keys, n));
    DL_TRY(sc.get(&keys_sorted, n));
    DL_TRY(sc.get(&bounds, 8));
    DL_TRY(sc.get(&sorted, n));
    DL_TRY(sc.get(&nn, n));
    DL_TRY(sc.get(&merge, n + 1));
    DL_TRY(sc.get(&keep, n + 1));
    DL_TRY(sc.get(&midx, n + 1));
    DL_TRY(sc.get(&kidx, n + 1));
    DL_TRY(sc.get(&ncount, n));
    DL_TRY(sc.get(&ecount, n));
    DL_TRY(sc.get(&parent_int, n));
    DL_TRY(sc.get(&parent_leaf, n));
    DL_TRY(sc.get(&leaf_pos, n));
    DL_TRY(sc.get(&node_pos, n));
    DL_TRY(sc.get(&node_rank, n));
    DL_TRY(sc.get(&child, n));
    DL_TRY(sc.get(&nleaf, n));
    DL_TRY(sc.get(&ncost, n));
    size_t sort_bytes = 0, scan_bytes = 0;
    uint32_t* ids = nullptr;
    DL_TRY(sc.get(&ids, n));
    DL_HIP_CHECK(c, hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, keys, keys_sorted, ids, sorted, n, 0, 63,
                                                       c->stream));
    DL_HIP_CHECK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, merge, midx, n + 1, c->stream));
    void* temp = nullptr;
    DL_TRY(sc.get((char**)&temp, std::max(sort_bytes, scan_bytes)));
    void *pn = nullptr, *pt = nullptr;
    DL_HIP_CHECK(c, dmalloc(c, &pn, (size_t)(n - 1) * kNodeBytes));
    DL_HIP_CHECK(c, dmalloc(c, &pt, (size_t)n * kTriBytes));
    c->ds.nodes = (float4*)pn;
    c->ds.tris = (float4*)pt;
    int* host_cnt = nullptr;
    DL_HIP_CHECK(c, hipHostMalloc((void**)&host_cnt, 4 * sizeof(int), hipHostMallocDefault));
    struct HostFree {
        int* p;
        ~HostFree() { (void)hipHostFree(p); }
    } hf{host_cnt};

    const int grid = std::max(1, c->num_cus) * 8;
    DL_HIP_CHECK(c, hipEventRecord(c->ev0, c->stream));
    k_init_bounds<<<1, 64, 0, c->stream>>>(bounds);
    k_prim_boxes<<<grid, kB, 0, c->stream>>>(c->ds.pos, c->ds.idx, (uint32_t)n, blo, bhi, bounds);
    k_keys63<<<grid, kB, 0, c->stream>>>(blo, bhi, (uint32_t)n, bounds, keys, ids);
    // stable: equal codes keep increasing primitive ids
    DL_HIP_CHECK(c, hipcub::DeviceRadixSort::SortPairs(temp, sort_bytes, keys, keys_sorted, ids, sorted, n, 0, 63,
                                                       c->stream));
    k_ploc_init<<<grid, kB, 0, c->stream>>>(n, sorted, blo, bhi, cref[0], clo[0], chi[0]);
    int cur = 0, count = n, top = n - 2, iters = 0;
    while (count > 1) {
        if (++iters > 4 * n) return set_error(c, DL_ERR_STATE, "PLOC made no progress");
        k_ploc_nn<<<grid, kB, 0, c->stream>>>(count, clo[cur], chi[cur], nn);
        k_ploc_flags<<<grid, kB, 0, c->stream>>>(count, nn, merge, keep);
        DL_HIP_CHECK(c, hipcub::DeviceScan::ExclusiveSum(temp, scan_bytes, merge, midx, count + 1, c->stream));
        DL_HIP_CHECK(c, hipcub::DeviceScan::ExclusiveSum(temp, scan_bytes, keep, kidx, count + 1, c->stream));
        DL_HIP_CHECK(c, hipMemcpyAsync(&host_cnt[0], midx + count, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        DL_HIP_CHECK(c, hipMemcpyAsync(&host_cnt[1], kidx + count, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        k_ploc_merge<<<grid, kB, 0, c->stream>>>(count, top, nn, merge, midx, cref[cur], clo[cur], chi[cur], child, nlo,
                                                 nhi, ncount, ecount, parent_int, parent_leaf, ncost, nleaf);
        k_ploc_compact<<<grid, kB, 0, c->stream>>>(count, keep, kidx, cref[cur], clo[cur], chi[cur], cref[cur ^ 1],
                                                   clo[cur ^ 1], chi[cur ^ 1]);
        DL_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        const int merges = host_cnt[0], kept = host_cnt[1];
        if (merges <= 0 || kept != count - merges) return set_error(c, DL_ERR_STATE, "PLOC merge step inconsistent");
        top -= merges;
        count = kept;
        cur ^= 1;
    }
    if (top != -1) return set_error(c, DL_ERR_STATE, "PLOC node count mismatch");
    k_ploc_offsets<<<grid, kB, 0, c->stream>>>(n, child, ncount, ecount, parent_int, parent_leaf, leaf_pos, node_pos,
                                               node_rank, bounds + 6);
    k_ploc_emit<<<grid, kB, 0, c->stream>>>(n, child, nlo, nhi, ncount, nleaf, parent_int, leaf_pos, node_pos,
                                            node_rank, sorted, blo, bhi, (float4*)pn, bounds + 7);
    k_ploc_gather<<<grid, kB, 0, c->stream>>>(n, sorted, leaf_pos, c->ds.pos, c->ds.idx, (float4*)pt);
    DL_HIP_CHECK(c, hipGetLastError());
    DL_HIP_CHECK(c, hipEventRecord(c->ev1, c->stream));
    DL_HIP_CHECK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    DL_HIP_CHECK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    uint32_t hb[8];
    DL_HIP_CHECK(c, hipMemcpy(hb, bounds, sizeof(hb), hipMemcpyDeviceToHost));
    int emitted = 0;
    DL_HIP_CHECK(c, hipMemcpy(&emitted, ecount, sizeof(int), hipMemcpyDeviceToHost));  // root = node 0
    stats->num_nodes = (uint32_t)emitted;
    stats->num_leaves = hb[7];
    stats->max_depth = hb[6];
    stats->build_ms = ms;
    if (hb[6] > (uint32_t)kMaxBinaryDepth) {
        c->bvh_too_deep = true;
        return set_error(c, DL_ERR_STATE, "PLOC tree deeper than the traversal stack");
    }
    return DL_OK;
}

// ===========================================================================
// Collapse of a binary BVH (any builder) into 4-wide nodes: every 4-wide
// node takes a binary node's two children and opens the internal child of
// largest surface area (lowest slot on ties) until it has four children or
// only leaves.  Top-down, one level per step: the frontier holds binary
// nodes that become 4-wide nodes; their internal children get consecutive
// indices (prefix sum), so the 4-wide nodes are in breadth-first order.
// ===========================================================================
namespace {

struct Cand {
    float lo[3], hi[3];
    int ref, cnt;
};

__device__ __forceinline__ void bin_child(const float4* __restrict__ bn, int b, int k, Cand& c) {
    const float4* np = bn + 4 * (size_t)b;
    const float4 r0 = np[0], r1 = np[1], r2 = np[2];
    const int4 r3 = reinterpret_cast<const int4*>(np)[3];
    c.lo[0] = k ? r0.z : r0.x;
    c.hi[0] = k ? r0.w : r0.y;
    c.lo[1] = k ? r1.z : r1.x;
    c.hi[1] = k ? r1.w : r1.y;
    c.lo[2] = k ? r2.z : r2.x;
    c.hi[2] = k ? r2.w : r2.y;
    c.ref = k ? r3.y : r3.x;
    c.cnt = k ? r3.w : r3.z;
}

__device__ __forceinline__ float cand_area(const Cand& c) {
    const float dx = c.hi[0] - c.lo[0], dy = c.hi[1] - c.lo[1], dz = c.hi[2] - c.lo[2];
    return dx * dy + dy * dz + dz * dx;
}

// Children of the 4-wide node made from binary node b; returns how many.
__device__ __forceinline__ int gather4(const float4* __restrict__ bn, int b, Cand (&c)[4]) {
    bin_child(bn, b, 0, c[0]);
    bin_child(bn, b, 1, c[1]);
    int m = 2;
    while (m < 4) {
        int best = -1;
        float ba = -1.0f;
        for (int k = 0; k < m; ++k)
            if (c[k].ref >= 0) {
                const float a = cand_area(c[k]);
                if (a > ba) {
                    ba = a;
                    best = k;
                }
            }
        if (best < 0) break;
        const int r = c[best].ref;
        bin_child(bn, r, 1, c[m]);
        bin_child(bn, r, 0, c[best]);
        ++m;
    }
    return m;
}

__global__ __launch_bounds__(kB) void k_c4_count(const float4* __restrict__ bn, const int* __restrict__ fbin, int nf,
                                                 int* __restrict__ nint) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < nf; i += gridDim.x * kB) {
        Cand c[4];
        const int m = gather4(bn, fbin[i], c);
        int k = 0;
        for (int j = 0; j < m; ++j) k += c[j].ref >= 0 ? 1 : 0;
        nint[i] = k;
        if (i == 0) nint[nf] = 0;
    }
}

__global__ __launch_bounds__(kB) void k_c4_emit(const float4* __restrict__ bn, const int* __restrict__ fbin,
                                                const int* __restrict__ fout, int nf, const int* __restrict__ off,
                                                int base, float4* __restrict__ n4, int* __restrict__ nbin,
                                                int* __restrict__ nout) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < nf; i += gridDim.x * kB) {
        Cand c[4];
        const int m = gather4(bn, fbin[i], c);
        float q[6][4];
        int ref[4], cnt[4];
        int k = off[i];
        for (int j = 0; j < 4; ++j) {
            if (j < m) {
                for (int a = 0; a < 3; ++a) {
                    q[2 * a][j] = c[j].lo[a];
                    q[2 * a + 1][j] = c[j].hi[a];
                }
                if (c[j].ref >= 0) {
                    nbin[k] = c[j].ref;
                    nout[k] = base + k;
                    ref[j] = base + k;
                    cnt[j] = 0;
                    ++k;
                } else {
                    ref[j] = c[j].ref;
                    cnt[j] = c[j].cnt;
                }
            } else {  // unused slot: an empty leaf
                for (int a = 0; a < 6; ++a) q[a][j] = 0.0f;
                ref[j] = ~0;
                cnt[j] = 0;
            }
        }
        float4* np = n4 + 8 * (size_t)fout[i];
        for (int a = 0; a < 6; ++a) np[a] = make_float4(q[a][0], q[a][1], q[a][2], q[a][3]);
        reinterpret_cast<int4*>(np)[6] = make_int4(ref[0], ref[1], ref[2], ref[3]);
        reinterpret_cast<int4*>(np)[7] = make_int4(cnt[0], cnt[1], cnt[2], cnt[3]);
    }
}

// Quantises 4-wide nodes with float child boxes (8 float4 records: lo.x,
// hi.x, lo.y, hi.y, lo.z, hi.z of children 0..3, refs, counts) into the
// 64-byte traversal layout.  Per axis the origin is the least lo of the
// used children and the step the smallest power of two whose 255 steps
// span them; qlo / qhi are rounded outward and then checked with the
// traversal's own decode (origin + q * step in binary32), stepping down /
// up until the decoded box contains the float box (a wider step when 255
// steps do not reach), so quantisation never shrinks a box.
__device__ float q_decode(float o, uint32_t q, float s) { return o + (float)q * s; }

__global__ __launch_bounds__(kB) void k_quantize4(const float4* __restrict__ in, float4* __restrict__ out, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4* np = in + 8 * (size_t)i;
        float b[6][4];
        for (int a = 0; a < 6; ++a) {
            const float4 v = np[a];
            b[a][0] = v.x;
            b[a][1] = v.y;
            b[a][2] = v.z;
            b[a][3] = v.w;
        }
        const int4 ref = reinterpret_cast<const int4*>(np)[6];
        const int4 cnt = reinterpret_cast<const int4*>(np)[7];
        const int refs[4] = {ref.x, ref.y, ref.z, ref.w}, cnts[4] = {cnt.x, cnt.y, cnt.z, cnt.w};
        bool used[4];
        for (int j = 0; j < 4; ++j) used[j] = !(refs[j] == ~0 && cnts[j] == 0);
        float origin[3];
        uint32_t ebits[3], qlo[3] = {0, 0, 0}, qhi[3] = {0, 0, 0};
        for (int ax = 0; ax < 3; ++ax) {
            float o = INFINITY, m = -INFINITY;
            for (int j = 0; j < 4; ++j)
                if (used[j]) {
                    o = fminf(o, b[2 * ax][j]);
                    m = fmaxf(m, b[2 * ax + 1][j]);
                }
            if (!(o <= m)) o = m = 0.0f;
            origin[ax] = o;
            // smallest exponent whose 255 steps cover m - o
            const double ext = (double)m - (double)o;
            int e = -126;
            if (ext > 0.0) {
                int ex;
                (void)frexp(ext / 255.0, &ex);  // ext / 255 < 2^ex
                e = max(-126, ex);
            }
            uint32_t lo4 = 0, hi4 = 0;
            for (;; ++e) {
                const float s = __uint_as_float((uint32_t)(e + 127) << 23);
                bool fits = true;
                lo4 = hi4 = 0;
                for (int j = 0; j < 4; ++j) {
                    uint32_t ql = 0, qh = 0;
                    if (used[j]) {
                        const float lo = b[2 * ax][j], hi = b[2 * ax + 1][j];
                        double fl = floor(((double)lo - (double)o) / (double)s);
                        double fh = ceil(((double)hi - (double)o) / (double)s);
                        ql = (uint32_t)fmin(fmax(fl, 0.0), 255.0);
                        qh = (uint32_t)fmin(fmax(fh, 0.0), 255.0);
                        while (ql > 0 && q_decode(o, ql, s) > lo) --ql;
                        while (qh < 255 && q_decode(o, qh, s) < hi) ++qh;
                        if (q_decode(o, ql, s) > lo || q_decode(o, qh, s) < hi) fits = false;
                    }
                    lo4 |= ql << (8 * j);
                    hi4 |= qh << (8 * j);
                }
                if (fits || e >= 127) break;
            }
            ebits[ax] = (uint32_t)(e + 127);
            qlo[ax] = lo4;
            qhi[ax] = hi4;
        }
        uint32_t c4 = 0;
        for (int j = 0; j < 4; ++j) c4 |= ((uint32_t)cnts[j] & 0xffu) << (8 * j);
        float4* dst = out + 4 * (size_t)i;
        dst[0] = make_float4(origin[0], origin[1], origin[2], __uint_as_float(ebits[0] | ebits[1] << 8 | ebits[2] << 16));
        reinterpret_cast<uint4*>(dst)[1] = make_uint4(qlo[0], qhi[0], qlo[1], qhi[1]);
        reinterpret_cast<uint4*>(dst)[2] = make_uint4(qlo[2], qhi[2], c4, 0u);
        reinterpret_cast<int4*>(dst)[3] = ref;
    }
}

constexpr size_t kNode4FloatBytes = 128;  // 4-wide node with float child boxes (collapse output)

}  // namespace

// Replaces the context's binary nodes (num_binary internal nodes, root 0)
// with quantised 4-wide nodes; stats get the 4-wide node count and depth.
dl_status collapse_bvh4(Ctx* c, uint32_t num_binary, dl_accel_stats* stats) {
    const int cap = (int)std::max<uint32_t>(1, num_binary);
    Scratch sc{c};
    int *fbin[2], *fout[2], *nint, *off;
    for (int b = 0; b < 2; b++) {
        DL_TRY(sc.get(&fbin[b], cap));
        DL_TRY(sc.get(&fout[b], cap));
    }
    DL_TRY(sc.get(&nint, cap + 1));
    DL_TRY(sc.get(&off, cap + 1));
    size_t scan_bytes = 0;
    DL_HIP_CHECK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, nint, off, cap + 1, c->stream));
    void* temp = nullptr;
    DL_TRY(sc.get((char**)&temp, std::max<size_t>(1, scan_bytes)));
    void* p4 = nullptr;
    DL_HIP_CHECK(c, dmalloc(c, &p4, (size_t)cap * kNode4FloatBytes));
    int* host_cnt = nullptr;
    DL_HIP_CHECK(c, hipHostMalloc((void**)&host_cnt, sizeof(int), hipHostMallocDefault));
    struct HostFree {
        int* p;
        ~HostFree() { (void)hipHostFree(p); }
    } hf{host_cnt};
    const int zero = 0;
    DL_HIP_CHECK(c, hipMemcpyAsync(fbin[0], &zero, sizeof(int), hipMemcpyHostToDevice, c->stream));
    DL_HIP_CHECK(c, hipMemcpyAsync(fout[0], &zero, sizeof(int), hipMemcpyHostToDevice, c->stream));
    const int grid = std::max(1, c->num_cus) * 8;
    int nf = 1, next = 1, cur = 0, levels = 0;
    while (nf > 0) {
        if (++levels > kMaxBinaryDepth || next > cap) {
            dfree(c, p4);
            return set_error(c, DL_ERR_STATE, "4-wide BVH collapse exceeded its bounds");
        }
        k_c4_count<<<grid, kB, 0, c->stream>>>(c->ds.nodes, fbin[cur], nf, nint);
        DL_HIP_CHECK(c, hipcub::DeviceScan::ExclusiveSum(temp, scan_bytes, nint, off, nf + 1, c->stream));
        k_c4_emit<<<grid, kB, 0, c->stream>>>(c->ds.nodes, fbin[cur], fout[cur], nf, off, next, (float4*)p4,
                                              fbin[cur ^ 1], fout[cur ^ 1]);
        DL_HIP_CHECK(c, hipMemcpyAsync(host_cnt, off + nf, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        DL_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        DL_HIP_CHECK(c, hipGetLastError());
        next += host_cnt[0];
        nf = host_cnt[0];
        cur ^= 1;
    }
    dfree(c, c->ds.nodes);
    c->ds.nodes = nullptr;
    void* pq = nullptr;
    if (dmalloc(c, &pq, (size_t)next * kNode4Bytes) != hipSuccess) {
        dfree(c, p4);
        return set_error(c, DL_ERR_OOM, "out of device memory for the 4-wide BVH");
    }
    k_quantize4<<<grid, kB, 0, c->stream>>>((const float4*)p4, (float4*)pq, next);
    DL_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    DL_HIP_CHECK(c, hipGetLastError());
    dfree(c, p4);
    c->ds.nodes = (float4*)pq;
    stats->num_nodes = (uint32_t)next;
    stats->max_depth = (uint32_t)levels;
    return DL_OK;
}

dl_status build_bvh_gpu(Ctx* c, dl_accel_stats* stats, int method) {
    return method == DL_BUILD_GPU_LBVH ? build_bvh_lbvh(c, stats) : build_bvh_ploc(c, stats);
}

}  // namespace dl
