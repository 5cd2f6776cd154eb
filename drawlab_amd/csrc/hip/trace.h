// BVH traversal for gfx950: one ray per lane, 4-wide nodes (bvh_layout.h:
// one 64-byte node per visit, four slab tests, the nearest hit internal
// child next and the others stacked farthest first), persistent lanes with
// per-lane ray refill,
// per-lane traversal stack staged in LDS with a lane-interleaved layout
// (entry k of lane l at stack[k * blockDim + l]: every wave touches 64
// consecutive dwords whatever the stack depths, so the accesses are
// bank-conflict free).
//
// Closest hit: the binary32 Moller-Trumbore test of the reference
// (src/tracer/mesh/mesh.cpp:180-220), accepting t in [tmin, t_best] and
// keeping the lowest primitive id on equal t, so the result equals the
// brute-force argmin over all triangles bit-for-bit.  Any hit: the first
// accepted triangle (src/tracer/accel/octree.cpp:112-133).
#pragma once
#include "../bvh_layout.h"
#include "backend.h"
#include "dl_math.h"

namespace dl {

constexpr int kTraceBlock = kTraceBlockHost;
// Slack on the far slab distance: absorbs the rounding of the slab
// arithmetic and of Moller-Trumbore's t (Ize, "Robust BVH ray traversal", 2013).
constexpr float kFarSlack = 1.0f + 1.0f / 65536.0f;

struct BvhView {
    const float4* __restrict__ nodes;
    const float4* __restrict__ tris;
};

struct TraceCounters {
    uint32_t nodes = 0;
    uint32_t tris = 0;
};

// Lane-private traversal stack: entries [0, kLdsStack) in LDS
// (lane-interleaved), [kLdsStack, kMaxStack) in a global overflow buffer
// (entry k of grid thread g at ovf[k * grid_threads + g], coalesced).
struct LaneStack {
    int* __restrict__ lds;
    int lane, stride;
    int* __restrict__ ovf;
    uint32_t ostride;
    // No tree is deeper than kMaxStack, so the last branch is never taken;
    // it keeps every access in bounds regardless.
    __device__ __forceinline__ void push(int& sp, int v) {
        if (sp < kLdsStack)
            lds[sp * stride + lane] = v;
        else if (sp < kMaxStack)
            ovf[(size_t)(sp - kLdsStack) * ostride] = v;
        else
            return;
        ++sp;
    }
    __device__ __forceinline__ int pop(int& sp) {
        --sp;
        return sp < kLdsStack ? lds[sp * stride + lane] : ovf[(size_t)(sp - kLdsStack) * ostride];
    }
};

// Compare-exchange of (key, ref) pairs: smaller key first.
__device__ __forceinline__ void cswap(float& ka, int& ra, float& kb, int& rb) {
    const bool s = kb < ka;
    const float k = s ? kb : ka;
    kb = s ? ka : kb;
    ka = k;
    const int r = s ? rb : ra;
    rb = s ? ra : rb;
    ra = r;
}

// Slab test of one child box (lo, hi per axis): entry distance n and
// whether [n, far] meets [tmin, tmax] (far widened by kFarSlack).
__device__ __forceinline__ bool slab1(float lx, float hx, float ly, float hy, float lz, float hz, const vec3& o,
                                      const vec3& rcp, float tmin, float tmax, float& n) {
    const float a0 = (lx - o.x) * rcp.x, b0 = (hx - o.x) * rcp.x;
    const float a1 = (ly - o.y) * rcp.y, b1 = (hy - o.y) * rcp.y;
    const float a2 = (lz - o.z) * rcp.z, b2 = (hz - o.z) * rcp.z;
    n = fmaxf(fmaxf(fminf(a0, b0), fminf(a1, b1)), fmaxf(fminf(a2, b2), tmin));
    const float f = fminf(fminf(fmaxf(a0, b0), fmaxf(a1, b1)), fminf(fmaxf(a2, b2), tmax));
    return n <= f * kFarSlack;
}

// Slab test of the child whose quantised planes sit at bit `sh` of the
// node's records r1, r2: each plane is origin + q * step.  q * step is
// exact (an 8-bit integer times a power of two), so the fused
// multiply-add rounds once, exactly like the separate product and sum the
// quantiser checked.
__device__ __forceinline__ float qplane(uint32_t word, int sh, float step, float origin) {
    return __builtin_fmaf((float)((word >> sh) & 0xffu), step, origin);
}
__device__ __forceinline__ bool slab_q(const float4& r0, const uint4& r1, const uint4& r2, int sh, float sx, float sy,
                                       float sz, const vec3& o, const vec3& rcp, float tmin, float tmax, float& n) {
    return slab1(qplane(r1.x, sh, sx, r0.x), qplane(r1.y, sh, sx, r0.x), qplane(r1.z, sh, sy, r0.y),
                 qplane(r1.w, sh, sy, r0.y), qplane(r2.x, sh, sz, r0.z), qplane(r2.y, sh, sz, r0.z), o, rcp, tmin,
                 tmax, n);
}

// Moller-Trumbore on a pre-gathered triangle; returns false on a miss.
// Branch-free: every quantity is computed and the reference's early-out
// tests (mesh.cpp:195, 204, 213) are combined into the returned predicate,
// which accepts exactly the rays the early-out sequence accepts.  The triangle
// is therefore fetched by three loads issued together.
__device__ __forceinline__ bool mt_test(const float4& a, const float4& b, const float4& c, const vec3& o,
                                        const vec3& d, float& u, float& v, float& t) {
    const vec3 p0 = v3(a.x, a.y, a.z);
    const vec3 e1 = v3(a.w, b.x, b.y);
    const vec3 e2 = v3(b.z, b.w, c.x);
    const vec3 pvec = cross(d, e2);
    const float det = dot(e1, pvec);
    const float inv = 1.0f / det;
    const vec3 tvec = sub(o, p0);
    u = dot(tvec, pvec) * inv;
    const vec3 qvec = cross(tvec, e1);
    v = dot(d, qvec) * inv;
    t = dot(e2, qvec) * inv;
    const bool parallel = det > -1e-8f && det < 1e-8f;
    const bool out_u = u < 0.0f || u > 1.0f;
    const bool out_v = v < 0.0f || u + v > 1.0f;
    return !(parallel | out_u | out_v);
}

// Adds v of every lane of a converged wave to *dst with one atomic (none
// when the sum is 0: the waves of a short launch mostly traced nothing).
__device__ __forceinline__ void wave_add(unsigned long long* dst, uint32_t v) {
    unsigned long long sum = v;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
    if (__lane_id() == 0 && sum != 0) atomicAdd(dst, sum);
}

// Relative VALU cost of the traversal's two step kinds, triangle step :
// node visit (2 : 5 measured best on the 2M-triangle displaced-sphere
// render with one triangle per step, whose instruction count to a node
// visit's is about 130 : 300; with two triangles per step 1 : 2, 3 : 5 and
// 4 : 5 measure within 0.5% of it).
constexpr uint32_t kTriCost = 2, kNodeCost = 5;

// Persistent traversal with per-lane ray refill.  Every lane of a wave stays
// in the loop until the whole wave has run out of rays; lanes whose rays
// finished start their next rays as soon as enough of them wait (lane g
// takes rays g, g + G, g + 2G, ...: no atomic round trip on the refill
// path), so a wave never idles behind its slowest ray.
//
// A step visits one node, or tests the next one or two of the triangles of
// the leaves the lane's last node visit hit.  A node visit decides the next
// node (the nearest hit internal child, or a stack pop) with the best t of
// that moment, then the lane works through that visit's triangles, two per
// step, before it visits the next node -- the sequence of node visits and
// triangle tests of a ray is the one a node-by-node loop makes (and the CPU
// replay follows), but a wave does not wait at every node for its lane with
// the most leaf triangles: a pair of triangle tests costs one step of the
// lane that needs it, one dependent fetch.  Each iteration the wave takes steps
// of one kind: the lanes whose next step is of that kind take it, the
// others hold theirs.
//
// IO provides: kRefillMin, kRefillMinAny (lanes of a wave that start new
// rays together, for closest-hit and any-hit queries: a refill waits on the
// ray's loads, so a wave whose loads take long refills once that many of its
// lanes are idle, or none is busy, rather than at every finished ray),
// entry(slot) (the ray of work slot `slot`), load(e, o, d,
// tmin, tmax), root(e) (the node traversal starts from), store_hit(e, prim,
// t, u, v, nodes, tris) for closest hits and store_occ(e, occluded, nodes,
// tris) for any-hit queries, e being an entry.
template <bool ANY, bool COUNT, class IO>
__device__ __forceinline__ void traverse_persistent(const BvhView& bvh, uint32_t n, uint32_t* work, LaneStack& stk,
                                                    IO& io, TraceCounters& total) {
    (void)work;  // a claim counter the static-stride refill does not need (callers still zero it)
    const uint32_t G = gridDim.x * blockDim.x;
    uint32_t next = blockIdx.x * blockDim.x + threadIdx.x;
    // the IO entry of the lane's next slot, fetched one refill ahead (a
    // queue's entry is a load the ray's own loads depend on)
    uint32_t pnext = next < n ? io.entry(next) : 0u;
    uint32_t ray = 0;
    bool active = false, exhausted = false;
    vec3 o = v3(0, 0, 0), d = v3(0, 0, 0), rcp = v3(0, 0, 0);
    float tmin = 0.f, best_t = 0.f, bu = 0.f, bv = 0.f;
    uint32_t best = 0xffffffffu;
    int sp = 0, cur = 0;
    // triangles of the last visited node's hit leaves: leaf references, the
    // triangle count of each hit leaf child packed one byte each (0 for the
    // others; a leaf holds at most kMaxLeafCount = 255), their total and the
    // next one to test; `last`: no node left to visit once they are tested
    int4 lref = make_int4(0, 0, 0, 0);
    uint32_t lcnt = 0;
    int ltri = 0, lj = 0;
    bool last = false;
    TraceCounters cnt;
    for (;;) {
        const bool need = !active && !exhausted;
        const uint32_t refill_min = ANY ? IO::kRefillMinAny : IO::kRefillMin;
        const bool refill = (uint32_t)__popcll(__ballot(need)) >= refill_min || !__ballot(active);
        if (need && refill) {
            const uint32_t slot = next;
            next += G;
            {
                if (slot < n) {
                    ray = pnext;
                    if (next < n) pnext = io.entry(next);
                    float tmax;
                    io.load(ray, o, d, tmin, tmax);
                    rcp = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
                    best_t = tmax;
                    best = 0xffffffffu;
                    bu = bv = 0.0f;
                    sp = 0;
                    cur = io.root(ray);
                    lcnt = 0;
                    ltri = lj = 0;
                    last = false;
                    cnt = TraceCounters();
                    // A NaN anywhere in the ray makes every Moller-Trumbore
                    // comparison false (no hit), but would pass every slab
                    // test: finish it at once as a miss.
                    if (isnan(o.x) || isnan(o.y) || isnan(o.z) || isnan(d.x) || isnan(d.y) || isnan(d.z) || isnan(tmin) ||
                        isnan(tmax)) {
                        if (ANY)
                            io.store_occ(ray, false, 0, 0);
                        else
                            io.store_hit(ray, best, best_t, 0.0f, 0.0f, 0, 0);
                    } else {
                        active = true;
                    }
                } else {
                    exhausted = true;
                }
            }
        }
        if (!__ballot(active)) break;
        // One kind of step per wave and iteration: the lanes whose next
        // step is a triangle test and those whose next step is a node visit
        // would otherwise run both branches every iteration under partial
        // masks.  Triangle steps run when enough lanes wait for one (a
        // triangle step costs about kTriCost / kNodeCost of a node visit);
        // the other lanes hold their step, so every ray still makes its own
        // sequence of visits and tests.
        const bool tri_step = lj < ltri;
        const uint32_t ntri = (uint32_t)__popcll(__ballot(active && tri_step));
        const uint32_t nnode = (uint32_t)__popcll(__ballot(active && !tri_step));
        const bool tri_mode = ntri != 0 && ntri * kNodeCost >= nnode * kTriCost;
        if (!active || tri_step != tri_mode) continue;

        bool done = false, hit_any = false;
        if (tri_step) {
            // the next one or two triangles of the last visit's hit leaves, in
            // slot order: both fetched before the first is tested, so a pair
            // costs one dependent wait; the second is accepted against the
            // best t the first left, as a sequential loop would
            const int ls0 = (int)(lcnt & 0xffu);
            const int ls1 = ls0 + (int)((lcnt >> 8) & 0xffu);
            const int ls2 = ls1 + (int)((lcnt >> 16) & 0xffu);
            const int j = lj;
            const bool two = j + 1 < ltri;
            const int k0 = j < ls0 ? ~lref.x + j : j < ls1 ? ~lref.y + (j - ls0) : j < ls2 ? ~lref.z + (j - ls1)
                                                                                     : ~lref.w + (j - ls2);
            const int j1 = two ? j + 1 : j;
            const int k1 = j1 < ls0 ? ~lref.x + j1 : j1 < ls1 ? ~lref.y + (j1 - ls0) : j1 < ls2 ? ~lref.z + (j1 - ls1)
                                                                                          : ~lref.w + (j1 - ls2);
            const float4* tp0 = bvh.tris + 3 * k0;
            const float4* tp1 = bvh.tris + 3 * k1;
            const float4 ta = tp0[0], tb = tp0[1], tc = tp0[2];
            const float4 sa = tp1[0], sb = tp1[1], sc = tp1[2];
            float u, v, t;
            const bool h0 = mt_test(ta, tb, tc, o, d, u, v, t) && t >= tmin && t <= best_t;
            if (ANY) {
                hit_any = h0;
                if (two && !h0) {
                    hit_any = mt_test(sa, sb, sc, o, d, u, v, t) && t >= tmin && t <= best_t;
                    lj++;
                    if (COUNT) cnt.tris++;
                }
            } else {
                if (h0) {
                    const uint32_t id = __float_as_uint(tc.y);
                    if (t < best_t || id < best) {
                        best_t = t;
                        best = id;
                        bu = u;
                        bv = v;
                    }
                }
                if (two) {
                    lj++;
                    if (COUNT) cnt.tris++;
                    if (mt_test(sa, sb, sc, o, d, u, v, t) && t >= tmin && t <= best_t) {
                        const uint32_t id = __float_as_uint(sc.y);
                        if (t < best_t || id < best) {
                            best_t = t;
                            best = id;
                            bu = u;
                            bv = v;
                        }
                    }
                }
            }
            lj++;
            if (COUNT) cnt.tris++;
            done = hit_any || (last && lj == ltri);
        } else {
            const float4* np = bvh.nodes + 4 * (size_t)cur;
            const float4 r0 = np[0];
            const uint4 r1 = reinterpret_cast<const uint4*>(np)[1];
            const uint4 r2 = reinterpret_cast<const uint4*>(np)[2];
            const int4 ref = reinterpret_cast<const int4*>(np)[3];
            if (COUNT) cnt.nodes++;
            // child boxes: origin + q * step (q * step exact, one rounding)
            const uint32_t eb = __float_as_uint(r0.w);
            const float sx = __uint_as_float((eb & 0xffu) << 23), sy = __uint_as_float(((eb >> 8) & 0xffu) << 23),
                        sz = __uint_as_float(((eb >> 16) & 0xffu) << 23);
            float n0, n1, n2, n3;
            const bool h0 = slab_q(r0, r1, r2, 0, sx, sy, sz, o, rcp, tmin, best_t, n0);
            const bool h1 = slab_q(r0, r1, r2, 8, sx, sy, sz, o, rcp, tmin, best_t, n1);
            const bool h2 = slab_q(r0, r1, r2, 16, sx, sy, sz, o, rcp, tmin, best_t, n2);
            const bool h3 = slab_q(r0, r1, r2, 24, sx, sy, sz, o, rcp, tmin, best_t, n3);
            // triangles of the hit leaf children, tested in the next
            // iterations (the closest hit is the minimum over (t, id), so
            // their order does not change it)
            lref = ref;
            lcnt = r2.z & (((h0 && ref.x < 0) ? 0xffu : 0u) | ((h1 && ref.y < 0) ? 0xff00u : 0u) |
                           ((h2 && ref.z < 0) ? 0xff0000u : 0u) | ((h3 && ref.w < 0) ? 0xff000000u : 0u));
            ltri = (int)((lcnt & 0xffu) + ((lcnt >> 8) & 0xffu) + ((lcnt >> 16) & 0xffu) + (lcnt >> 24));
            lj = 0;
            // hit internal children: keys = entry distances (+inf: not
            // taken), sorted by a 5-exchange network; the nearest is visited
            // next and the others are stacked farthest first
            float k0 = (h0 && ref.x >= 0) ? n0 : INFINITY, k1 = (h1 && ref.y >= 0) ? n1 : INFINITY;
            float k2 = (h2 && ref.z >= 0) ? n2 : INFINITY, k3 = (h3 && ref.w >= 0) ? n3 : INFINITY;
            int c0 = ref.x, c1 = ref.y, c2 = ref.z, c3 = ref.w;
            cswap(k0, c0, k1, c1);
            cswap(k2, c2, k3, c3);
            cswap(k0, c0, k2, c2);
            cswap(k1, c1, k3, c3);
            cswap(k1, c1, k2, c2);
            if (k0 != INFINITY) {
                if (k3 != INFINITY) stk.push(sp, c3);
                if (k2 != INFINITY) stk.push(sp, c2);
                if (k1 != INFINITY) stk.push(sp, c1);
                cur = c0;
            } else if (sp == 0) {
                last = true;
            } else {
                cur = stk.pop(sp);
            }
            done = last && ltri == 0;
        }
        if (done) {
            if (ANY)
                io.store_occ(ray, hit_any, cnt.nodes, cnt.tris);
            else
                io.store_hit(ray, best, best_t, bu, bv, cnt.nodes, cnt.tris);  // a miss keeps t = tmax
            if (COUNT) {
                total.nodes += cnt.nodes;
                total.tris += cnt.tris;
            }
            active = false;
        }
    }
}

}  // namespace dl
