// Host binned-SAH builder for the child-pair BVH (layout in bvh_layout.h).
//
// Replaces the reference's octree build (src/tracer/accel/octree.cpp:7-51,
// 135-151) as the DL_BUILD_HOST_SAH option; the default device builder is in
// hip/bvh_gpu.hip.  Traversal results do not depend on the tree shape: the
// kernels return the lowest-index closest Moller-Trumbore hit whatever the
// tree, so this builder only affects speed.
#include "bvh_build.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <future>
#include <vector>

namespace dl {

namespace {

struct Ref {
    float lo[3], hi[3], c[3];
    uint32_t id;
};

struct Aabb {
    float lo[3], hi[3];
    void reset() {
        for (int i = 0; i < 3; i++) {
            lo[i] = FLT_MAX;
            hi[i] = -FLT_MAX;
        }
    }
    void grow(const float* l, const float* h) {
        for (int i = 0; i < 3; i++) {
            lo[i] = std::min(lo[i], l[i]);
            hi[i] = std::max(hi[i], h[i]);
        }
    }
    void grow_pt(const float* p) { grow(p, p); }
    float half_area() const {
        float d[3];
        for (int i = 0; i < 3; i++) d[i] = std::max(0.0f, hi[i] - lo[i]);
        return d[0] * d[1] + d[1] * d[2] + d[2] * d[0];
    }
};

struct Child {
    int32_t ref;
    int32_t count;
    Aabb box;
};

constexpr int kBins = 32;

struct Builder {
    int sah_depth = kMaxDepth;
    std::vector<Ref> refs;
    std::vector<BvhNodeH> nodes;
    std::atomic<uint32_t> next_node{0};
    std::atomic<uint32_t> leaves{0};
    std::atomic<uint32_t> max_depth{0};

    static void pad_box(const Aabb& b, float* lo, float* hi) {
        float m = 0.0f;
        for (int i = 0; i < 3; i++) m = std::max(m, std::max(std::fabs(b.lo[i]), std::fabs(b.hi[i])));
        float pad = std::max(m, FLT_MIN) * (1.0f / 65536.0f);
        for (int i = 0; i < 3; i++) {
            lo[i] = b.lo[i] - pad;
            hi[i] = b.hi[i] + pad;
        }
    }

    void write_node(uint32_t idx, const Child& a, const Child& b) {
        BvhNodeH& n = nodes[idx];
        float alo[3], ahi[3], blo[3], bhi[3];
        pad_box(a.box, alo, ahi);
        pad_box(b.box, blo, bhi);
        float* rows[3] = {n.r0, n.r1, n.r2};
        for (int k = 0; k < 3; k++) {
            rows[k][0] = alo[k];
            rows[k][1] = ahi[k];
            rows[k][2] = blo[k];
            rows[k][3] = bhi[k];
        }
        n.r3[0] = a.ref;
        n.r3[1] = b.ref;
        n.r3[2] = a.count;
        n.r3[3] = b.count;
    }

    Child make_leaf(uint32_t b, uint32_t e, const Aabb& box, uint32_t depth) {
        Child c;
        c.ref = ~(int32_t)b;
        c.count = (int32_t)(e - b);
        c.box = box;
        leaves++;
        uint32_t d = max_depth.load();
        while (depth > d && !max_depth.compare_exchange_weak(d, depth)) {
        }
        return c;
    }

    // Binned SAH split of refs[b, e).  Returns the split position or b when
    // no split separates the primitives.
    uint32_t split(uint32_t b, uint32_t e, const Aabb& cb) {
        uint32_t n = e - b;
        float best_cost = FLT_MAX;
        int best_axis = -1, best_bin = 0;
        for (int ax = 0; ax < 3; ax++) {
            float ext = cb.hi[ax] - cb.lo[ax];
            if (!(ext > 0.0f)) continue;
            float k = (float)kBins / ext;
            Aabb bb[kBins];
            uint32_t cnt[kBins] = {0};
            for (int i = 0; i < kBins; i++) bb[i].reset();
            for (uint32_t i = b; i < e; i++) {
                int bin = std::min(kBins - 1, (int)((refs[i].c[ax] - cb.lo[ax]) * k));
                cnt[bin]++;
                bb[bin].grow(refs[i].lo, refs[i].hi);
            }
            float rarea[kBins];
            uint32_t rcnt[kBins];
            Aabb acc;
            acc.reset();
            uint32_t c = 0;
            for (int i = kBins - 1; i > 0; i--) {
                acc.grow(bb[i].lo, bb[i].hi);
                c += cnt[i];
                rarea[i] = acc.half_area();
                rcnt[i] = c;
            }
            acc.reset();
            c = 0;
            for (int i = 0; i < kBins - 1; i++) {
                acc.grow(bb[i].lo, bb[i].hi);
                c += cnt[i];
                if (c == 0 || rcnt[i + 1] == 0) continue;
                float cost = acc.half_area() * (float)c + rarea[i + 1] * (float)rcnt[i + 1];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = ax;
                    best_bin = i;
                }
            }
        }
        if (best_axis < 0) return b;
        float ext = cb.hi[best_axis] - cb.lo[best_axis];
        float k = (float)kBins / ext;
        auto it = std::partition(refs.begin() + b, refs.begin() + e, [&](const Ref& r) {
            int bin = std::min(kBins - 1, (int)((r.c[best_axis] - cb.lo[best_axis]) * k));
            return bin <= best_bin;
        });
        uint32_t mid = (uint32_t)(it - refs.begin());
        if (mid == b || mid == e) return b;
        (void)n;
        return mid;
    }

    Child build(uint32_t b, uint32_t e, uint32_t depth) {
        Aabb box, cb;
        box.reset();
        cb.reset();
        for (uint32_t i = b; i < e; i++) {
            box.grow(refs[i].lo, refs[i].hi);
            cb.grow_pt(refs[i].c);
        }
        uint32_t n = e - b;
        if (n <= (uint32_t)kMaxLeafSize) return make_leaf(b, e, box, depth);
        // Below the SAH depth cap, object-median splits until a leaf fits
        // the 4-wide node's 8-bit triangle count.
        const bool capped = depth + 1 >= (uint32_t)sah_depth;
        if (capped && n <= (uint32_t)kMaxLeafCount) return make_leaf(b, e, box, depth);
        uint32_t mid = capped ? b : split(b, e, cb);
        if (mid == b) {
            // Coincident centroids (or the depth cap): median split, ties in
            // index order.
            int ax = 0;
            for (int i = 1; i < 3; i++)
                if (cb.hi[i] - cb.lo[i] > cb.hi[ax] - cb.lo[ax]) ax = i;
            mid = b + n / 2;
            std::nth_element(refs.begin() + b, refs.begin() + mid, refs.begin() + e,
                             [&](const Ref& x, const Ref& y) { return x.c[ax] < y.c[ax] || (x.c[ax] == y.c[ax] && x.id < y.id); });
        }
        uint32_t idx = next_node++;
        Child l, r;
        if (n > 65536 && depth < 8) {
            auto fut = std::async(std::launch::async, [&] { return build(b, mid, depth + 1); });
            r = build(mid, e, depth + 1);
            l = fut.get();
        } else {
            l = build(b, mid, depth + 1);
            r = build(mid, e, depth + 1);
        }
        write_node(idx, l, r);
        Child c;
        c.ref = (int32_t)idx;
        c.count = 0;
        c.box = box;
        return c;
    }
};

}  // namespace

void build_bvh_host(const float* tri_verts, uint32_t num_tris, HostBvh& out, int sah_depth) {
    Builder bld;
    bld.sah_depth = std::min(sah_depth, kMaxDepth);
    bld.refs.resize(num_tris);
    for (uint32_t i = 0; i < num_tris; i++) {
        const float* p = tri_verts + 9 * (size_t)i;
        Ref& r = bld.refs[i];
        for (int k = 0; k < 3; k++) {
            r.lo[k] = std::min(p[k], std::min(p[3 + k], p[6 + k]));
            r.hi[k] = std::max(p[k], std::max(p[3 + k], p[6 + k]));
            r.c[k] = 0.5f * (r.lo[k] + r.hi[k]);
        }
        r.id = i;
    }
    bld.nodes.resize(std::max<uint32_t>(1, num_tris));
    Child root = bld.build(0, num_tris, 1);
    if (root.ref < 0) {
        // The whole scene fits one leaf: node 0 gets it as its first child
        // and an empty leaf as its second.
        uint32_t idx = bld.next_node++;
        Child empty;
        empty.ref = ~(int32_t)0;
        empty.count = 0;
        empty.box = root.box;
        bld.write_node(idx, root, empty);
    }
    uint32_t nn = bld.next_node.load();
    // Children were allocated after their parent, so the root is node 0.
    out.nodes.assign(bld.nodes.begin(), bld.nodes.begin() + nn);
    out.tris.resize(num_tris);
    for (uint32_t k = 0; k < num_tris; k++) {
        const Ref& r = bld.refs[k];
        const float* p = tri_verts + 9 * (size_t)r.id;
        float e1[3], e2[3];
        for (int j = 0; j < 3; j++) {
            e1[j] = p[3 + j] - p[j];
            e2[j] = p[6 + j] - p[j];
        }
        BvhTriH& t = out.tris[k];
        t.t0[0] = p[0];
        t.t0[1] = p[1];
        t.t0[2] = p[2];
        t.t0[3] = e1[0];
        t.t1[0] = e1[1];
        t.t1[1] = e1[2];
        t.t1[2] = e2[0];
        t.t1[3] = e2[1];
        t.t2[0] = e2[2];
        uint32_t id = r.id;
        std::memcpy(&t.t2[1], &id, 4);
        t.t2[2] = 0.0f;
        t.t2[3] = 0.0f;
    }
    out.num_leaves = bld.leaves.load();
    out.max_depth = bld.max_depth.load();
    // SAH cost estimate (traversal 1, intersection 1 per triangle), relative to the root box.
    out.sah_cost = 0.0;
}

}  // namespace dl
