// Structural checks of the host BVH builder, run on the CPU before any GPU
// traversal: every child reference is in range, leaves cover each primitive
// exactly once, every triangle lies inside all its ancestors' child boxes,
// no leaf holds more triangles than a 4-wide node can count, and no path is
// deeper than the traversal stack.
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../drawlab_amd/csrc/host/bvh_build.h"

using namespace dl;

extern "C" int bvh_check(const float* tri_verts, uint32_t n, uint32_t* stats, int sah_depth) {
    HostBvh b;
    build_bvh_host(tri_verts, n, b, sah_depth);
    const uint32_t nn = (uint32_t)b.nodes.size();
    std::vector<int> seen(n, 0);
    struct Item { int node; int depth; float lo[3], hi[3]; };
    std::vector<Item> st;
    Item root{0, 1, {-1e30f, -1e30f, -1e30f}, {1e30f, 1e30f, 1e30f}};
    st.push_back(root);
    uint32_t maxd = 0, leaves = 0;
    while (!st.empty()) {
        Item it = st.back();
        st.pop_back();
        if (it.node < 0 || (uint32_t)it.node >= nn) return 1;
        if (it.depth > kMaxBinaryDepth) return 2;
        maxd = it.depth > (int)maxd ? it.depth : maxd;
        const BvhNodeH& nd = b.nodes[it.node];
        for (int c = 0; c < 2; c++) {
            float lo[3] = {nd.r0[2 * c], nd.r1[2 * c], nd.r2[2 * c]};
            float hi[3] = {nd.r0[2 * c + 1], nd.r1[2 * c + 1], nd.r2[2 * c + 1]};
            int ref = nd.r3[c], cnt = nd.r3[2 + c];
            if (ref >= 0) {
                if (cnt != 0) return 3;
                Item ch{ref, it.depth + 1, {lo[0], lo[1], lo[2]}, {hi[0], hi[1], hi[2]}};
                st.push_back(ch);
            } else {
                leaves++;
                uint32_t first = (uint32_t)(~ref);
                if (cnt < 0 || first + (uint32_t)cnt > n) return 4;
                if (cnt > kMaxLeafCount) return 10;
                for (int k = 0; k < cnt; k++) {
                    const BvhTriH& t = b.tris[first + k];
                    uint32_t id;
                    std::memcpy(&id, &t.t2[1], 4);
                    if (id >= n) return 5;
                    seen[id]++;
                    const float* p = tri_verts + 9 * (size_t)id;
                    for (int v = 0; v < 3; v++)
                        for (int a = 0; a < 3; a++) {
                            float x = p[3 * v + a];
                            if (!(x > lo[a] && x < hi[a])) return 6;      // strictly inside the padded box
                            if (!(x >= it.lo[a] && x <= it.hi[a])) return 7;
                        }
                    if (t.t0[0] != p[0] || t.t0[3] != p[3] - p[0] || t.t2[0] != p[8] - p[2]) return 8;
                }
            }
        }
    }
    for (uint32_t i = 0; i < n; i++)
        if (seen[i] != 1) return 9;
    stats[0] = nn;
    stats[1] = leaves;
    stats[2] = maxd;
    return 0;
}
