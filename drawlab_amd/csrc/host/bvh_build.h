#pragma once
#include <stdint.h>

#include <vector>

#include "../bvh_layout.h"

namespace dl {

struct HostBvh {
    std::vector<BvhNodeH> nodes;
    std::vector<BvhTriH> tris;
    uint32_t num_leaves = 0;
    uint32_t max_depth = 0;
    double sah_cost = 0.0;
};

// tri_verts: 9 floats per triangle (p0, p1, p2), indexed by global primitive id.
// sah_depth: depth below which nodes are split at the object median instead
// of by SAH (kMaxDepth; smaller values only in the builder's tests).
void build_bvh_host(const float* tri_verts, uint32_t num_tris, HostBvh& out, int sah_depth = kMaxDepth);

}  // namespace dl
