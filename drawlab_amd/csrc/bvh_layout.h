// HBM layout of the acceleration structure used by the gfx950 traversal
// kernels.  Shared by the host builder and the device code.
//
// The builders produce binary "child-pair" nodes; a device pass then
// collapses them into the 4-wide nodes the traversal kernels read.
//
// Binary node, 64 bytes = 4 x 16-byte records:
//   r0 = (c0.lo.x, c0.hi.x, c1.lo.x, c1.hi.x)
//   r1 = (c0.lo.y, c0.hi.y, c1.lo.y, c1.hi.y)
//   r2 = (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z)
//   r3 = (c0.ref, c1.ref, c0.count, c1.count)  as int32
// ref >= 0 is an interior node index; ref < 0 is a leaf whose triangles are
// tris[~ref .. ~ref + count).  Node 0 is the root.  Child boxes are padded
// outward (relative 2^-16 of their magnitude), so the slab test never
// rejects a box that holds a triangle Moller-Trumbore accepts, including
// rays with zero direction components (no 0 * inf).
//
// 4-wide node (traversal), 64 bytes = 4 x 16-byte records; child boxes on
// a per-node grid of 8-bit steps (2^ex, 2^ey, 2^ez) from the node's
// origin:
//   r0 = (origin.x, origin.y, origin.z, exponent bits) -- exponent bits:
//        ex | ey << 8 | ez << 16, each the biased binary32 exponent of the
//        step, so the step is the float with those exponent bits
//   r1 = (qlo.x, qhi.x, qlo.y, qhi.y), r2 = (qlo.z, qhi.z, count, 0)
//        -- each word holds children 0..3 in bytes 0..3
//   r3 = ref of children 0..3 (int32)
// with ref / count as in the binary node; an unused slot is an empty leaf
// (ref = ~0, count = 0).  A child's box is lo = origin + qlo * step,
// hi = origin + qhi * step in binary32: q * step is exact, the sum rounds
// once, and the quantiser chooses qlo, qhi by that same arithmetic so the
// decoded box contains the child's (padded) float box.
// Node 0 is the root; boxes are the binary nodes' padded boxes.
//
// Triangle: 48 bytes = 3 x 16-byte records, pre-gathered in leaf order:
//   t0 = (p0.x, p0.y, p0.z, e1.x)
//   t1 = (e1.y, e1.z, e2.x, e2.y)
//   t2 = (e2.z, prim_id (bits), 0, 0)
// with e1 = p1 - p0 and e2 = p2 - p0 computed once in binary32, which is
// exactly what the reference computes per test (src/tracer/mesh/mesh.cpp:187).
#pragma once
#include <stdint.h>

namespace dl {

constexpr int kNodeBytes = 64;    // binary node (builders)
constexpr int kNode4Bytes = 64;   // 4-wide node (traversal)
constexpr int kTriBytes = 48;
constexpr int kMaxLeafSize = 4;
// Traversal stack: kLdsStack entries in LDS per lane (8 KB per 128-lane
// block, so LDS does not limit occupancy below the 8 waves/SIMD the VGPR
// budget allows), the rest in a global overflow buffer.  No binary tree is
// deeper than kMaxBinaryDepth internal levels (the host builder: SAH splits
// down to depth kMaxDepth, then object-median splits until a leaf holds at
// most kMaxLeafCount triangles, at most 24 more levels for 2^32 triangles;
// the GPU builders' depth is checked and a deeper tree is rebuilt by the
// host builder), so the 4-wide tree has at most as many levels, and a
// 4-wide visit pushes at most 3 entries: kMaxStack = 3 x kMaxBinaryDepth
// entries always suffice.
constexpr int kLdsStack = 16;
constexpr int kMaxBinaryDepth = 64;
constexpr int kMaxStack = 3 * kMaxBinaryDepth;
constexpr int kMaxDepth = 40;
// Triangles of one leaf: the 4-wide node stores each child's count in 8 bits.
constexpr int kMaxLeafCount = 255;
static_assert(kMaxDepth + 24 <= kMaxBinaryDepth, "median levels below the SAH depth cap");

struct BvhNodeH {
    float r0[4], r1[4], r2[4];
    int32_t r3[4];
};
struct Bvh4NodeH {
    float origin[3];
    uint32_t exps;
    uint8_t qlo_x[4], qhi_x[4], qlo_y[4], qhi_y[4];
    uint8_t qlo_z[4], qhi_z[4], cnt[4], pad[4];
    int32_t ref[4];
};
struct BvhTriH {
    float t0[4], t1[4], t2[4];
};
static_assert(sizeof(BvhNodeH) == kNodeBytes, "node layout");
static_assert(sizeof(Bvh4NodeH) == kNode4Bytes, "4-wide node layout");
static_assert(sizeof(BvhTriH) == kTriBytes, "tri layout");

}  // namespace dl
