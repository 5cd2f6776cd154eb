// CPU replay of the HIP backend's BVH traversal -- TEST INFRASTRUCTURE ONLY.
//
// Restates, operation for operation, the device traversal of the 4-wide
// node layout (drawlab_amd/csrc/bvh_layout.h, drawlab_amd/csrc/hip/trace.h)
// on the node and triangle arrays exported from a built context, so the
// node visits and triangle tests of every ray -- the per-ray counts behind
// the roofline's algorithmic bytes -- can be recomputed on the CPU and
// compared with the in-kernel counters, together with the hits themselves.
//
// Layout read here (all little-endian 32-bit words):
//   node (16 words): origin x, y, z (floats), step exponent bits (ex | ey << 8
//     | ez << 16); qlo.x, qhi.x, qlo.y, qhi.y; qlo.z, qhi.z, count, 0 (each
//     word: children 0..3 in bytes 0..3); ref[4] (int: >= 0 child node, < 0
//     leaf ~first triangle).  Child plane = origin + q * step in binary32.
//   triangle (12 words): p0.xyz, e1.x | e1.yz, e2.xy | e2.z, id (bits), 0, 0
//   rays: 8 planes of n floats (ox, oy, oz, dx, dy, dz, tmin, tmax)
// Arithmetic: binary32, no contraction (built with -ffp-contract=off),
// IEEE fminf/fmaxf (a NaN operand yields the other one), as on the device.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace {

constexpr float kFarSlack = 1.0f + 1.0f / 65536.0f;

struct F3 {
    float x, y, z;
};
inline F3 sub(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float dot(F3 a, F3 b) {
    float s = 0.0f;
    s = s + a.x * b.x;
    s = s + a.y * b.y;
    s = s + a.z * b.z;
    return s;
}
inline F3 cross(F3 a, F3 b) { return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

inline bool slab1(float lx, float hx, float ly, float hy, float lz, float hz, F3 o, F3 rcp, float tmin, float tmax,
                  float& n) {
    const float a0 = (lx - o.x) * rcp.x, b0 = (hx - o.x) * rcp.x;
    const float a1 = (ly - o.y) * rcp.y, b1 = (hy - o.y) * rcp.y;
    const float a2 = (lz - o.z) * rcp.z, b2 = (hz - o.z) * rcp.z;
    n = std::fmax(std::fmax(std::fmin(a0, b0), std::fmin(a1, b1)), std::fmax(std::fmin(a2, b2), tmin));
    const float f = std::fmin(std::fmin(std::fmax(a0, b0), std::fmax(a1, b1)), std::fmin(std::fmax(a2, b2), tmax));
    return n <= f * kFarSlack;
}

inline void cswap(float& ka, int& ra, float& kb, int& rb) {
    const bool s = kb < ka;
    const float k = s ? kb : ka;
    kb = s ? ka : kb;
    ka = k;
    const int r = s ? rb : ra;
    rb = s ? ra : rb;
    ra = r;
}

inline bool mt_test(const float* tr, F3 o, F3 d, float& u, float& v, float& t) {
    const F3 p0{tr[0], tr[1], tr[2]};
    const F3 e1{tr[3], tr[4], tr[5]};
    const F3 e2{tr[6], tr[7], tr[8]};
    const F3 pvec = cross(d, e2);
    const float det = dot(e1, pvec);
    const float inv = 1.0f / det;
    const F3 tvec = sub(o, p0);
    u = dot(tvec, pvec) * inv;
    const F3 qvec = cross(tvec, e1);
    v = dot(d, qvec) * inv;
    t = dot(e2, qvec) * inv;
    const bool parallel = det > -1e-8f && det < 1e-8f;
    const bool out_u = u < 0.0f || u > 1.0f;
    const bool out_v = v < 0.0f || u + v > 1.0f;
    return !(parallel | out_u | out_v);
}

}  // namespace

extern "C" {

// Closest hit (any == 0: prim, t, u, v) or any hit (any != 0: occ) of n rays,
// with the node visits and triangle tests of each ray.
void dlo_bvh4_replay(const float* nodes, const float* tris, const float* rays, uint32_t n, int any, uint32_t* prim,
                     float* tout, float* uout, float* vout, uint8_t* occ, uint32_t* cnodes, uint32_t* ctris) {
    std::vector<int> stack;
    for (uint32_t i = 0; i < n; i++) {
        const F3 o{rays[i], rays[(size_t)n + i], rays[2 * (size_t)n + i]};
        const F3 d{rays[3 * (size_t)n + i], rays[4 * (size_t)n + i], rays[5 * (size_t)n + i]};
        const float tmin = rays[6 * (size_t)n + i], tmax = rays[7 * (size_t)n + i];
        uint32_t best = 0xffffffffu, nv = 0, nt = 0;
        float best_t = tmax, bu = 0.0f, bv = 0.0f;
        bool hit_any = false;
        const bool nan_ray = std::isnan(o.x) || std::isnan(o.y) || std::isnan(o.z) || std::isnan(d.x) ||
                             std::isnan(d.y) || std::isnan(d.z) || std::isnan(tmin) || std::isnan(tmax);
        if (!nan_ray) {
            const F3 rcp{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
            stack.clear();
            int cur = 0;
            for (;;) {
                const float* np = nodes + 16 * (size_t)cur;
                uint32_t w[16];
                std::memcpy(w, np, sizeof(w));
                int ref[4], cnt[4];
                std::memcpy(ref, np + 12, sizeof(ref));
                for (int c = 0; c < 4; c++) cnt[c] = (int)((w[10] >> (8 * c)) & 0xffu);
                nv++;
                float step[3];
                for (int a = 0; a < 3; a++) {
                    const uint32_t bits = ((w[3] >> (8 * a)) & 0xffu) << 23;
                    std::memcpy(&step[a], &bits, 4);
                }
                auto plane = [&](int axis, uint32_t word, int c) {
                    return np[axis] + (float)((word >> (8 * c)) & 0xffu) * step[axis];
                };
                float nn[4];
                bool h[4];
                for (int c = 0; c < 4; c++)
                    h[c] = slab1(plane(0, w[4], c), plane(0, w[5], c), plane(1, w[6], c), plane(1, w[7], c),
                                 plane(2, w[8], c), plane(2, w[9], c), o, rcp, tmin, best_t, nn[c]);
                // triangles of the hit leaf children, slot order
                hit_any = false;
                for (int c = 0; c < 4 && !(any && hit_any); c++) {
                    if (!(h[c] && ref[c] < 0)) continue;
                    for (int j = 0; j < cnt[c]; j++) {
                        const float* tp = tris + 12 * (size_t)(~ref[c] + j);
                        nt++;
                        float u, v, t;
                        if (mt_test(tp, o, d, u, v, t) && t >= tmin && t <= best_t) {
                            if (any) {
                                hit_any = true;
                                break;
                            }
                            uint32_t id;
                            std::memcpy(&id, tp + 9, 4);
                            if (t < best_t || id < best) {
                                best_t = t;
                                best = id;
                                bu = u;
                                bv = v;
                            }
                        }
                    }
                }
                if (any && hit_any) break;
                const float inf = std::numeric_limits<float>::infinity();
                float k0 = (h[0] && ref[0] >= 0) ? nn[0] : inf, k1 = (h[1] && ref[1] >= 0) ? nn[1] : inf;
                float k2 = (h[2] && ref[2] >= 0) ? nn[2] : inf, k3 = (h[3] && ref[3] >= 0) ? nn[3] : inf;
                int c0 = ref[0], c1 = ref[1], c2 = ref[2], c3 = ref[3];
                cswap(k0, c0, k1, c1);
                cswap(k2, c2, k3, c3);
                cswap(k0, c0, k2, c2);
                cswap(k1, c1, k3, c3);
                cswap(k1, c1, k2, c2);
                if (k0 != inf) {
                    if (k3 != inf) stack.push_back(c3);
                    if (k2 != inf) stack.push_back(c2);
                    if (k1 != inf) stack.push_back(c1);
                    cur = c0;
                } else if (stack.empty()) {
                    break;
                } else {
                    cur = stack.back();
                    stack.pop_back();
                }
            }
        }
        if (any) {
            occ[i] = hit_any ? 1 : 0;
        } else {
            prim[i] = best;
            tout[i] = best_t;
            uout[i] = bu;
            vout[i] = bv;
        }
        if (cnodes) cnodes[i] = nan_ray ? 0 : nv;
        if (ctris) ctris[i] = nan_ray ? 0 : nt;
    }
}

}  // extern "C"
