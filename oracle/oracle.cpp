// CPU oracle -- TEST INFRASTRUCTURE ONLY (see omath.h for the rule).
//
// A restatement, in plain C++, of the DrawLab reference CPU path that the
// HIP backend replaces: Moller-Trumbore + octree traversal, the unidirectional
// MIS path integrator, its BSDFs, emitters, camera, pcg32 sampling and the
// filtered-splat film.  Every function cites the reference file:line it
// follows.  Where the reference CPU path is a stub (environment map, the
// anisotropic GGX BSDF) the restatement follows the reference device code,
// evaluated with the CPU path's arithmetic conventions.  The decisions where
// the two reference backends disagree are listed in oracle.h.
//
// How each part is pinned (beyond restating the cited lines):
//   warps, fresnel/refract/reflect  bit-exact against the reference's own
//       wrap.cpp / math.cpp compiled from its sources (oracle/ref.mk)
//   pcg32       published known answers and a big-integer restatement
//   MT, octree  dyadic hand cases; octree == brute force on edge, axis and
//       random ray batches (tests/test_oracle_cpu.py)
//   BSDFs       integral of pdf = valid-sample fraction, mean weight =
//       integral of f cos, chi-square of sampled directions, reciprocity,
//       textbook closed forms of D, G and Fresnel, Snell / TIR / Fresnel
//       frequency (tests/test_oracle_pins_cpu.py)
//   emitters    area pick chi-square, Lambert polygon irradiance, point
//       irradiance, environment direction chi-square, emitter pick
//   Li          white furnaces (closed emitting box, spheres under a
//       constant sky) and a point-lit plane, per path
//   Each of these fails when a mistake is planted (dlo_set_mutation).
#include "oracle.h"
#include "omath.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

using namespace dlo;

namespace {

// Planted mistake (DLO_MUT_*), set only by the tests of the oracle's pins.
int g_mut = DLO_MUT_NONE;
inline bool mut(int m) { return g_mut == m; }

// ---------------------------------------------------------------------------
// Scene containers
// ---------------------------------------------------------------------------
struct Box {
    float mn[3], mx[3];
    void reset() {
        for (int i = 0; i < 3; i++) {
            mn[i] = INFINITY;
            mx[i] = -INFINITY;
        }
    }
    void expand(const float* p) {
        for (int i = 0; i < 3; i++) {
            mn[i] = std::min(mn[i], p[i]);
            mx[i] = std::max(mx[i], p[i]);
        }
    }
    void expand(const Box& b) {
        for (int i = 0; i < 3; i++) {
            mn[i] = std::min(mn[i], b.mn[i]);
            mx[i] = std::max(mx[i], b.mx[i]);
        }
    }
};

struct Mesh {
    std::vector<float> V, N, UV, T;
    std::vector<uint32_t> F;
    int bsdf = -1;
    int emitter = -1;          // index into Scene::emitters, -1 if none
    std::vector<float> cdf;    // DiscretePDF over face areas (dpdf.h:18-178)
    float normalization = 0;   // 1 / total area
    Box bbox;
    uint32_t prim_offset = 0;  // global primitive index of face 0
    uint32_t ntri() const { return (uint32_t)(F.size() / 3); }
    V3 pos(uint32_t f, int k) const {
        uint32_t p = F[3 * f + k];
        return mk(V[3 * p], V[3 * p + 1], V[3 * p + 2]);
    }
    V3 nrm(uint32_t f, int k) const {
        uint32_t p = F[3 * f + k];
        return mk(N[3 * p], N[3 * p + 1], N[3 * p + 2]);
    }
    V3 tan(uint32_t f, int k) const {
        uint32_t p = F[3 * f + k];
        return mk(T[3 * p], T[3 * p + 1], T[3 * p + 2]);
    }
    V2 tex(uint32_t f, int k) const {
        uint32_t p = F[3 * f + k];
        return V2{UV[2 * p], UV[2 * p + 1]};
    }
};

struct Bsdf {
    int type = DLO_BSDF_DIFFUSE;
    float p[16] = {0};
    float ks = 0;  // microfacet: 1 - max(kd)   microfacet.cpp:32
    // texture slots (1 + texture index, 0 = the constant): diffuse albedo;
    // aniso_ggx pd, ps, axay, normal, tangent (diffuse.cpp:69-75,
    // aniso_ggx.cpp:43-58); normal / tangent maps in object space if set
    int tex[5] = {0, 0, 0, 0, 0};
    int object_space = 0;
};

// Bitmap texture as the reference's GPU path samples it (cuda_texture.cpp:
// normalized coordinates, bilinear, wrap); rgb row 0 at v = 0.
struct Tex {
    int w = 0, h = 0;
    std::vector<float> rgb;
};

struct Emitter {
    int type = DLO_EMITTER_AREA;
    int mesh = -1;
    V3 radiance{0, 0, 0};  // area
    V3 position{0, 0, 0};  // point
    V3 intensity{0, 0, 0};
    // environment map (lat-long, row 0 = theta 0 = +y)
    int w = 0, h = 0;
    std::vector<float> rgb;
    float scale = 1.0f;
    bool visual = true;
    std::vector<float> q, pdf;
    std::vector<uint32_t> alias;
};

struct Camera {
    int type = DLO_CAMERA_VIRTUAL;
    int width = 0, height = 0;
    V3 eye, U, V, W;
    V2 inv_size;
    float S[16] = {0}, C[16] = {0};  // perspective: sample->camera, camera->world (row-major); opencv: C
    float near_clip = 1e-4f, far_clip = 1e4f;
    float intr[4] = {0, 0, 1, 1};     // opencv: cx, cy, fx, fy
};

struct Filter {
    float radius = 2.0f;
    int border = 2;
    float table[33];
    float lookup = 16.0f;
};

struct OcNode {
    Box box;
    int child[8];
    std::vector<std::pair<int, int>> tris;
};

struct Scene {
    std::vector<Mesh> meshes;
    std::vector<Bsdf> bsdfs;
    std::vector<Tex> textures;
    std::vector<Emitter> emitters;
    int env = -1;
    Camera cam;
    Filter filter;
    V3 bg{0, 0, 0};
    bool normals = false;      // NormalIntegrator instead of PathIntegrator
    bool hide_lights = false;  // area-light surfaces transparent to path rays
    float eps = kEpsilon;      // ray offset
    std::vector<OcNode> oct;
    int oct_root = -1;
    bool built = false;
    uint32_t nprims = 0;
};

// ---------------------------------------------------------------------------
// Ray and Moller-Trumbore
// ---------------------------------------------------------------------------
struct Ray {
    V3 o, d, rcp;
    float mint, maxt;
};
// TRay(o, d, mint, maxt) computes dRcp = 1 / d     ray.h:41-55, vector.h:78-83
static inline Ray make_ray(V3 o, V3 d, float mint, float maxt) {
    Ray r;
    r.o = o;
    r.d = d;
    r.rcp = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    r.mint = mint;
    r.maxt = maxt;
    return r;
}

// Mesh::rayIntersect  src/tracer/mesh/mesh.cpp:180-220
static bool mt(const Mesh& m, uint32_t f, const Ray& ray, float& u, float& v, float& t) {
    V3 p0 = m.pos(f, 0), p1 = m.pos(f, 1), p2 = m.pos(f, 2);
    V3 e1 = p1 - p0, e2 = p2 - p0;
    V3 pvec = cross(ray.d, e2);
    float det = dot(e1, pvec);
    if (det > -1e-8f && det < 1e-8f) return false;
    float inv = 1.0f / det;
    V3 tvec = ray.o - p0;
    u = dot(tvec, pvec) * inv;
    if (u < 0.0 || u > 1.0) return false;
    V3 qvec = cross(tvec, e1);
    v = dot(ray.d, qvec) * inv;
    if (v < 0.0 || u + v > 1.0) return false;
    t = dot(e2, qvec) * inv;
    return t >= ray.mint && t <= ray.maxt;
}

// Conservative variant of a box for the robust octree traversal: padded by
// 2^-16 of its largest coordinate magnitude.
static Box padded(const Box& b) {
    float m = 0.0f;
    for (int i = 0; i < 3; i++) m = std::max(m, std::max(std::fabs(b.mn[i]), std::fabs(b.mx[i])));
    float pad = std::max(m, 1.17549435e-38f) * (1.0f / 65536.0f);
    Box r;
    for (int i = 0; i < 3; i++) {
        r.mn[i] = b.mn[i] - pad;
        r.mx[i] = b.mx[i] + pad;
    }
    return r;
}

// TBoundingBox::rayIntersect(ray)  include/core/math/bbox.h:258-285
static bool slab(const Box& b, const Ray& ray) {
    float nearT = -INFINITY, farT = INFINITY;
    const float o[3] = {ray.o.x, ray.o.y, ray.o.z};
    const float d[3] = {ray.d.x, ray.d.y, ray.d.z};
    const float r[3] = {ray.rcp.x, ray.rcp.y, ray.rcp.z};
    for (int i = 0; i < 3; i++) {
        if (d[i] == 0) {
            if (o[i] < b.mn[i] || o[i] > b.mx[i]) return false;
        } else {
            float t1 = (b.mn[i] - o[i]) * r[i];
            float t2 = (b.mx[i] - o[i]) * r[i];
            if (t1 > t2) std::swap(t1, t2);
            nearT = std::max(t1, nearT);
            farT = std::min(t2, farT);
            if (!(nearT <= farT)) return false;
        }
    }
    return ray.mint <= farT && nearT <= ray.maxt;
}

// TBoundingBox::rayIntersect(ray, nearT, farT)  bbox.h:288-315.  The partial
// nearT of an early exit is observable: the octree sorts children by it.
static bool slab_interval(const Box& b, const Ray& ray, float& nearT, float& farT) {
    nearT = -INFINITY;
    farT = INFINITY;
    const float o[3] = {ray.o.x, ray.o.y, ray.o.z};
    const float d[3] = {ray.d.x, ray.d.y, ray.d.z};
    const float r[3] = {ray.rcp.x, ray.rcp.y, ray.rcp.z};
    for (int i = 0; i < 3; i++) {
        if (d[i] == 0) {
            if (o[i] < b.mn[i] || o[i] > b.mx[i]) return false;
        } else {
            float t1 = (b.mn[i] - o[i]) * r[i];
            float t2 = (b.mx[i] - o[i]) * r[i];
            if (t1 > t2) std::swap(t1, t2);
            nearT = std::max(t1, nearT);
            farT = std::min(t2, farT);
            if (!(nearT <= farT)) return false;
        }
    }
    return true;
}

static Box tri_box(const Mesh& m, uint32_t f) {
    Box b;
    V3 p0 = m.pos(f, 0), p1 = m.pos(f, 1), p2 = m.pos(f, 2);
    b.mn[0] = b.mx[0] = p0.x;  // BoundingBox3f(p0)  mesh.cpp:116-121
    b.mn[1] = b.mx[1] = p0.y;
    b.mn[2] = b.mx[2] = p0.z;
    float a[3] = {p1.x, p1.y, p1.z}, c[3] = {p2.x, p2.y, p2.z};
    b.expand(a);
    b.expand(c);
    return b;
}

// ---------------------------------------------------------------------------
// Octree  src/tracer/accel/octree.cpp
// ---------------------------------------------------------------------------
static int oc_build(Scene& s, const Box& box, std::vector<std::pair<int, int>>& tris) {
    // recursiveBuild  octree.cpp:7-51
    if (tris.empty()) return -1;
    if (tris.size() < 20) {
        OcNode n;
        n.box = box;
        n.tris = tris;
        s.oct.push_back(std::move(n));
        return (int)s.oct.size() - 1;
    }
    std::vector<std::pair<int, int>> lists[8];
    float center[3];
    for (int k = 0; k < 3; k++) center[k] = (box.mx[k] + box.mn[k]) * 0.5f;  // bbox.h getCenter
    for (int i = 0; i < 8; i++) {
        Box sub;
        for (int k = 0; k < 3; k++) {
            float corner = (i & (1 << k)) ? box.mx[k] : box.mn[k];  // getCorner  bbox.h:250-255
            sub.mn[k] = std::min(corner, center[k]);
            sub.mx[k] = std::max(corner, center[k]);
        }
        for (auto& mt_ : tris) {
            Box tb = tri_box(s.meshes[mt_.first], (uint32_t)mt_.second);
            // tb.overlaps(sub)  bbox.h:120-126
            bool ov = true;
            for (int k = 0; k < 3; k++)
                if (!(sub.mn[k] <= tb.mx[k] && sub.mx[k] >= tb.mn[k])) ov = false;
            if (ov) lists[i].push_back(mt_);
        }
        if (tris.size() == lists[i].size()) {
            OcNode n;
            n.box = box;
            n.tris = tris;
            s.oct.push_back(std::move(n));
            return (int)s.oct.size() - 1;
        }
    }
    int self = (int)s.oct.size();
    s.oct.push_back(OcNode());
    s.oct[self].box = box;
    for (int i = 0; i < 8; i++) {
        Box cb;
        cb.reset();
        for (auto& mt_ : lists[i]) cb.expand(tri_box(s.meshes[mt_.first], (uint32_t)mt_.second));
        for (int k = 0; k < 3; k++) cb.mn[k] -= 0.1f;  // octree.cpp:46
        int c = oc_build(s, cb, lists[i]);
        s.oct[self].child[i] = c;
    }
    return self;
}

struct HitRec {
    int mesh = -1;
    int tri = -1;
    float u = 0, v = 0, t = 0;
};

static inline uint32_t gid(const Scene& s, int m, int f) { return s.meshes[m].prim_offset + (uint32_t)f; }

// recursiveIntersect  octree.cpp:53-110.  lowest=false reproduces the
// reference exactly (an equal later t replaces the hit, children visited in
// stable nearT order); lowest=true keeps the lowest global primitive index
// on equal t.
static bool oc_closest(const Scene& s, int node, Ray& ray, HitRec& hit, bool lowest, bool robust) {
    if (node < 0) return false;
    const OcNode& n = s.oct[node];
    if (robust) {
        // conservative: padded box, far distance relaxed by 2^-16
        Ray r2 = ray;
        r2.maxt = ray.maxt * (1.0f + 1.0f / 65536.0f);
        if (!slab(padded(n.box), r2)) return false;
    } else if (!slab(n.box, ray)) {
        return false;
    }
    if (!n.tris.empty()) {
        bool any = false;
        float u, v, t;
        for (auto& mf : n.tris) {
            if (mt(s.meshes[mf.first], (uint32_t)mf.second, ray, u, v, t)) {
                if (lowest && hit.mesh >= 0 && t == ray.maxt &&
                    gid(s, mf.first, mf.second) > gid(s, hit.mesh, hit.tri))
                    continue;
                ray.maxt = t;
                hit.mesh = mf.first;
                hit.tri = mf.second;
                hit.u = u;
                hit.v = v;
                hit.t = t;
                any = true;
            }
        }
        return any;
    }
    float nearTs[8];
    int index[8];
    for (int i = 0; i < 8; i++) {
        if (n.child[i] < 0)
            nearTs[i] = INFINITY;
        else {
            float nearT, farT;
            slab_interval(s.oct[n.child[i]].box, ray, nearT, farT);
            nearTs[i] = nearT;
        }
        index[i] = i;
    }
    // std::sort on 8 elements is an insertion sort (stable) in libstdc++.
    std::stable_sort(index, index + 8, [&](int a, int b) { return nearTs[a] < nearTs[b]; });
    bool any = false;
    for (int k = 0; k < 8; k++) {
        int idx = index[k];
        if (!robust && ray.maxt < nearTs[idx]) continue;
        if (oc_closest(s, n.child[idx], ray, hit, lowest, robust)) any = true;
    }
    return any;
}

// recursiveAnyhit  octree.cpp:112-133
static bool oc_any(const Scene& s, int node, const Ray& ray, bool robust) {
    if (node < 0) return false;
    const OcNode& n = s.oct[node];
    if (robust) {
        Ray r2 = ray;
        r2.maxt = ray.maxt * (1.0f + 1.0f / 65536.0f);
        if (!slab(padded(n.box), r2)) return false;
    } else if (!slab(n.box, ray)) {
        return false;
    }
    if (!n.tris.empty()) {
        float u, v, t;
        for (auto& mf : n.tris)
            if (mt(s.meshes[mf.first], (uint32_t)mf.second, ray, u, v, t)) return true;
        return false;
    }
    for (int i = 0; i < 8; i++)
        if (oc_any(s, n.child[i], ray, robust)) return true;
    return false;
}

static bool brute_closest(const Scene& s, const Ray& ray, HitRec& hit) {
    bool any = false;
    float u, v, t;
    for (int m = 0; m < (int)s.meshes.size(); m++) {
        const Mesh& me = s.meshes[m];
        for (uint32_t f = 0; f < me.ntri(); f++) {
            if (mt(me, f, ray, u, v, t) && (!any || t < hit.t)) {
                any = true;
                hit.mesh = m;
                hit.tri = (int)f;
                hit.u = u;
                hit.v = v;
                hit.t = t;
            }
        }
    }
    return any;
}

static bool brute_any(const Scene& s, const Ray& ray) {
    float u, v, t;
    for (auto& me : s.meshes)
        for (uint32_t f = 0; f < me.ntri(); f++)
            if (mt(me, f, ray, u, v, t)) return true;
    return false;
}

static bool trace_closest(const Scene& s, int mode, const Ray& ray_in, HitRec& hit) {
    hit = HitRec();
    if (mode == DLO_ACCEL_BRUTE) return brute_closest(s, ray_in, hit);
    Ray ray = ray_in;  // OCTree::rayIntersect copies the ray   octree.cpp:155
    return oc_closest(s, s.oct_root, ray, hit, mode != DLO_ACCEL_OCTREE, mode == DLO_ACCEL_OCTREE_ROBUST);
}
static bool trace_any(const Scene& s, int mode, const Ray& ray) {
    if (mode == DLO_ACCEL_BRUTE) return brute_any(s, ray);
    return oc_any(s, s.oct_root, ray, mode == DLO_ACCEL_OCTREE_ROBUST);
}

// ---------------------------------------------------------------------------
// Intersection record  octree.cpp:166-209 (+ tangent frame for aniso_ggx,
// src/optix/cuda/shader_common.h:121-133, include/optix/math/onb.h:11-33)
// ---------------------------------------------------------------------------
struct Its {
    V3 p;
    float t;
    V2 uv;
    V2 tc;  // texture coordinates as the GPU path interpolates them (0 without)
    Frame sh, geo;
    int mesh;
    int tri;
    V3 bary;
};

// ---------------------------------------------------------------------------
// Bitmap textures, sampled as on the reference GPU path
// (src/optix/host/cuda_texture.cpp:57-82, src/optix/cuda/shader_common.h)
// ---------------------------------------------------------------------------
static int wrap_texel(float f, int n) {
    float r = std::fmod(f, (float)n);
    if (r < 0.0f) r += (float)n;
    int i = (int)r;
    return i < n ? i : n - 1;
}

static V3 tex_sample(const Tex& t, float u, float v) {
    if (!std::isfinite(u) || !std::isfinite(v)) return mk(0, 0, 0);
    const int w = t.w, h = t.h;
    const float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
    const float fx = std::floor(x), fy = std::floor(y);
    const float ax = x - fx, ay = y - fy;
    const int x0 = wrap_texel(fx, w), y0 = wrap_texel(fy, h);
    const int x1 = x0 + 1 < w ? x0 + 1 : 0, y1 = y0 + 1 < h ? y0 + 1 : 0;
    const float* T = t.rgb.data();
    const float bx = 1.0f - ax, by = 1.0f - ay;
    float out[3];
    for (int c = 0; c < 3; c++) {
        const float t00 = T[3 * ((size_t)y0 * w + x0) + c], t10 = T[3 * ((size_t)y0 * w + x1) + c];
        const float t01 = T[3 * ((size_t)y1 * w + x0) + c], t11 = T[3 * ((size_t)y1 * w + x1) + c];
        out[c] = (t00 * bx + t10 * ax) * by + (t01 * bx + t11 * ax) * ay;
    }
    return mk(out[0], out[1], out[2]);
}

static bool textured(const Bsdf& b) { return (b.tex[0] | b.tex[1] | b.tex[2] | b.tex[3] | b.tex[4]) != 0; }

// BSDF with its textured parameters replaced by their values at texture
// coordinates tc (diffuse.cu:16-19; aniso_ggx.cu:17-36).
static Bsdf resolve_bsdf(const Scene& s, const Bsdf& b, V2 tc) {
    Bsdf r = b;
    auto put = [&](int slot, int first, int n) {
        if (!b.tex[slot]) return;
        const V3 a = tex_sample(s.textures[b.tex[slot] - 1], tc.x, tc.y);
        const float v[3] = {a.x, a.y, a.z};
        for (int k = 0; k < n; k++) r.p[first + k] = v[k];
    };
    if (b.type == DLO_BSDF_DIFFUSE) put(0, 0, 3);
    if (b.type == DLO_BSDF_ANISO_GGX) {
        put(0, 0, 3);
        put(1, 3, 3);
        put(2, 6, 2);
    }
    return r;
}

// Onb(normal, tangent)  include/optix/math/onb.h:11-33
static Frame onb_frame(V3 n, V3 t) {
    Frame f;
    f.n = n;
    if (t.x == 0.0f && t.y == 0.0f && t.z == 0.0f) {
        V3 b = std::fabs(n.x) > std::fabs(n.z) ? mk(-n.y, n.x, 0.0f) : mk(0.0f, -n.z, n.y);
        f.b = normalized(b);
        f.t = cross(f.b, n);
    } else {
        f.t = t;
        f.b = normalized(cross(n, t));
    }
    return f;
}

static V3 onb_to_world(const Frame& f, V3 p) { return (f.t * p.x + f.b * p.y) + f.n * p.z; }

// aniso_ggx shading frame with normal / tangent maps (shader_common.h:46-147)
static Frame map_frame(const Scene& s, const Mesh& m, uint32_t f, V3 bary, const Bsdf& b, V2 tc) {
    const float b0 = bary.x, u = bary.y, v = bary.z;
    const V3 p0 = m.pos(f, 0), p1 = m.pos(f, 1), p2 = m.pos(f, 2);
    const V3 gn = normalized(cross(p1 - p0, p2 - p0));
    const bool frame_ok = !m.T.empty() && !m.N.empty();
    Frame tbn{};
    if (frame_ok) {
        const V3 n = normalized((m.nrm(f, 0) * b0 + m.nrm(f, 1) * u) + m.nrm(f, 2) * v);
        const V3 t = normalized((m.tan(f, 0) * b0 + m.tan(f, 1) * u) + m.tan(f, 2) * v);
        tbn = onb_frame(n, t);
    }
    V3 sn = gn;
    if (b.tex[3]) {
        const V3 c = tex_sample(s.textures[b.tex[3] - 1], tc.x, tc.y);
        const V3 l = mk(c.x * 2.0f - 1.0f, c.y * 2.0f - 1.0f, c.z * 2.0f - 1.0f);
        if (b.object_space)
            sn = l;
        else if (frame_ok)
            sn = onb_to_world(tbn, l);
    } else if (!m.N.empty()) {
        sn = (m.nrm(f, 0) * b0 + m.nrm(f, 1) * u) + m.nrm(f, 2) * v;
    }
    sn = normalized(sn);
    V3 st = mk(0, 0, 0);
    if (b.tex[4]) {
        const V3 c = tex_sample(s.textures[b.tex[4] - 1], tc.x, tc.y);
        const V3 l = mk(c.x * 2.0f - 1.0f, c.y * 2.0f - 1.0f, c.z * 2.0f - 1.0f);
        if (b.object_space)
            st = l;
        else if (frame_ok)
            st = onb_to_world(tbn, l);
        if (st.x != 0.0f || st.y != 0.0f || st.z != 0.0f) st = normalized(st);
    } else if (!m.T.empty()) {
        const V3 t0 = m.tan(f, 0);
        if (!(std::isnan(t0.x) || std::isnan(t0.y) || std::isnan(t0.z)))
            st = normalized((t0 * b0 + m.tan(f, 1) * u) + m.tan(f, 2) * v);
    }
    return onb_frame(sn, st);
}

static void make_its(const Scene& s, const HitRec& h, Its& its) {
    const Mesh& m = s.meshes[h.mesh];
    uint32_t f = (uint32_t)h.tri;
    its.mesh = h.mesh;
    its.tri = h.tri;
    its.t = h.t;
    its.uv = V2{h.u, h.v};
    V3 bary = mk((1.0f - h.u) - h.v, h.u, h.v);
    its.bary = bary;
    V3 p0 = m.pos(f, 0), p1 = m.pos(f, 1), p2 = m.pos(f, 2);
    its.p = (p0 * bary.x + p1 * bary.y) + p2 * bary.z;
    if (!m.UV.empty()) {
        V2 t0 = m.tex(f, 0), t1 = m.tex(f, 1), t2 = m.tex(f, 2);
        its.uv = V2{(t0.x * bary.x + t1.x * bary.y) + t2.x * bary.z,
                    (t0.y * bary.x + t1.y * bary.y) + t2.y * bary.z};
    }
    its.geo = frame_from_normal(normalized(cross(p1 - p0, p2 - p0)));
    if (!m.N.empty()) {
        // Only the third vertex normal is normalised, and the sum is not:
        // `bary.z() * n2.normalized()` binds before the additions.
        V3 n = (m.nrm(f, 0) * bary.x + m.nrm(f, 1) * bary.y) + normalized(m.nrm(f, 2)) * bary.z;
        its.sh = frame_from_normal(n);
    } else {
        its.sh = its.geo;
    }
    const Bsdf& b = s.bsdfs[m.bsdf];
    if (b.type == DLO_BSDF_ANISO_GGX && !m.T.empty()) {
        V3 t0 = m.tan(f, 0);
        if (!(std::isnan(t0.x) || std::isnan(t0.y) || std::isnan(t0.z))) {
            V3 st = normalized((t0 * bary.x + m.tan(f, 1) * bary.y) + m.tan(f, 2) * bary.z);
            V3 n = its.sh.n;
            its.sh.t = st;
            its.sh.b = normalized(cross(n, st));
        }
    }
    its.tc = V2{0.0f, 0.0f};
    if (textured(b)) {
        if (!m.UV.empty()) {
            V2 t0 = m.tex(f, 0), t1 = m.tex(f, 1), t2 = m.tex(f, 2);
            its.tc = V2{(bary.x * t0.x + bary.y * t1.x) + bary.z * t2.x, (bary.x * t0.y + bary.y * t1.y) + bary.z * t2.y};
        }
        if (b.type == DLO_BSDF_ANISO_GGX && (b.tex[3] | b.tex[4]) != 0) its.sh = map_frame(s, m, f, bary, b, its.tc);
    }
}

// ---------------------------------------------------------------------------
// Warps  src/core/math/wrap.cpp
// ---------------------------------------------------------------------------
// squareToUniformDisk  wrap.cpp:33-37
static V2 sq_disk(V2 s) {
    float r = dsqrt(s.x);
    float theta = 2 * kPi * s.y;
    return V2{(float)((double)r * dcos(theta)), (float)((double)r * dsin(theta))};
}
// squareToCosineHemisphere  wrap.cpp:66-70
static V3 sq_cos_hemi(V2 s) {
    V2 p = sq_disk(s);
    float sum = 0.0f;
    sum += p.x * p.x;
    sum += p.y * p.y;
    float z = dsqrt(1 - sum);
    return mk(p.x, p.y, z);
}
static float sq_cos_hemi_pdf(V3 v) { return v.z <= 0 ? 0.0f : v.z * kInvPi; }  // wrap.cpp:72-74
// squareToBeckmann  wrap.cpp:77-81
static V3 sq_beckmann(V2 s, float alpha) {
    float theta = (float)std::atan(std::sqrt((double)(-alpha * alpha) * std::log((double)(1 - s.x))));
    float phi = 2 * kPi * s.y;
    double st = dsin(theta);
    return mk((float)(st * dcos(phi)), (float)(st * dsin(phi)), (float)dcos(theta));
}
// squareToBeckmannPdf  wrap.cpp:83-91
static float sq_beckmann_pdf(V3 m, float alpha) {
    if (m.z <= 0) return 0;
    float c = m.z;
    float tan2 = (1 - c * c) / (c * c);
    double e = std::exp((double)(-tan2 / (alpha * alpha)));
    float den = alpha * alpha * c * c * c;
    return (float)((double)kInvTwoPi * (2 * e) / (double)den);
}
// squareToUniformTriangle  wrap.cpp:94-97
static V3 sq_triangle(V2 s) {
    float t = dsqrt(1.f - s.x);
    return mk(1 - t, t * s.y, t - t * s.y);
}

// fresnel / refract / reflect  src/core/math/math.cpp:8-51
static float fresnel(float cosThetaI, float extIOR, float intIOR) {
    float etaI = extIOR, etaT = intIOR;
    if (extIOR == intIOR) return 0.0f;
    if (cosThetaI < 0.0f) {
        std::swap(etaI, etaT);
        cosThetaI = -cosThetaI;
    }
    float eta = etaI / etaT, sinThetaTSqr = eta * eta * (1 - cosThetaI * cosThetaI);
    if (sinThetaTSqr > 1.0f) return 1.0f;
    float cosThetaT = dsqrt(1.0f - sinThetaTSqr);
    float Rs = (etaI * cosThetaI - etaT * cosThetaT) / (etaI * cosThetaI + etaT * cosThetaT);
    float Rp = (etaT * cosThetaI - etaI * cosThetaT) / (etaT * cosThetaI + etaI * cosThetaT);
    return (Rs * Rs + Rp * Rp) / 2.0f;
}
static V3 refract(V3 wi, float eta) {
    float c = wi.z;
    bool outside = c > 0.f;
    eta = outside ? eta : 1 / eta;
    float ct2 = 1 - eta * eta * (1 - c * c);
    float ct = dsqrt(ct2);
    ct = outside ? -ct : ct;
    return mk(-eta * wi.x, -eta * wi.y, ct);
}
static V3 reflect(V3 wi, V3 n) { return n * (2 * dot(wi, n)) - wi; }

// ---------------------------------------------------------------------------
// BSDFs
// ---------------------------------------------------------------------------
enum Measure { kUnknown = 0, kSolidAngle, kDiscrete };
struct BRec {
    V3 wi, wo;
    float eta = 1.f;
    Measure measure = kUnknown;
};

static inline V3 col(const float* p) { return mk(p[0], p[1], p[2]); }

// Beckmann G1 rational approximation  microfacet.cpp:35-44 (binary64 tail)
static float mf_g1(V3 wv, V3 wh, float alpha) {
    float c = dot(wv, wh) / wv.z;
    if (c <= 0) return 0.0f;
    float tmp = 1 - wv.z * wv.z;
    float tanT = tmp <= 0.0f ? 0.0f : dsqrt(tmp) / wv.z;  // Frame::tanTheta frame.h:65-70
    float b = (float)(1.0 / (double)(alpha * tanT));
    double bd = b;
    return bd < 1.6 ? (float)((3.535 * bd + 2.181 * bd * bd) / (1 + 2.276 * bd + 2.577 * bd * bd)) : 1.0f;
}

// Anisotropic GGX, written from src/optix/cuda/material/aniso_ggx.cu:9-121
// and include/optix/math/wrap.h:54-77 with the CPU arithmetic conventions.
static float ggx_d(V3 m, float ax, float ay) {
    if (m.z <= 0) return 0;
    V3 vh = mk(m.x / ax, m.y / ay, m.z);
    float len2 = dot(vh, vh);
    return 1.f / (kPi * ax * ay * len2 * len2);
}
// Wrap::squareToGGX  wrap.h:70-77: x = ax * sqrt(z1) / sqrt(1 - z1) * cos(2 pi z2)
static V3 ggx_sample(V2 s, float ax, float ay) {
    double sz = std::sqrt((double)s.x), sw = std::sqrt((double)(1 - s.x));
    float ang = 2 * kPi * s.y;
    float x = (float)((double)ax * sz / sw * dcos(ang));
    float y = (float)((double)ay * sz / sw * dsin(ang));
    return normalized(mk(-x, -y, 1));
}
static float ggx_g1(V3 v, float ax, float ay) {
    if (v.z <= 0.0) return 0.0f;
    V3 vv = mk(v.x * ax, v.y * ay, v.z);
    return (float)(2.0 * (double)v.z / (double)(v.z + length(vv)));
}
static float schlick(float ldoth) {
    const float F0 = 0.04f;
    double tmp = 1.0 - (double)ldoth;
    float temp = std::min(std::max(0.0f, (float)tmp), 1.0f);
    double t = temp;
    return (float)((double)F0 + (1.0 - (double)F0) * t * t * t * t * t);
}

// Conductor Fresnel for a complex IOR eta + i k (unpolarised, exact).
static float fresnel_conductor(float cosi, float eta, float k) {
    float c2 = cosi * cosi;
    float s2 = 1.0f - c2;
    float e2 = eta * eta, k2 = k * k;
    float t0 = e2 - k2 - s2;
    float a2b2 = dsqrt(t0 * t0 + 4.0f * e2 * k2);
    float t1 = a2b2 + c2;
    float a = dsqrt(0.5f * (a2b2 + t0));
    float t2 = 2.0f * cosi * a;
    float rs = (t1 - t2) / (t1 + t2);
    float t3 = c2 * a2b2 + s2 * s2;
    float t4 = t2 * s2;
    float rp = rs * (t3 - t4) / (t3 + t4);
    return 0.5f * (rp + rs);
}

static bool bsdf_is_diffuse(const Bsdf& b) {
    // diffuse.cpp:78, microfacet.cpp:106-111, aniso_ggx.cpp:61; mirror and
    // dielectric use the BSDF default (false)  bsdf.h:104
    return b.type == DLO_BSDF_DIFFUSE || b.type == DLO_BSDF_MICROFACET ||
           b.type == DLO_BSDF_ANISO_GGX || b.type == DLO_BSDF_CONDUCTOR;
}

static V3 bsdf_eval(const Bsdf& b, const BRec& r) {
    switch (b.type) {
        case DLO_BSDF_DIFFUSE:  // diffuse.cpp:22-31
            if (r.measure != kSolidAngle || r.wi.z <= 0 || r.wo.z <= 0) return mk(0, 0, 0);
            return col(b.p) * kInvPi;
        case DLO_BSDF_MICROFACET: {  // microfacet.cpp:47-58
            float alpha = b.p[0];
            V3 wh = normalized(r.wi + r.wo);
            float D = sq_beckmann_pdf(wh, alpha);
            float F = fresnel(dot(r.wi, wh), b.p[2], b.p[1]);
            float G = mf_g1(r.wi, wh, alpha) * mf_g1(mut(DLO_MUT_MICROFACET_G_WI_TWICE) ? r.wi : r.wo, wh, alpha);
            float deno = 1 / (4 * r.wi.z * r.wo.z);
            return add_scalar(col(b.p + 3) * kInvPi, b.ks * D * F * G * deno);
        }
        case DLO_BSDF_ANISO_GGX: {  // aniso_ggx.cu:38-62
            if (r.wi.z <= 0 || r.wo.z <= 0) return mk(0, 0, 0);
            float ax = b.p[6], ay = b.p[7];
            V3 wh = normalized(r.wi + r.wo);
            float D = ggx_d(wh, ax, ay);
            float F = schlick(dot(r.wi, wh));
            float G = ggx_g1(r.wi, ax, ay) * ggx_g1(r.wo, ax, ay);
            float deno = 1.f / (4.f * r.wi.z * r.wo.z);
            return col(b.p) * kInvPi + (((col(b.p + 3) * D) * F) * G) * deno;
        }
        case DLO_BSDF_CONDUCTOR: {
            if (r.wi.z <= 0 || r.wo.z <= 0) return mk(0, 0, 0);
            float a = b.p[0];
            V3 wh = normalized(r.wi + r.wo);
            float D = ggx_d(wh, a, a);
            float G = ggx_g1(r.wi, a, a) * ggx_g1(r.wo, a, a);
            float cosi = dot(r.wi, wh);
            float deno = 1.f / (4.f * r.wi.z * r.wo.z);
            float s = (D * G) * deno;
            const int ie = mut(DLO_MUT_CONDUCTOR_SWAP_ETA_K) ? 4 : 1, ik = 5 - ie;
            return mk(fresnel_conductor(cosi, b.p[ie], b.p[ik]) * s,
                      fresnel_conductor(cosi, b.p[ie + 1], b.p[ik + 1]) * s,
                      fresnel_conductor(cosi, b.p[ie + 2], b.p[ik + 2]) * s);
        }
        default:  // mirror / dielectric: discrete   mirror.cpp:12-15, dielectric.cpp:40-43
            return mk(0, 0, 0);
    }
}

static float bsdf_pdf(const Bsdf& b, const BRec& r) {
    switch (b.type) {
        case DLO_BSDF_DIFFUSE:  // diffuse.cpp:34-48
            if (r.measure != kSolidAngle || r.wi.z <= 0 || r.wo.z <= 0) return 0.0f;
            if (mut(DLO_MUT_DIFFUSE_UNIFORM_PDF)) return kInvTwoPi;
            return kInvPi * r.wo.z;
        case DLO_BSDF_MICROFACET: {  // microfacet.cpp:61-73
            if (r.measure != kSolidAngle || r.wi.z <= 0 || r.wo.z <= 0) return 0.0f;
            V3 wh = normalized(r.wi + r.wo);
            float diff = (1 - b.ks) * sq_cos_hemi_pdf(r.wo);
            float spec = b.ks * sq_beckmann_pdf(wh, b.p[0]) / 4 / dot(wh, r.wo);
            return diff + spec;
        }
        case DLO_BSDF_ANISO_GGX: {  // aniso_ggx.cu:64-78
            if (r.measure != kSolidAngle || r.wi.z <= 0 || r.wo.z <= 0) return 0.f;
            V3 wh = normalized(r.wi + r.wo);
            float diff = (1 - 0.5f) * sq_cos_hemi_pdf(r.wo);
            float whz = mut(DLO_MUT_GGX_PDF_NO_WHZ) ? 1.0f : wh.z;
            float spec = 0.5f * ggx_d(wh, b.p[6], b.p[7]) * whz / 4.f / dot(wh, r.wo);
            return diff + spec;
        }
        case DLO_BSDF_CONDUCTOR: {
            if (r.measure != kSolidAngle || r.wi.z <= 0 || r.wo.z <= 0) return 0.f;
            V3 wh = normalized(r.wi + r.wo);
            float whz = mut(DLO_MUT_GGX_PDF_NO_WHZ) ? 1.0f : wh.z;
            return ggx_d(wh, b.p[0], b.p[0]) * whz / 4.f / dot(wh, r.wo);
        }
        default:
            return 0.0f;
    }
}

static V3 bsdf_sample(const Bsdf& b, BRec& r, V2 s) {
    switch (b.type) {
        case DLO_BSDF_DIFFUSE:  // diffuse.cpp:51-67
            if (r.wi.z <= 0) return mk(0, 0, 0);
            r.measure = kSolidAngle;
            r.wo = sq_cos_hemi(s);
            r.eta = 1.0f;
            return col(b.p);
        case DLO_BSDF_MIRROR:  // mirror.cpp:17-29
            if (r.wi.z <= 0) return mk(0, 0, 0);
            r.wo = mk(-r.wi.x, -r.wi.y, r.wi.z);
            r.measure = kDiscrete;
            r.eta = 1.0f;
            return mk(1, 1, 1);
        case DLO_BSDF_DIELECTRIC: {  // dielectric.cpp:45-58
            r.measure = kDiscrete;
            float c = r.wi.z;
            const bool swap = mut(DLO_MUT_DIELECTRIC_SWAP_IOR);
            const float int_ior = swap ? b.p[1] : b.p[0], ext_ior = swap ? b.p[0] : b.p[1];
            float ri = mut(DLO_MUT_DIELECTRIC_NO_FRESNEL) ? 0.0f : fresnel(c, ext_ior, int_ior);
            if (s.x < ri) {
                r.wo = mk(-r.wi.x, -r.wi.y, r.wi.z);
                r.eta = 1;
            } else {
                float eta = ext_ior / int_ior;
                r.wo = normalized(refract(r.wi, eta));
                r.eta = c > 0.f ? eta : 1 / eta;
            }
            return mk(1.f, 1.f, 1.f);
        }
        case DLO_BSDF_MICROFACET: {  // microfacet.cpp:76-104
            if (r.wi.z <= 0) return mk(0, 0, 0);
            r.measure = kSolidAngle;
            V2 sm = s;
            float sx = s.x;
            if (sx > b.ks) {
                sm.x = (sx - b.ks) / (1 - b.ks);
                r.wo = sq_cos_hemi(sm);
            } else {
                sm.x = sx / b.ks;
                V3 wh = sq_beckmann(sm, b.p[0]);
                r.wo = reflect(r.wi, wh);
            }
            if (r.wo.z <= 0) return mk(0, 0, 0);
            return bsdf_eval(b, r) * r.wo.z / bsdf_pdf(b, r);
        }
        case DLO_BSDF_ANISO_GGX: {  // aniso_ggx.cu:91-121, lobe picked by sample reuse
            if (r.wi.z <= 0) return mk(0, 0, 0);
            r.measure = kSolidAngle;
            V2 sm = s;
            if (s.x > 0.5f) {
                sm.x = (s.x - 0.5f) / (1 - 0.5f);
                r.wo = sq_cos_hemi(sm);
            } else {
                sm.x = s.x / 0.5f;
                V3 wh = ggx_sample(sm, b.p[6], b.p[7]);
                r.wo = reflect(r.wi, wh);
            }
            if (r.wo.z <= 0) return mk(0, 0, 0);
            return bsdf_eval(b, r) * r.wo.z / bsdf_pdf(b, r);
        }
        case DLO_BSDF_CONDUCTOR: {
            if (r.wi.z <= 0) return mk(0, 0, 0);
            r.measure = kSolidAngle;
            V3 wh = ggx_sample(s, b.p[0], b.p[0]);
            r.wo = reflect(r.wi, wh);
            if (r.wo.z <= 0) return mk(0, 0, 0);
            return bsdf_eval(b, r) * r.wo.z / bsdf_pdf(b, r);
        }
    }
    return mk(0, 0, 0);
}

// ---------------------------------------------------------------------------
// Emitters
// ---------------------------------------------------------------------------
struct DirSample {
    V3 o, d, n;
    float dist = 0, pdf = 0;
    int emitter = -1;
    bool delta = false;
};
// DirectionSample(ori, dst, n)  mesh.h:89-104
static DirSample dir_sample(V3 o, V3 dst, V3 n) {
    DirSample ds;
    ds.o = o;
    ds.n = n;
    V3 vec = dst - o;
    ds.d = normalized(vec);
    ds.dist = length(vec);
    return ds;
}

// DiscretePDF::sample / sampleReuse  dpdf.h:95-136
static uint32_t dpdf_sample_reuse(const std::vector<float>& cdf, float& x) {
    size_t pos = std::lower_bound(cdf.begin(), cdf.end(), x) - cdf.begin();
    size_t index = (size_t)std::max((ptrdiff_t)0, (ptrdiff_t)pos - 1);
    index = std::min(index, cdf.size() - 2);
    x = (x - cdf[index]) / (cdf[index + 1] - cdf[index]);
    return (uint32_t)index;
}

// Mesh::samplePosition  mesh.cpp:222-244
static void sample_position(const Mesh& m, V2 s, V3& pos, V3& nrm) {
    uint32_t f = dpdf_sample_reuse(m.cdf, s.x);
    V3 p0 = m.pos(f, 0), p1 = m.pos(f, 1), p2 = m.pos(f, 2);
    V3 b = sq_triangle(s);
    pos = (p0 * b.x + p1 * b.y) + p2 * b.z;
    if (!m.N.empty())
        nrm = (m.nrm(f, 0) * b.x + m.nrm(f, 1) * b.y) + m.nrm(f, 2) * b.z;
    else
        nrm = cross(p1 - p0, p2 - p1);
    nrm = normalized(nrm);
}

// Bilinear lookup with wrap addressing on a lat-long image, row 0 = top.
static V3 env_lookup(const Emitter& e, float u, float v) {
    float x = u * (float)e.w - 0.5f;
    float y = v * (float)e.h - 0.5f;
    float fx0 = std::floor(x), fy0 = std::floor(y);
    float ax = x - fx0, ay = y - fy0;
    long ix = (long)fx0, iy = (long)fy0;
    long x0 = ((ix % e.w) + e.w) % e.w, x1 = (((ix + 1) % e.w) + e.w) % e.w;
    long y0 = ((iy % e.h) + e.h) % e.h, y1 = (((iy + 1) % e.h) + e.h) % e.h;
    auto px = [&](long xx, long yy) {
        const float* p = &e.rgb[3 * ((size_t)yy * e.w + xx)];
        return mk(p[0], p[1], p[2]);
    };
    float w00 = (1.0f - ax) * (1.0f - ay), w10 = ax * (1.0f - ay);
    float w01 = (1.0f - ax) * ay, w11 = ax * ay;
    V3 r = ((px(x0, y0) * w00 + px(x1, y0) * w10) + px(x0, y1) * w01) + px(x1, y1) * w11;
    return r * e.scale;
}

// Envmap::environment_coords  include/optix/light/envmap.h:25-30
static V2 env_coords(V3 d) {
    float u = (float)std::atan2((double)d.z, (double)d.x) * (float)(0.5 / (double)kPi) + 0.5f;
    float yc = std::max(std::min(d.y, 1.0f), -1.0f);
    float v = (float)std::acos((double)yc) * kInvPi;
    return V2{u, v};
}

// Envmap::eval  envmap.h:89-94
static V3 env_eval(const Emitter& e, V3 wi) {
    V2 uv = env_coords(wi);
    return env_lookup(e, uv.x, uv.y);
}
// Envmap::pdfDirection  envmap.h:81-87
static float env_pdf(const Emitter& e, V3 d) {
    V2 uv = env_coords(d);
    int x = std::min((int)(uv.x * (float)e.w), e.w - 1);
    int y = std::min((int)(uv.y * (float)e.h), e.h - 1);
    return e.pdf[(size_t)y * e.w + x];
}

struct Sampler {
    Pcg32 rng;
    V2 next2d() { return rng.next2d(); }
    float next1d() { return rng.next_float(); }
};
// A sampler whose one-dimensional draw is given (emitter queries).
struct FixedSampler {
    float x;
    float next1d() const { return x; }
};

// Emitter::sampleDirection dispatch:
//  area   src/tracer/emitter/area.cpp:11-37
//  point  include/optix/light/point.h:15-31 (receiver cosine, delta)
//  env    include/optix/light/envmap.h:32-79 (third uniform drawn here)
template <class SamplerT>
static V3 emitter_sample(const Scene& s, const Emitter& e, V2 smp, const Its& its, DirSample& ds,
                         SamplerT& sampler) {
    if (e.type == DLO_EMITTER_AREA) {
        const Mesh& m = s.meshes[e.mesh];
        V3 position, normal;
        sample_position(m, smp, position, normal);
        ds = dir_sample(its.p, position, normal);
        float pA = m.normalization;
        float dp = std::max(dot(normal, -ds.d), 0.0f);
        // pA * dist * dist / dp, evaluated left to right as area.cpp:29
        float pw = dp == 0 ? 0 : mut(DLO_MUT_AREA_PDF_NO_DIST2) ? pA / dp : pA * ds.dist * ds.dist / dp;
        ds.pdf = pw;
        return pw != 0 ? (e.radiance * std::max(dot(its.sh.n, ds.d), 0.0f)) / pw : mk(0, 0, 0);
    }
    if (e.type == DLO_EMITTER_POINT) {
        ds = dir_sample(its.p, e.position, mk(0, 0, 0));
        ds.pdf = 1;
        ds.delta = true;
        float inv = 1.0f / ds.dist;
        float c = std::max(dot(its.sh.n, ds.d), 0.0f);
        if (mut(DLO_MUT_POINT_INV_DIST)) return c > 0.f ? (e.intensity * c) * inv : mk(0, 0, 0);
        return c > 0.f ? ((e.intensity * c) * inv) * inv : mk(0, 0, 0);
    }
    // environment map
    float xz = sampler.next1d();
    uint32_t size = (uint32_t)e.w * (uint32_t)e.h;
    uint32_t idx = std::min((uint32_t)(smp.x * (float)size), size - 1);
    uint32_t env_idx;
    float xy = smp.y;
    if (xy < e.q[idx]) {
        env_idx = idx;
        xy /= e.q[idx];
    } else {
        env_idx = e.alias[idx];
        xy = (xy - e.q[idx]) / (1.0f - e.q[idx]);
    }
    uint32_t py = env_idx / (uint32_t)e.w, px = env_idx % (uint32_t)e.w;
    ds.pdf = e.pdf[env_idx];
    float u = ((float)px + xy) / (float)e.w;
    float phi = (float)((double)u * 2.0 * (double)kPi - (double)kPi);
    float step = kPi / (float)e.h;
    float th0 = (float)py * step;
    float c0 = (float)dcos(th0), c1 = (float)dcos(th0 + step);
    float ct = c0 * (1.0f - xz) + c1 * xz;
    float theta = (float)std::acos((double)ct);
    float st = (float)dsin(theta);
    V3 dir = mk((float)dcos(phi) * st, ct, (float)dsin(phi) * st);
    ds.o = its.p;
    ds.d = dir;
    ds.dist = 1e16f;
    ds.delta = false;
    ds.n = -dir;
    float v = theta * kInvPi;
    V3 val = env_lookup(e, u, v);
    float c = std::max(dot(its.sh.n, ds.d), 0.0f);
    return c > 0.f ? (val * c) / ds.pdf : mk(0, 0, 0);
}

// Scene::sampleEmitterDirection  src/tracer/scene/scene.cpp:139-170
static V3 sample_emitter_direction(const Scene& s, const Its& its, V2 smp, DirSample& ds, Sampler& sampler,
                                   int mode) {
    V3 spectrum;
    size_t n = s.emitters.size();
    if (n == 0) {
        spectrum = mk(0, 0, 0);
    } else if (n == 1) {
        spectrum = emitter_sample(s, s.emitters[0], smp, its, ds, sampler);
        ds.emitter = 0;
    } else {
        float epdf = 1.f / (float)n;
        size_t index = std::min((size_t)(smp.x * (float)n), n - 1);
        smp.x = smp.x * (float)n - (float)index;
        spectrum = emitter_sample(s, s.emitters[index], smp, its, ds, sampler);
        spectrum = spectrum * (float)n;
        ds.pdf *= epdf;
        ds.emitter = (int)index;
    }
    return spectrum;
}

// Scene::pdfEmitterDirection  scene.cpp:172-179 (+ env pdf, envmap.h:81-87)
static float pdf_emitter_direction(const Scene& s, int e, const DirSample& ds) {
    const Emitter& em = s.emitters[e];
    float p;
    if (em.type == DLO_EMITTER_AREA) {
        // AreaLight::pdfDirection  area.cpp:39-44
        float pA = s.meshes[em.mesh].normalization;
        float dp = std::max(dot(ds.n, -ds.d), 0.0f);
        p = dp == 0 ? 0 : mut(DLO_MUT_AREA_PDF_NO_DIST2) ? pA / dp : pA * ds.dist * ds.dist / dp;
    } else if (em.type == DLO_EMITTER_POINT) {
        p = 0;
    } else {
        p = env_pdf(em, ds.d);
    }
    if (s.emitters.size() == 1) return p;
    return p / (float)s.emitters.size();
}

static inline float mis_weight(float a, float b) {  // path.cpp:104-108
    a *= a;
    b *= b;
    return a / (a + b);
}

// ---------------------------------------------------------------------------
// PathIntegrator::Li  src/tracer/integrator/path.cpp:13-100
// ---------------------------------------------------------------------------
struct RenderOpts {
    int max_depth = -1;
    int rr_depth = 4;
    int accel = DLO_ACCEL_OCTREE_ROBUST;
};

static bool intersect(const Scene& s, const RenderOpts& o, const Ray& ray, Its& its) {
    HitRec h;
    if (!trace_closest(s, o.accel, ray, h)) return false;
    make_its(s, h, its);
    return true;
}

static V3 area_eval(const Emitter& e, const Its& its, V3 wi) {  // area.cpp:46-49
    float c = to_local(its.sh, wi).z;
    return c > 0.f ? e.radiance : mk(0, 0, 0);
}

// With hide_lights, a ray that hits an area light's surface continues from
// the hit point in the same direction (src/optix/cuda/hitgroup_pgs.cu:55-64:
// fr = 1, eta = 1, p = its.p, wo = ray direction), drawing no random numbers
// and counting no bounce; the camera ray does not see the light.
static bool hidden(const Scene& s, const Its& its) { return s.hide_lights && s.meshes[its.mesh].emitter >= 0; }

static V3 Li(const Scene& s, const RenderOpts& o, Sampler& sampler, Ray ray, uint64_t* counters) {
    float eta = 1.f;
    V3 result = mk(0, 0, 0), thr = mk(1, 1, 1);
    Its its;
    for (;;) {
        counters[0]++;
        if (!intersect(s, o, ray, its)) {
            if (s.env >= 0) {  // miss at depth 0: src/optix/cuda/miss.cu:23-29
                const Emitter& e = s.emitters[s.env];
                result = e.visual ? env_eval(e, ray.d) : s.bg;
            }
            return result;
        }
        if (!hidden(s, its)) break;
        ray = make_ray(its.p, ray.d, s.eps, INFINITY);
    }
    {
        const Mesh& m = s.meshes[its.mesh];
        if (m.emitter >= 0) result = result + cmul(thr, area_eval(s.emitters[m.emitter], its, -ray.d));
    }
    for (int depth = 0;; ++depth) {
        if (o.max_depth >= 0 && depth + 2 > o.max_depth) break;
        const Bsdf bsdf = resolve_bsdf(s, s.bsdfs[s.meshes[its.mesh].bsdf], its.tc);
        // Emitter sampling  path.cpp:32-56
        DirSample ds;
        V2 lsmp = sampler.next2d();
        V3 emitter_val = sample_emitter_direction(s, its, lsmp, ds, sampler, o.accel);
        if (csum(emitter_val) > 0) {  // visibility  scene.cpp:165-169
            Ray sr = make_ray(its.p, ds.d, s.eps, ds.dist - s.eps);
            counters[1]++;
            float occluded = trace_any(s, o.accel, sr) ? 1.0f : 0.0f;
            emitter_val = emitter_val * (1 - occluded);
        }
        V3 wi = to_local(its.sh, -normalized(ray.d));
        V3 wo = to_local(its.sh, ds.d);
        BRec q;
        q.wi = wi;
        q.wo = wo;
        q.measure = kSolidAngle;
        if (ds.pdf > 0) {
            V3 bv = bsdf_eval(bsdf, q);
            float bp = bsdf_pdf(bsdf, q);
            float mis = ds.delta ? 1.0f : mis_weight(ds.pdf, bp);
            result = result + cmul(cmul(thr * mis, bv), emitter_val);
        }
        // BSDF sampling  path.cpp:58-69
        BRec br;  // BSDFQueryRecord(its, wi): wo value-initialised to zero
        br.wi = wi;
        br.wo = mk(0, 0, 0);
        V3 fr = bsdf_sample(bsdf, br, sampler.next2d());
        thr = cmul(thr, fr);
        eta *= br.eta;
        if (is_zero(thr)) break;  // every later contribution is multiplied by zero
        ray = make_ray(its.p, to_world(its.sh, br.wo), s.eps, INFINITY);
        Its nx;
        bool hit;
        for (;;) {
            counters[0]++;
            hit = intersect(s, o, ray, nx);
            if (!hit) break;
            const Mesh& m = s.meshes[nx.mesh];
            if (m.emitter >= 0) {  // path.cpp:72-84
                const Emitter& e = s.emitters[m.emitter];
                V3 ev = area_eval(e, nx, -ray.d);
                DirSample hs = dir_sample(ray.o, nx.p, nx.sh.n);
                float epdf = pdf_emitter_direction(s, m.emitter, hs);
                float bpdf = bsdf_pdf(bsdf, br);
                float mis = bsdf_is_diffuse(bsdf) && !mut(DLO_MUT_EMITTER_HIT_NO_MIS) ? mis_weight(bpdf, epdf) : 1.f;
                result = result + cmul(thr * mis, ev);
            }
            if (!hidden(s, nx)) break;
            ray = make_ray(nx.p, ray.d, s.eps, INFINITY);
        }
        if (hit) {
            its = nx;
        } else {
            if (s.env >= 0) {  // miss at depth > 0: miss.cu:31-42
                const Emitter& e = s.emitters[s.env];
                V3 ev = env_eval(e, ray.d);
                DirSample hs;
                hs.d = ray.d;
                float epdf = pdf_emitter_direction(s, s.env, hs);
                float bpdf = bsdf_pdf(bsdf, br);
                float mis = bsdf_is_diffuse(bsdf) && !mut(DLO_MUT_EMITTER_HIT_NO_MIS) ? mis_weight(bpdf, epdf) : 1.f;
                result = result + cmul(thr * mis, ev);
            }
            break;
        }
        // Russian roulette  path.cpp:89-97
        if (depth >= o.rr_depth) {
            float p = std::min(cmax((thr * eta) * eta), 0.99f);
            if (sampler.next1d() > p) break;
            if (!mut(DLO_MUT_RR_NO_DIVIDE)) thr = thr / p;
        }
    }
    return result;
}

// Matrix4f * Vector4f: each row accumulated from zero in index order
// (include/core/math/matrix.h:255-262 via vector.h:329-336).
static void mat4_apply(const float* m, const float in[4], float out[4]) {
    for (int j = 0; j < 4; j++) {
        float s = 0.0f;
        for (int k = 0; k < 4; k++) s += m[4 * j + k] * in[k];
        out[j] = s;
    }
}

// PerspectiveCamera::sampleRay  src/tracer/camera/perspective.cpp:75-95:
// near-plane point = sampleToCamera * (pos / size, 0) (homogeneous divide),
// d = normalize(nearP), o = cameraToWorld * (0,0,0), d' = cameraToWorld * d,
// [mint, maxt] = [near, far] / d.z.
static Ray perspective_ray(const Camera& c, V2 pos) {
    const float in[4] = {pos.x * c.inv_size.x, pos.y * c.inv_size.y, 0.0f, 1.0f};
    float q[4];
    mat4_apply(c.S, in, q);
    const V3 d = normalized(mk(q[0] / q[3], q[1] / q[3], q[2] / q[3]));
    const float inv_z = 1.0f / d.z;
    const float o_in[4] = {0.0f, 0.0f, 0.0f, 1.0f}, d_in[4] = {d.x, d.y, d.z, 0.0f};
    float o4[4], d4[4];
    mat4_apply(c.C, o_in, o4);
    mat4_apply(c.C, d_in, d4);
    return make_ray(mk(o4[0] / o4[3], o4[1] / o4[3], o4[2] / o4[3]), mk(d4[0], d4[1], d4[2]), c.near_clip * inv_z,
                    c.far_clip * inv_z);
}

// OpencvCamera::sampleRay as the GPU path casts it (include/optix/camera/
// opencv.h:28-46): (pos - principal point) / focal length at z = 1, w = 1,
// mapped by the inverse extrinsic matrix; the ray leaves the eye toward it.
static Ray opencv_ray(const Camera& c, V2 pos, float eps) {
    const float in[4] = {(pos.x - c.intr[0]) / c.intr[2], (pos.y - c.intr[1]) / c.intr[3], 1.0f, 1.0f};
    float q[4];
    mat4_apply(c.C, in, q);
    const V3 d = normalized(mk(q[0], q[1], q[2]) - c.eye);
    return make_ray(c.eye, d, eps, INFINITY);
}

// VirtualCamera::sampleRay  src/tracer/camera/virtual.cpp:52-65
static Ray camera_ray(const Camera& c, V2 pos, float eps) {
    if (c.type == DLO_CAMERA_PERSPECTIVE) return perspective_ray(c, pos);
    if (c.type == DLO_CAMERA_OPENCV) return opencv_ray(c, pos, eps);
    float lx = pos.x, ly = ((float)c.height - pos.y) - 1;
    float px = lx * c.inv_size.x * 2 - 1;
    float py = ly * c.inv_size.y * 2 - 1;
    V3 d = normalized((c.U * px + c.V * py) + c.W);
    return make_ray(c.eye, d, eps, INFINITY);
}

// NormalIntegrator::Li  src/tracer/integrator/normal.cpp:9-19
static V3 Li_normals(const Scene& s, const RenderOpts& o, const Ray& ray, uint64_t* counters) {
    Its its;
    counters[0]++;
    if (!intersect(s, o, ray, its)) return mk(0, 0, 0);
    return mk(std::fabs(its.sh.n.x), std::fabs(its.sh.n.y), std::fabs(its.sh.n.z));
}

static inline uint64_t stream_state(uint32_t pixel, uint32_t seed) {
    return (uint64_t)pixel + ((uint64_t)seed << 32);
}

// One camera sample: the body of CPURenderer::renderBlock's inner loop
// src/tracer/backend/cpu_renderer.cpp:36-51, with a per-(pixel, sample)
// pcg32 stream.
static V3 render_sample(const Scene& s, const RenderOpts& o, int x, int y, uint32_t sidx, uint32_t seed,
                        V2& pos, uint64_t* counters) {
    Sampler sampler;
    sampler.rng.seed(stream_state((uint32_t)y * (uint32_t)s.cam.width + (uint32_t)x, seed), sidx);
    V2 ps = sampler.next2d();
    pos = V2{(float)x + ps.x, (float)y + ps.y};
    sampler.next2d();  // aperture sample, unused by every camera
    Ray ray = camera_ray(s.cam, pos, s.eps);
    if (s.normals) return Li_normals(s, o, ray, counters);
    return Li(s, o, sampler, ray, counters);
}

// ---------------------------------------------------------------------------
// Film: ImageBlock  src/core/bitmap/block.cpp:10-129
// ---------------------------------------------------------------------------
struct Block {
    int ox, oy, sx, sy;  // offset and size of the pixel region
    int border;
    std::vector<float> data;  // (sx + 2b) * (sy + 2b) * 4
    int cols() const { return sx + 2 * border; }
    int rows() const { return sy + 2 * border; }
};

static bool valid(V3 c) {  // Color3f::isValid  color.h:50-58
    float v[3] = {c.x, c.y, c.z};
    for (int i = 0; i < 3; i++)
        if (v[i] < 0 || !std::isfinite(v[i])) return false;
    return true;
}

static void block_put(const Filter& f, Block& b, V2 P, V3 value) {
    if (!valid(value)) return;
    float posx = P.x - 0.5f - (float)(b.ox - b.border);
    float posy = P.y - 0.5f - (float)(b.oy - b.border);
    int x0 = (int)std::ceil(posx - f.radius), y0 = (int)std::ceil(posy - f.radius);
    int x1 = (int)std::floor(posx + f.radius), y1 = (int)std::floor(posy + f.radius);
    x0 = std::max(x0, 0);
    y0 = std::max(y0, 0);
    x1 = std::min(x1, b.cols() - 1);
    y1 = std::min(y1, b.rows() - 1);
    float wx[16], wy[16];
    for (int x = x0, i = 0; x <= x1; ++x) wx[i++] = f.table[(int)(std::fabs((float)x - posx) * f.lookup)];
    for (int y = y0, i = 0; y <= y1; ++y) wy[i++] = f.table[(int)(std::fabs((float)y - posy) * f.lookup)];
    for (int y = y0, yr = 0; y <= y1; ++y, ++yr)
        for (int x = x0, xr = 0; x <= x1; ++x, ++xr) {
            float* c = &b.data[4 * ((size_t)y * b.cols() + x)];
            c[0] += (value.x * wx[xr]) * wy[yr];
            c[1] += (value.y * wx[xr]) * wy[yr];
            c[2] += (value.z * wx[xr]) * wy[yr];
            c[3] += (1.0f * wx[xr]) * wy[yr];
        }
}

// BlockGenerator spiral  block.cpp:131-177
static std::vector<std::pair<int, int>> spiral_blocks(int W, int H, int bs) {
    int nx = (int)std::ceil(W / (float)bs), ny = (int)std::ceil(H / (float)bs);
    int left = nx * ny;
    int dir = 0, bx = nx / 2, by = ny / 2, steps_left = 1, nsteps = 1;
    std::vector<std::pair<int, int>> out;
    while (left > 0) {
        out.push_back({bx * bs, by * bs});
        if (--left == 0) break;
        do {
            switch (dir) {
                case 0: ++bx; break;
                case 1: ++by; break;
                case 2: --bx; break;
                case 3: --by; break;
            }
            if (--steps_left == 0) {
                dir = (dir + 1) % 4;
                if (dir == 2 || dir == 0) ++nsteps;
                steps_left = nsteps;
            }
        } while (bx < 0 || by < 0 || bx >= nx || by >= ny);
    }
    return out;
}

struct RenderJob {
    const Scene* s;
    RenderOpts o;
    int spp;
    uint32_t seed;
    std::vector<std::pair<int, int>> blocks;
    std::vector<Block> results;
    std::atomic<int> next{0};
    std::vector<uint64_t> counters;  // per thread: [closest, shadow]
};

static void render_worker(RenderJob* job, int tid) {
    const Scene& s = *job->s;
    const int BS = 32;
    uint64_t* cnt = &job->counters[2 * tid];
    for (;;) {
        int bi = job->next.fetch_add(1);
        if (bi >= (int)job->blocks.size()) break;
        Block& b = job->results[bi];
        b.ox = job->blocks[bi].first;
        b.oy = job->blocks[bi].second;
        b.sx = std::min(BS, s.cam.width - b.ox);
        b.sy = std::min(BS, s.cam.height - b.oy);
        b.border = s.filter.border;
        b.data.assign((size_t)b.cols() * b.rows() * 4, 0.0f);
        for (int y = 0; y < b.sy; ++y)
            for (int x = 0; x < b.sx; ++x)
                for (int i = 0; i < job->spp; ++i) {
                    V2 pos;
                    V3 v = render_sample(s, job->o, x + b.ox, y + b.oy, (uint32_t)i, job->seed, pos, cnt);
                    block_put(s.filter, b, pos, v);
                }
    }
}

}  // namespace

// ===========================================================================
// C API
// ===========================================================================
namespace {
template <typename F>
static void parallel_for(uint32_t n, F&& fn) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 64 || nt == 1) {
        fn(0u, n);
        return;
    }
    std::vector<std::thread> pool;
    uint32_t chunk = (n + nt - 1) / nt;
    for (unsigned k = 0; k < nt; k++) {
        uint32_t b = k * chunk, e = std::min(n, b + chunk);
        if (b < e) pool.emplace_back([&fn, b, e] { fn(b, e); });
    }
    for (auto& th : pool) th.join();
}

}  // namespace

extern "C" {

void* dlo_scene_new(void) { return new Scene(); }
void dlo_scene_free(void* h) { delete (Scene*)h; }

int dlo_scene_add_bsdf(void* h, int type, const float* params, int n) {
    Scene& s = *(Scene*)h;
    Bsdf b;
    b.type = type;
    for (int i = 0; i < n && i < 16; i++) b.p[i] = params[i];
    if (type == DLO_BSDF_MICROFACET) b.ks = 1 - cmax(col(b.p + 3));
    s.bsdfs.push_back(b);
    return (int)s.bsdfs.size() - 1;
}

int dlo_scene_add_texture(void* h, const float* rgb, int w, int hh) {
    Scene& s = *(Scene*)h;
    Tex t;
    t.w = w;
    t.h = hh;
    t.rgb.assign(rgb, rgb + 3 * (size_t)w * (size_t)hh);
    s.textures.push_back(std::move(t));
    return (int)s.textures.size() - 1;
}

// Texture lookups at n (u, v) pairs (tests of the sampler).
void dlo_texture_sample(const float* rgb, int w, int hh, const float* uv, uint32_t n, float* out) {
    Tex t;
    t.w = w;
    t.h = hh;
    t.rgb.assign(rgb, rgb + 3 * (size_t)w * (size_t)hh);
    for (uint32_t i = 0; i < n; i++) {
        const V3 c = tex_sample(t, uv[2 * i], uv[2 * i + 1]);
        out[3 * i] = c.x;
        out[3 * i + 1] = c.y;
        out[3 * i + 2] = c.z;
    }
}

void dlo_scene_set_bsdf_textures(void* h, int bsdf, const int* tex5, int object_space) {
    Scene& s = *(Scene*)h;
    Bsdf& b = s.bsdfs[bsdf];
    for (int k = 0; k < 5; k++) b.tex[k] = tex5[k];
    b.object_space = object_space;
}

// Mesh::activate  src/tracer/mesh/mesh.cpp:15-45 (area CDF, tangents)
int dlo_scene_add_mesh(void* h, const float* V, uint32_t nv, const uint32_t* F, uint32_t nf, const float* N,
                       const float* UV, int bsdf, const float* radiance) {
    Scene& s = *(Scene*)h;
    Mesh m;
    m.V.assign(V, V + 3 * (size_t)nv);
    m.F.assign(F, F + 3 * (size_t)nf);
    if (N) m.N.assign(N, N + 3 * (size_t)nv);
    if (UV) m.UV.assign(UV, UV + 2 * (size_t)nv);
    m.bsdf = bsdf;
    m.bbox.reset();
    for (uint32_t i = 0; i < nv; i++) m.bbox.expand(&m.V[3 * i]);
    m.cdf.reserve(nf + 1);
    m.cdf.push_back(0.0f);
    for (uint32_t f = 0; f < nf; f++) {
        V3 p0 = m.pos(f, 0), p1 = m.pos(f, 1), p2 = m.pos(f, 2);
        float area = 0.5f * length(cross(p1 - p0, p2 - p0));  // Mesh::surfaceArea mesh.cpp:102-108
        m.cdf.push_back(m.cdf.back() + area);
    }
    float sum = m.cdf.back();
    if (sum > 0) {  // DiscretePDF::normalize  dpdf.h:72-84
        m.normalization = 1.0f / sum;
        for (size_t i = 1; i < m.cdf.size(); ++i) m.cdf[i] *= m.normalization;
        m.cdf.back() = 1.0f;
    } else {
        m.normalization = 0.0f;
    }
    if (UV) {
        // Mesh::calTangent  mesh.cpp:144-177; each face overwrites its vertices
        m.T.assign(3 * (size_t)nv, 0.0f);
        for (uint32_t f = 0; f < nf; f++) {
            V3 e1 = m.pos(f, 1) - m.pos(f, 0), e2 = m.pos(f, 2) - m.pos(f, 0);
            V2 t0 = m.tex(f, 0), t1 = m.tex(f, 1), t2 = m.tex(f, 2);
            float a = t1.x - t0.x, b = t1.y - t0.y, c = t2.x - t0.x, d = t2.y - t0.y;
            float det = 1 / (a * d - b * c);
            float i00 = d / det, i01 = -b / det, i10 = -c / det, i11 = a / det;
            float r0[3], ex[3] = {e1.x, e1.y, e1.z}, ey[3] = {e2.x, e2.y, e2.z};
            for (int j = 0; j < 3; j++) {
                float acc = 0.0f;
                acc += i00 * ex[j];
                acc += i01 * ey[j];
                r0[j] = acc;
            }
            (void)i10;
            (void)i11;
            V3 T = normalized(mk(r0[0], r0[1], r0[2]));
            for (int k = 0; k < 3; k++) {
                uint32_t p = m.F[3 * f + k];
                m.T[3 * p] = T.x;
                m.T[3 * p + 1] = T.y;
                m.T[3 * p + 2] = T.z;
            }
        }
    }
    m.prim_offset = s.nprims;
    s.nprims += nf;
    if (radiance) {
        Emitter e;
        e.type = DLO_EMITTER_AREA;
        e.radiance = mk(radiance[0], radiance[1], radiance[2]);
        e.mesh = (int)s.meshes.size();
        m.emitter = (int)s.emitters.size();
        s.emitters.push_back(e);
    }
    s.meshes.push_back(std::move(m));
    s.built = false;
    return (int)s.meshes.size() - 1;
}

int dlo_scene_add_point_light(void* h, const float* position, const float* intensity) {
    Scene& s = *(Scene*)h;
    Emitter e;
    e.type = DLO_EMITTER_POINT;
    e.position = mk(position[0], position[1], position[2]);
    e.intensity = mk(intensity[0], intensity[1], intensity[2]);
    s.emitters.push_back(e);
    return (int)s.emitters.size() - 1;
}

// EnvironmentMap::activate + build_alias_map  src/tracer/emitter/envmap.cpp:59-141
// (importance rows taken in the same orientation the lookup uses).
int dlo_scene_add_envmap(void* h, const float* rgb, int w, int hgt, float scale, int visual) {
    Scene& s = *(Scene*)h;
    if (s.env >= 0) return -1;
    Emitter e;
    e.type = DLO_EMITTER_ENV;
    e.w = w;
    e.h = hgt;
    e.rgb.assign(rgb, rgb + 3 * (size_t)w * hgt);
    e.scale = scale;
    e.visual = visual != 0;
    uint32_t size = (uint32_t)w * (uint32_t)hgt;
    std::vector<float> imp(size);
    float cos0 = 1.0f;
    const float step_phi = (float)(2.0 * (double)kPi) / (float)w;
    const float step_theta = kPi / (float)hgt;
    for (int y = 0; y < hgt; ++y) {
        float th1 = (float)(y + 1) * step_theta;
        float cos1 = (float)dcos(th1);
        float area = (cos0 - cos1) * step_phi;
        cos0 = cos1;
        for (int x = 0; x < w; ++x) {
            size_t i = (size_t)y * w + x;
            const float* p = &e.rgb[3 * i];
            imp[i] = area * std::max(p[0], std::max(p[1], p[2]));
        }
    }
    e.q.resize(size);
    e.alias.resize(size);
    e.pdf.resize(size);
    float sum = 0.0f;
    for (uint32_t i = 0; i < size; ++i) sum += imp[i];
    for (uint32_t i = 0; i < size; ++i) e.q[i] = (float)size * imp[i] / sum;
    std::vector<uint32_t> part(size);
    uint32_t sm = 0u, large = size;
    for (uint32_t i = 0; i < size; ++i) {
        e.alias[i] = i;
        part[(e.q[i] < 1.0f) ? (sm++) : (--large)] = i;
    }
    for (sm = 0; sm < large && large < size; ++sm) {
        const uint32_t j = part[sm], k = part[large];
        e.alias[j] = k;
        e.q[k] += e.q[j] - 1.0f;
        large = (e.q[k] < 1.0f) ? (large + 1u) : large;
    }
    const float inv = 1.0f / sum;
    for (uint32_t i = 0; i < size; ++i) {
        const float* p = &e.rgb[3 * i];
        e.pdf[i] = (mut(DLO_MUT_ENV_PDF_TIMES_AREA) ? imp[i] : std::max(p[0], std::max(p[1], p[2]))) * inv;
    }
    s.env = (int)s.emitters.size();
    s.emitters.push_back(std::move(e));
    return s.env;
}

// VirtualCamera ctor  src/tracer/camera/virtual.cpp:16-42
void dlo_scene_set_camera_virtual(void* h, const float* eye, const float* lookat, const float* up, float fov,
                                  int width, int height) {
    Scene& s = *(Scene*)h;
    Camera& c = s.cam;
    c.type = DLO_CAMERA_VIRTUAL;
    c.width = width;
    c.height = height;
    c.inv_size = V2{1.f / width, 1.f / height};
    float aspect = width / (float)height;
    c.eye = mk(eye[0], eye[1], eye[2]);
    V3 la = mk(lookat[0], lookat[1], lookat[2]), upv = mk(up[0], up[1], up[2]);
    c.W = la - c.eye;
    float wlen = length(c.W);
    c.U = normalized(cross(c.W, upv));
    c.V = normalized(cross(c.U, c.W));
    float vlen = wlen * tanf(0.5f * fov * kPi / 180.f);
    c.V = c.V * vlen;
    float ulen = vlen * aspect;
    c.U = c.U * ulen;
}

// A virtual camera given by its basis (eye, U, V, W as the VirtualCamera
// constructor leaves them).
void dlo_scene_set_camera_basis(void* h, const float* eye_u_v_w, int width, int height) {
    Scene& s = *(Scene*)h;
    Camera& c = s.cam;
    c.type = DLO_CAMERA_VIRTUAL;
    c.width = width;
    c.height = height;
    c.inv_size = V2{1.f / width, 1.f / height};
    V3* dst[4] = {&c.eye, &c.U, &c.V, &c.W};
    for (int i = 0; i < 4; i++) *dst[i] = mk(eye_u_v_w[3 * i], eye_u_v_w[3 * i + 1], eye_u_v_w[3 * i + 2]);
}

// The film filter given by its ImageBlock tabulation (radius and the 33
// table entries).
void dlo_scene_set_filter_table(void* h, float radius, const float* table33) {
    Scene& s = *(Scene*)h;
    Filter& f = s.filter;
    f.radius = radius;
    f.border = (int)std::ceil(radius - 0.5f);
    f.lookup = 32.0f / radius;
    for (int i = 0; i < 33; i++) f.table[i] = table33[i];
}

// PerspectiveCamera after activate(): the matrices are inputs (row-major).
void dlo_scene_set_camera_perspective(void* h, const float* sample_to_camera, const float* camera_to_world,
                                      int width, int height, float near_clip, float far_clip) {
    Scene& s = *(Scene*)h;
    Camera& c = s.cam;
    c.type = DLO_CAMERA_PERSPECTIVE;
    c.width = width;
    c.height = height;
    c.inv_size = V2{1.f / width, 1.f / height};
    for (int i = 0; i < 16; i++) {
        c.S[i] = sample_to_camera[i];
        c.C[i] = camera_to_world[i];
    }
    c.near_clip = near_clip;
    c.far_clip = far_clip;
}

// OpencvCamera (opencv.cpp:9-24) with its inverse extrinsic, eye and
// intrinsics given.
void dlo_scene_set_camera_opencv(void* h, const float* camera_to_world, const float* eye, const float* intrinsics,
                                 int width, int height) {
    Scene& s = *(Scene*)h;
    Camera& c = s.cam;
    c.type = DLO_CAMERA_OPENCV;
    c.width = width;
    c.height = height;
    c.inv_size = V2{1.f / width, 1.f / height};
    for (int i = 0; i < 16; i++) c.C[i] = camera_to_world[i];
    c.eye = mk(eye[0], eye[1], eye[2]);
    for (int i = 0; i < 4; i++) c.intr[i] = intrinsics[i];
}

void dlo_scene_set_options(void* h, int normals, int hide_lights, float eps) {
    Scene& s = *(Scene*)h;
    s.normals = normals != 0;
    s.hide_lights = hide_lights != 0;
    s.eps = eps > 0.0f ? eps : kEpsilon;
}

void dlo_scene_get_camera(void* h, float* out12) {
    Scene& s = *(Scene*)h;
    V3 v[4] = {s.cam.eye, s.cam.U, s.cam.V, s.cam.W};
    for (int i = 0; i < 4; i++) {
        out12[3 * i] = v[i].x;
        out12[3 * i + 1] = v[i].y;
        out12[3 * i + 2] = v[i].z;
    }
}

// ReconstructionFilter table, ImageBlock ctor  block.cpp:12-27; filters
// src/tracer/rfilter/rfilter.cpp:13-98
void dlo_scene_set_filter(void* h, int type, float radius, float p0, float p1) {
    Scene& s = *(Scene*)h;
    Filter& f = s.filter;
    if (type == DLO_FILTER_BOX) radius = 0.5f;
    if (type == DLO_FILTER_TENT) radius = 1.0f;
    f.radius = radius;
    f.border = (int)std::ceil(radius - 0.5f);
    for (int i = 0; i < 32; ++i) {
        float x = (radius * i) / 32;
        float val = 0;
        if (type == DLO_FILTER_GAUSSIAN) {
            float alpha = -1.0f / (2.0f * p0 * p0);
            val = std::max(0.0f, std::exp(alpha * x * x) - std::exp(alpha * radius * radius));
        } else if (type == DLO_FILTER_BOX) {
            val = 1.0f;
        } else if (type == DLO_FILTER_TENT) {
            val = std::max(0.0f, 1.0f - std::fabs(x));
        } else {  // Mitchell-Netravali, B = p0, C = p1
            x = std::fabs(2.0f * x / radius);
            float x2 = x * x, x3 = x2 * x;
            if (x < 1)
                val = 1.0f / 6.0f * ((12 - 9 * p0 - 6 * p1) * x3 + (-18 + 12 * p0 + 6 * p1) * x2 + (6 - 2 * p0));
            else if (x < 2)
                val = 1.0f / 6.0f *
                      ((-p0 - 6 * p1) * x3 + (6 * p0 + 30 * p1) * x2 + (-12 * p0 - 48 * p1) * x + (8 * p0 + 24 * p1));
            else
                val = 0.0f;
        }
        f.table[i] = val;
    }
    f.table[32] = 0.0f;
    f.lookup = 32 / radius;
}

void dlo_scene_get_filter_table(void* h, float* out33) {
    Scene& s = *(Scene*)h;
    memcpy(out33, s.filter.table, sizeof(float) * 33);
}

void dlo_scene_set_background(void* h, const float* rgb) {
    Scene& s = *(Scene*)h;
    s.bg = mk(rgb[0], rgb[1], rgb[2]);
}

// OCTree::build  octree.cpp:135-151
int dlo_scene_build(void* h) {
    Scene& s = *(Scene*)h;
    s.oct.clear();
    Box bb;
    bb.reset();
    std::vector<std::pair<int, int>> tris;
    for (int m = 0; m < (int)s.meshes.size(); m++) {
        for (uint32_t f = 0; f < s.meshes[m].ntri(); f++) tris.push_back({m, (int)f});
        bb.expand(s.meshes[m].bbox);
    }
    s.oct.reserve(tris.size() / 4 + 16);
    s.oct_root = oc_build(s, bb, tris);
    s.built = true;
    return (int)s.oct.size();
}

uint32_t dlo_scene_num_prims(void* h) { return ((Scene*)h)->nprims; }

// Closest hits for a batch of rays stored as 8 planes of n floats
// (ox, oy, oz, dx, dy, dz, tmin, tmax).  prim = 0xffffffff on a miss.
void dlo_trace_closest(void* h, int mode, const float* rays, uint32_t n, uint32_t* prim, float* t, float* u,
                       float* v) {
    const Scene& s = *(Scene*)h;
    parallel_for(n, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; i++) {
            Ray r = make_ray(mk(rays[i], rays[n + i], rays[2 * n + i]),
                             mk(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]), rays[6 * n + i], rays[7 * n + i]);
            HitRec hr;
            if (trace_closest(s, mode, r, hr)) {
                prim[i] = gid(s, hr.mesh, hr.tri);
                t[i] = hr.t;
                u[i] = hr.u;
                v[i] = hr.v;
            } else {
                prim[i] = 0xffffffffu;
                t[i] = rays[7 * n + i];
                u[i] = v[i] = 0.0f;
            }
        }
    });
}

void dlo_trace_shadow(void* h, int mode, const float* rays, uint32_t n, uint8_t* occluded) {
    const Scene& s = *(Scene*)h;
    parallel_for(n, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; i++) {
            Ray r = make_ray(mk(rays[i], rays[n + i], rays[2 * n + i]),
                             mk(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]), rays[6 * n + i], rays[7 * n + i]);
            occluded[i] = trace_any(s, mode, r) ? 1 : 0;
        }
    });
}

// Camera rays for explicit sample positions (P = pixel + xi).
void dlo_camera_rays(void* h, const float* pos, uint32_t n, float* rays) {
    const Scene& s = *(Scene*)h;
    for (uint32_t i = 0; i < n; i++) {
        Ray r = camera_ray(s.cam, V2{pos[2 * i], pos[2 * i + 1]}, s.eps);
        float vals[8] = {r.o.x, r.o.y, r.o.z, r.d.x, r.d.y, r.d.z, r.mint, r.maxt};
        for (int k = 0; k < 8; k++) rays[(size_t)k * n + i] = vals[k];
    }
}

// Per-path radiance for (x, y, sample) triples: out = (Px, Py, r, g, b).
void dlo_render_samples(void* h, const int* xys, uint32_t n, uint32_t seed, int max_depth, int rr_depth, int accel,
                        float* out) {
    const Scene& s = *(Scene*)h;
    RenderOpts o;
    o.max_depth = max_depth;
    o.rr_depth = rr_depth;
    o.accel = accel;
    uint64_t cnt[2] = {0, 0};
    for (uint32_t i = 0; i < n; i++) {
        V2 pos;
        V3 L = render_sample(s, o, xys[3 * i], xys[3 * i + 1], (uint32_t)xys[3 * i + 2], seed, pos, cnt);
        float* o5 = out + 5 * (size_t)i;
        o5[0] = pos.x;
        o5[1] = pos.y;
        o5[2] = L.x;
        o5[3] = L.y;
        o5[4] = L.z;
    }
}

// CPURenderer::renderAsync  cpu_renderer.cpp:57-135: 32x32 blocks in spiral
// order on a thread pool (an atomic block counter stands in for
// tbb::parallel_for), merged into the image block and normalised.
// max_blocks > 0 renders only the first max_blocks blocks of the spiral.
// stats: [0] paths, [1] closest queries, [2] shadow queries, [3] seconds.
int dlo_render(void* h, int spp, uint32_t seed, int max_depth, int rr_depth, int accel, int threads, int max_blocks,
               float* rgb_out, double* stats) {
    Scene& s = *(Scene*)h;
    if (!s.built && accel != DLO_ACCEL_BRUTE) return -1;
    if (s.emitters.empty()) return -2;  // cpu_renderer.cpp:63-66
    auto job = std::make_unique<RenderJob>();
    job->s = &s;
    job->o.max_depth = max_depth;
    job->o.rr_depth = rr_depth;
    job->o.accel = accel;
    job->spp = spp;
    job->seed = seed;
    job->blocks = spiral_blocks(s.cam.width, s.cam.height, 32);
    if (max_blocks > 0 && max_blocks < (int)job->blocks.size()) job->blocks.resize(max_blocks);
    job->results.resize(job->blocks.size());
    if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
    if (threads <= 0) threads = 1;
    job->counters.assign(2 * (size_t)threads, 0);
    auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (int i = 0; i < threads; i++) pool.emplace_back(render_worker, job.get(), i);
    for (auto& th : pool) th.join();
    auto t1 = std::chrono::steady_clock::now();
    // ImageBlock::put(ImageBlock&) merge, then toBitmap  block.cpp:58-66,115-129
    const int W = s.cam.width, H = s.cam.height, b = s.filter.border;
    std::vector<float> img((size_t)(W + 2 * b) * (H + 2 * b) * 4, 0.0f);
    uint64_t paths = 0;
    for (auto& blk : job->results) {
        paths += (uint64_t)blk.sx * blk.sy * spp;
        for (int i = 0; i < blk.rows(); i++)
            for (int j = 0; j < blk.cols(); j++) {
                int y = blk.oy + i, x = blk.ox + j;  // offset shift cancels with equal borders
                float* d = &img[4 * ((size_t)y * (W + 2 * b) + x)];
                const float* sv = &blk.data[4 * ((size_t)i * blk.cols() + j)];
                for (int k = 0; k < 4; k++) d[k] += sv[k];
            }
    }
    if (rgb_out) {
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const float* c = &img[4 * ((size_t)(y + b) * (W + 2 * b) + (x + b))];
                float* o = rgb_out + 3 * ((size_t)y * W + x);
                if (c[3] != 0) {
                    o[0] = c[0] / c[3];
                    o[1] = c[1] / c[3];
                    o[2] = c[2] / c[3];
                } else {
                    o[0] = o[1] = o[2] = 0.0f;
                }
            }
    }
    if (stats) {
        uint64_t cq = 0, sq = 0;
        for (int i = 0; i < threads; i++) {
            cq += job->counters[2 * i];
            sq += job->counters[2 * i + 1];
        }
        stats[0] = (double)paths;
        stats[1] = (double)cq;
        stats[2] = (double)sq;
        stats[3] = std::chrono::duration<double>(t1 - t0).count();
        stats[4] = (double)threads;
    }
    return 0;
}

// --- unit-level entry points used by the known-answer tests --------------
void dlo_pcg32(uint64_t initstate, uint64_t initseq, uint32_t n, uint32_t* out) {
    Pcg32 r;
    r.seed(initstate, initseq);
    for (uint32_t i = 0; i < n; i++) out[i] = r.next_uint();
}
void dlo_pcg32_float(uint64_t initstate, uint64_t initseq, uint32_t n, float* out) {
    Pcg32 r;
    r.seed(initstate, initseq);
    for (uint32_t i = 0; i < n; i++) out[i] = r.next_float();
}
// which: 0 disk(2) 1 cosine hemisphere(3) 2 beckmann(3, alpha) 3 triangle(3)
//        4 beckmann pdf(1, alpha) 5 cosine pdf(1) 6 ggx sample(3, ax, ay)
//        7 ggx D(1, ax, ay)
void dlo_warp(int which, const float* in, uint32_t n, float a0, float a1, float* out) {
    for (uint32_t i = 0; i < n; i++) {
        V2 s{in[3 * i], in[3 * i + 1]};
        V3 v3{in[3 * i], in[3 * i + 1], in[3 * i + 2]};
        float* o = out + 3 * (size_t)i;
        V3 r{0, 0, 0};
        switch (which) {
            case 0: {
                V2 d = sq_disk(s);
                r = mk(d.x, d.y, 0);
            } break;
            case 1: r = sq_cos_hemi(s); break;
            case 2: r = sq_beckmann(s, a0); break;
            case 3: r = sq_triangle(s); break;
            case 4: r.x = sq_beckmann_pdf(v3, a0); break;
            case 5: r.x = sq_cos_hemi_pdf(v3); break;
            case 6: r = ggx_sample(s, a0, a1); break;
            case 7: r.x = ggx_d(v3, a0, a1); break;
        }
        o[0] = r.x;
        o[1] = r.y;
        o[2] = r.z;
    }
}
// which: 0 fresnel(cos, ext, int) 1 refract(wi, eta) 2 reflect(wi, n)
//        3 conductor fresnel(cos, eta, k)
void dlo_optics(int which, const float* in, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; i++) {
        const float* a = in + 6 * (size_t)i;
        float* o = out + 3 * (size_t)i;
        V3 r{0, 0, 0};
        if (which == 0)
            r.x = fresnel(a[0], a[1], a[2]);
        else if (which == 1)
            r = refract(mk(a[0], a[1], a[2]), a[3]);
        else if (which == 2)
            r = reflect(mk(a[0], a[1], a[2]), mk(a[3], a[4], a[5]));
        else
            r.x = fresnel_conductor(a[0], a[1], a[2]);
        o[0] = r.x;
        o[1] = r.y;
        o[2] = r.z;
    }
}
// BSDF queries in the local frame.  in: wi(3) wo(3) s(2); out: eval(3)
// pdf(1) sample(3) wo'(3) eta(1) measure(1) = 12 floats.
void dlo_bsdf_query(void* h, int bsdf, const float* in, uint32_t n, float* out) {
    const Scene& s = *(Scene*)h;
    const Bsdf& b = s.bsdfs[bsdf];
    for (uint32_t i = 0; i < n; i++) {
        const float* a = in + 8 * (size_t)i;
        float* o = out + 12 * (size_t)i;
        BRec q;
        q.wi = mk(a[0], a[1], a[2]);
        q.wo = mk(a[3], a[4], a[5]);
        q.measure = kSolidAngle;
        V3 e = bsdf_eval(b, q);
        float p = bsdf_pdf(b, q);
        BRec r;
        r.wi = q.wi;
        V3 f = bsdf_sample(b, r, V2{a[6], a[7]});
        float vals[12] = {e.x, e.y, e.z, p, f.x, f.y, f.z, r.wo.x, r.wo.y, r.wo.z, r.eta, (float)r.measure};
        memcpy(o, vals, sizeof(vals));
    }
}
// Emitter sampling at shading points (tests of the emitters): the scene's
// emitter pick, the emitter's sample, and the pdf of the sampled direction
// as the emitter-hit MIS evaluates it.  The environment map's third uniform
// is in[8].
void dlo_emitter_query(void* h, const float* in, uint32_t n, float* out) {
    const Scene& s = *(Scene*)h;
    for (uint32_t i = 0; i < n; i++) {
        const float* a = in + 9 * (size_t)i;
        float* o = out + 11 * (size_t)i;
        Its its;
        its.p = mk(a[0], a[1], a[2]);
        its.sh = frame_from_normal(mk(a[3], a[4], a[5]));
        its.geo = its.sh;
        its.mesh = -1;
        its.tri = -1;
        FixedSampler fs{a[8]};
        DirSample ds;
        ds.d = ds.n = ds.o = mk(0, 0, 0);
        V3 spec = mk(0, 0, 0);
        size_t ne = s.emitters.size();
        if (ne > 0) {
            V2 smp{a[6], a[7]};
            int index = 0;
            float scale = 1.0f;
            if (ne > 1) {
                index = (int)std::min((size_t)(smp.x * (float)ne), ne - 1);
                smp.x = smp.x * (float)ne - (float)index;
                scale = (float)ne;
            }
            spec = emitter_sample(s, s.emitters[index], smp, its, ds, fs);
            if (ne > 1) {
                spec = spec * scale;
                ds.pdf *= 1.f / (float)ne;
            }
            ds.emitter = index;
        }
        float pdir = ds.emitter >= 0 && !ds.delta ? pdf_emitter_direction(s, ds.emitter, ds) : 0.0f;
        float vals[11] = {spec.x, spec.y, spec.z, ds.d.x, ds.d.y, ds.d.z, ds.dist, ds.pdf, pdir, (float)ds.emitter,
                          ds.delta ? 1.0f : 0.0f};
        memcpy(o, vals, sizeof(vals));
    }
}
void dlo_set_mutation(int m) { g_mut = m; }
// Environment map internals: q, alias, pdf tables (n = w*h).
void dlo_env_tables(void* h, float* q, uint32_t* alias, float* pdf) {
    const Scene& s = *(Scene*)h;
    if (s.env < 0) return;
    const Emitter& e = s.emitters[s.env];
    memcpy(q, e.q.data(), sizeof(float) * e.q.size());
    memcpy(alias, e.alias.data(), sizeof(uint32_t) * e.alias.size());
    memcpy(pdf, e.pdf.data(), sizeof(float) * e.pdf.size());
}
// Block spiral order (pairs of offsets) for a W x H image.
uint32_t dlo_spiral_blocks(int W, int H, int* out) {
    auto v = spiral_blocks(W, H, 32);
    if (out)
        for (size_t i = 0; i < v.size(); i++) {
            out[2 * i] = v[i].first;
            out[2 * i + 1] = v[i].second;
        }
    return (uint32_t)v.size();
}

}  // extern "C"
