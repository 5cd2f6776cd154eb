// Device shading: intersection records, BSDF sample/eval/pdf, emitter
// sampling.  Written for the wavefront kernels in integrator.hip; every
// function reproduces the reference CPU path's binary32 arithmetic.
#pragma once
#include "backend.h"
#include "dl_math.h"

namespace dl {

struct SceneView {
    const float* __restrict__ pos;
    const float* __restrict__ nrm;
    const float* __restrict__ tan;
    const uint32_t* __restrict__ idx;
    const uint32_t* __restrict__ prim_mesh;
    const uint8_t* __restrict__ prim_info;  // material type | kPrimLight
    const MeshRec* __restrict__ meshes;
    const dl_material_desc* __restrict__ mats;
    const LightRec* __restrict__ lights;
    const float* __restrict__ light_cdf;
    const uint32_t* __restrict__ light_guide;
    const float4* __restrict__ env_rgb;
    const float* __restrict__ env_q;
    const uint32_t* __restrict__ env_alias;
    const float* __restrict__ env_pdf;
    const float* __restrict__ uv;        // texture coordinates, 2 per vertex
    const float4* __restrict__ texels;   // bitmap textures
    const int4* __restrict__ tex_info;   // per texture: first texel, width, height
    int num_lights;
    int env_light;
    float bg[3];
    // a layout holding several scenes: per-scene records and each path's
    // scene index (nullptr: one scene, described by the fields above)
    const SceneRec* __restrict__ scenes;
    const uint8_t* __restrict__ sid;
};

// The record of the scene path p belongs to.
__device__ __forceinline__ SceneRec path_scene(const SceneView& s, uint32_t p) {
    if (s.sid) return s.scenes[s.sid[p]];
    SceneRec r;
    r.light0 = 0;
    r.num_lights = s.num_lights;
    r.env_light = s.env_light;
    r.root = 0;
    r.bg[0] = s.bg[0];
    r.bg[1] = s.bg[1];
    r.bg[2] = s.bg[2];
    r.pad = 0;
    return r;
}

__device__ __forceinline__ vec3 ld3(const float* p, uint32_t i) { return v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }

struct Its {
    vec3 p;
    Frame sh;
    int mesh;
};

// OCTree::rayIntersect record construction (src/tracer/accel/octree.cpp:166-209);
// the anisotropic GGX frame follows src/optix/cuda/shader_common.h:121-133 and
// include/optix/math/onb.h:11-33.
__device__ __forceinline__ void make_its(const SceneView& s, uint32_t prim, float u, float v, bool want_tangent,
                                         Its& its) {
    const uint32_t m = s.prim_mesh[prim];
    const MeshRec& mr = s.meshes[m];
    const uint32_t i0 = s.idx[3 * prim], i1 = s.idx[3 * prim + 1], i2 = s.idx[3 * prim + 2];
    const float b0 = (1.0f - u) - v, b1 = u, b2 = v;
    const vec3 p0 = ld3(s.pos, i0), p1 = ld3(s.pos, i1), p2 = ld3(s.pos, i2);
    its.mesh = (int)m;
    its.p = add(add(scale(p0, b0), scale(p1, b1)), scale(p2, b2));
    if (mr.has_normals) {
        const vec3 n = add(add(scale(ld3(s.nrm, i0), b0), scale(ld3(s.nrm, i1), b1)), scale(normalized(ld3(s.nrm, i2)), b2));
        its.sh = frame_from_normal(n);
    } else {
        its.sh = frame_from_normal(normalized(cross(sub(p1, p0), sub(p2, p0))));
    }
    if (want_tangent && mr.has_tangents) {
        const vec3 t0 = ld3(s.tan, i0);
        if (!(isnan(t0.x) || isnan(t0.y) || isnan(t0.z))) {
            const vec3 st = normalized(add(add(scale(t0, b0), scale(ld3(s.tan, i1), b1)), scale(ld3(s.tan, i2), b2)));
            its.sh.t = st;
            its.sh.b = normalized(cross(its.sh.n, st));
        }
    }
}

// ---- bitmap textures (reference GPU path) ---------------------------------
// Texture coordinates of a hit as the reference's GPU path interpolates them
// (src/optix/cuda/shader_common.h:36-45): zero for a mesh without them.
__device__ __forceinline__ float2 hit_texcoord(const SceneView& s, uint32_t prim, float u, float v) {
    const MeshRec& mr = s.meshes[s.prim_mesh[prim]];
    if (!mr.has_texcoords) return make_float2(0.0f, 0.0f);
    const uint32_t i0 = s.idx[3 * prim], i1 = s.idx[3 * prim + 1], i2 = s.idx[3 * prim + 2];
    const float b0 = (1.0f - u) - v;
    return make_float2((b0 * s.uv[2 * i0] + u * s.uv[2 * i1]) + v * s.uv[2 * i2],
                       (b0 * s.uv[2 * i0 + 1] + u * s.uv[2 * i1 + 1]) + v * s.uv[2 * i2 + 1]);
}

// Integer texel coordinate f (a float holding an integer) wrapped into
// [0, n) -- cudaAddressModeWrap; fmod is exact.
__device__ __forceinline__ int wrap_texel(float f, int n) {
    float r = fmodf(f, (float)n);
    if (r < 0.0f) r += (float)n;
    const int i = (int)r;
    return i < n ? i : n - 1;
}

// Bilinear lookup of texture t (0-based) at normalized coordinates (u, v):
// texel centres at ((i + 1/2) / w, (j + 1/2) / h), wrap addressing
// (src/optix/host/cuda_texture.cpp:57-82), binary32 weights.
__device__ __forceinline__ vec3 tex_sample(const SceneView& s, int t, float u, float v) {
    const int4 ti = s.tex_info[t];
    const int w = ti.y, h = ti.z;
    if (!isfinite(u) || !isfinite(v)) return v3(0.0f, 0.0f, 0.0f);
    const float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float ax = x - fx, ay = y - fy;
    const int x0 = wrap_texel(fx, w), y0 = wrap_texel(fy, h);
    const int x1 = x0 + 1 < w ? x0 + 1 : 0, y1 = y0 + 1 < h ? y0 + 1 : 0;
    const float4* T = s.texels + ti.x;
    const float4 t00 = T[y0 * w + x0], t10 = T[y0 * w + x1], t01 = T[y1 * w + x0], t11 = T[y1 * w + x1];
    const float bx = 1.0f - ax, by = 1.0f - ay;
    return v3((t00.x * bx + t10.x * ax) * by + (t01.x * bx + t11.x * ax) * ay,
              (t00.y * bx + t10.y * ax) * by + (t01.y * bx + t11.y * ax) * ay,
              (t00.z * bx + t10.z * ax) * by + (t01.z * bx + t11.z * ax) * ay);
}

__device__ __forceinline__ bool has_textures(const dl_material_desc& m) {
    return (m.tex[0] | m.tex[1] | m.tex[2] | m.tex[3] | m.tex[4]) != 0;
}

// The first 8 material parameters at a surface point: textured slots
// replaced by their texture values (diffuse.cu:16-19 albedo;
// aniso_ggx.cu:17-36 pd, ps, axay).
__device__ __forceinline__ void resolve_params(const SceneView& s, const dl_material_desc& m, float2 tc, float* p) {
    for (int k = 0; k < 8; k++) p[k] = m.params[k];
    if (m.type == DL_BSDF_DIFFUSE && m.tex[DL_TEX_ALBEDO]) {
        const vec3 a = tex_sample(s, m.tex[DL_TEX_ALBEDO] - 1, tc.x, tc.y);
        p[0] = a.x;
        p[1] = a.y;
        p[2] = a.z;
    }
    if (m.type == DL_BSDF_ANISO_GGX) {
        if (m.tex[DL_TEX_PD]) {
            const vec3 a = tex_sample(s, m.tex[DL_TEX_PD] - 1, tc.x, tc.y);
            p[0] = a.x;
            p[1] = a.y;
            p[2] = a.z;
        }
        if (m.tex[DL_TEX_PS]) {
            const vec3 a = tex_sample(s, m.tex[DL_TEX_PS] - 1, tc.x, tc.y);
            p[3] = a.x;
            p[4] = a.y;
            p[5] = a.z;
        }
        if (m.tex[DL_TEX_AXAY]) {
            const vec3 a = tex_sample(s, m.tex[DL_TEX_AXAY] - 1, tc.x, tc.y);
            p[6] = a.x;
            p[7] = a.y;
        }
    }
}

// Onb(normal, tangent) of the reference GPU path (include/optix/math/onb.h:11-33).
__device__ __forceinline__ Frame onb_frame(vec3 n, vec3 t) {
    Frame f;
    f.n = n;
    if (t.x == 0.0f && t.y == 0.0f && t.z == 0.0f) {
        vec3 b = fabsf(n.x) > fabsf(n.z) ? v3(-n.y, n.x, 0.0f) : v3(0.0f, -n.z, n.y);
        f.b = normalized(b);
        f.t = cross(f.b, n);
    } else {
        f.t = t;
        f.b = normalized(cross(n, t));
    }
    return f;
}

__device__ __forceinline__ vec3 onb_to_world(const Frame& f, vec3 p) {
    return add(add(scale(f.t, p.x), scale(f.b, p.y)), scale(f.n, p.z));
}

// Shading frame of an aniso_ggx point with a normal and/or tangent map
// (src/optix/cuda/shader_common.h:46-147): the map value * 2 - 1, in the
// frame of the interpolated normal and tangent (tangent space, the default)
// or as is (object space); without a normal map the interpolated vertex
// normal, without either the geometric normal.
__device__ __forceinline__ Frame map_frame(const SceneView& s, uint32_t prim, float u, float v,
                                           const dl_material_desc& m, float2 tc) {
    const MeshRec& mr = s.meshes[s.prim_mesh[prim]];
    const uint32_t i0 = s.idx[3 * prim], i1 = s.idx[3 * prim + 1], i2 = s.idx[3 * prim + 2];
    const float b0 = (1.0f - u) - v;
    const vec3 p0 = ld3(s.pos, i0), p1 = ld3(s.pos, i1), p2 = ld3(s.pos, i2);
    const vec3 gn = normalized(cross(sub(p1, p0), sub(p2, p0)));
    const bool frame_ok = mr.has_tangents && mr.has_normals;
    Frame tbn{};
    if (frame_ok) {
        const vec3 n = normalized(add(add(scale(ld3(s.nrm, i0), b0), scale(ld3(s.nrm, i1), u)), scale(ld3(s.nrm, i2), v)));
        const vec3 t = normalized(add(add(scale(ld3(s.tan, i0), b0), scale(ld3(s.tan, i1), u)), scale(ld3(s.tan, i2), v)));
        tbn = onb_frame(n, t);
    }
    vec3 sn = gn;
    if (m.tex[DL_TEX_NORMAL]) {
        const vec3 c = tex_sample(s, m.tex[DL_TEX_NORMAL] - 1, tc.x, tc.y);
        const vec3 l = v3(c.x * 2.0f - 1.0f, c.y * 2.0f - 1.0f, c.z * 2.0f - 1.0f);
        if (m.maps_object_space)
            sn = l;
        else if (frame_ok)
            sn = onb_to_world(tbn, l);
    } else if (mr.has_normals) {
        sn = add(add(scale(ld3(s.nrm, i0), b0), scale(ld3(s.nrm, i1), u)), scale(ld3(s.nrm, i2), v));
    }
    sn = normalized(sn);
    vec3 st = v3(0.0f, 0.0f, 0.0f);
    if (m.tex[DL_TEX_TANGENT]) {
        const vec3 c = tex_sample(s, m.tex[DL_TEX_TANGENT] - 1, tc.x, tc.y);
        const vec3 l = v3(c.x * 2.0f - 1.0f, c.y * 2.0f - 1.0f, c.z * 2.0f - 1.0f);
        if (m.maps_object_space)
            st = l;
        else if (frame_ok)
            st = onb_to_world(tbn, l);
        // a tangent-space map on a mesh without tangents leaves no tangent
        // (the reference normalises the zero vector to NaN)
        if (st.x != 0.0f || st.y != 0.0f || st.z != 0.0f) st = normalized(st);
    } else if (mr.has_tangents) {
        const vec3 t0 = ld3(s.tan, i0);
        if (!(isnan(t0.x) || isnan(t0.y) || isnan(t0.z)))
            st = normalized(add(add(scale(t0, b0), scale(ld3(s.tan, i1), u)), scale(ld3(s.tan, i2), v)));
    }
    return onb_frame(sn, st);
}

// Hit record and the first 8 material parameters of a shading point: the
// reference CPU record (make_its); for a textured material the texture
// values, and for aniso_ggx with normal / tangent maps the mapped frame.
__device__ __forceinline__ void shading_point(const SceneView& s, uint32_t prim, float u, float v, int type, Its& its,
                                              int& mat, float (&p)[8]) {
    make_its(s, prim, u, v, type == DL_BSDF_ANISO_GGX, its);
    mat = s.meshes[its.mesh].material;
    const dl_material_desc& md = s.mats[mat];
    float2 tc = make_float2(0.0f, 0.0f);
    if (has_textures(md)) {
        tc = hit_texcoord(s, prim, u, v);
        if (type == DL_BSDF_ANISO_GGX && (md.tex[DL_TEX_NORMAL] | md.tex[DL_TEX_TANGENT]) != 0)
            its.sh = map_frame(s, prim, u, v, md, tc);
    }
    resolve_params(s, md, tc, p);
}

// Shading normal only (emitter hits need its.p and shFrame.n).
__device__ __forceinline__ void hit_point_normal(const SceneView& s, uint32_t prim, float u, float v, vec3& p, vec3& n) {
    const uint32_t m = s.prim_mesh[prim];
    const MeshRec& mr = s.meshes[m];
    const uint32_t i0 = s.idx[3 * prim], i1 = s.idx[3 * prim + 1], i2 = s.idx[3 * prim + 2];
    const float b0 = (1.0f - u) - v, b1 = u, b2 = v;
    const vec3 p0 = ld3(s.pos, i0), p1 = ld3(s.pos, i1), p2 = ld3(s.pos, i2);
    p = add(add(scale(p0, b0), scale(p1, b1)), scale(p2, b2));
    if (mr.has_normals)
        n = add(add(scale(ld3(s.nrm, i0), b0), scale(ld3(s.nrm, i1), b1)), scale(normalized(ld3(s.nrm, i2)), b2));
    else
        n = normalized(cross(sub(p1, p0), sub(p2, p0)));
}

// ---- warps (src/core/math/wrap.cpp), binary64 transcendentals ------------
__device__ __forceinline__ float dsqrtf(float x) { return (float)sqrt((double)x); }

__device__ __forceinline__ vec3 sq_cos_hemi(float sx, float sy) {
    const float r = dsqrtf(sx);
    const float th = 2 * kPi * sy;
    const float px = (float)((double)r * cos((double)th));
    const float py = (float)((double)r * sin((double)th));
    float sum = 0.0f;
    sum = sum + px * px;
    sum = sum + py * py;
    return v3(px, py, dsqrtf(1 - sum));
}
__device__ __forceinline__ float sq_cos_hemi_pdf(vec3 v) { return v.z <= 0 ? 0.0f : v.z * kInvPi; }
__device__ __forceinline__ vec3 sq_beckmann(float sx, float sy, float alpha) {
    const float theta = (float)atan(sqrt((double)(-alpha * alpha) * log((double)(1 - sx))));
    const float phi = 2 * kPi * sy;
    const double st = sin((double)theta);
    return v3((float)(st * cos((double)phi)), (float)(st * sin((double)phi)), (float)cos((double)theta));
}
__device__ __forceinline__ float sq_beckmann_pdf(vec3 m, float alpha) {
    if (m.z <= 0) return 0;
    const float c = m.z;
    const float tan2 = (1 - c * c) / (c * c);
    const double e = exp((double)(-tan2 / (alpha * alpha)));
    const float den = alpha * alpha * c * c * c;
    return (float)((double)kInvTwoPi * (2 * e) / (double)den);
}
__device__ __forceinline__ vec3 sq_triangle(float sx, float sy) {
    const float t = dsqrtf(1.f - sx);
    return v3(1 - t, t * sy, t - t * sy);
}

// ---- optics (src/core/math/math.cpp:8-51) --------------------------------
__device__ __forceinline__ float fresnel(float c, float extIOR, float intIOR) {
    float etaI = extIOR, etaT = intIOR;
    if (extIOR == intIOR) return 0.0f;
    if (c < 0.0f) {
        const float tmp = etaI;
        etaI = etaT;
        etaT = tmp;
        c = -c;
    }
    const float eta = etaI / etaT, s2 = eta * eta * (1 - c * c);
    if (s2 > 1.0f) return 1.0f;
    const float ct = dsqrtf(1.0f - s2);
    const float Rs = (etaI * c - etaT * ct) / (etaI * c + etaT * ct);
    const float Rp = (etaT * c - etaI * ct) / (etaT * c + etaI * ct);
    return (Rs * Rs + Rp * Rp) / 2.0f;
}
__device__ __forceinline__ vec3 refract(vec3 wi, float eta) {
    const float c = wi.z;
    const bool outside = c > 0.f;
    eta = outside ? eta : 1 / eta;
    const float ct2 = 1 - eta * eta * (1 - c * c);
    float ct = dsqrtf(ct2);
    ct = outside ? -ct : ct;
    return v3(-eta * wi.x, -eta * wi.y, ct);
}
__device__ __forceinline__ vec3 reflect(vec3 wi, vec3 n) { return sub(scale(n, 2 * dot(wi, n)), wi); }

// ---- BSDFs ----------------------------------------------------------------
enum { kMeasureUnknown = 0, kMeasureSolidAngle = 1, kMeasureDiscrete = 2 };

struct BRec {
    vec3 wi, wo;
    float eta;
    int measure;
};

__device__ __forceinline__ vec3 pcol(const float* p) { return v3(p[0], p[1], p[2]); }

__device__ __forceinline__ float mf_g1(vec3 wv, vec3 wh, float alpha) {  // microfacet.cpp:35-44
    const float c = dot(wv, wh) / wv.z;
    if (c <= 0) return 0.0f;
    const float tmp = 1 - wv.z * wv.z;
    const float tanT = tmp <= 0.0f ? 0.0f : dsqrtf(tmp) / wv.z;
    const float b = (float)(1.0 / (double)(alpha * tanT));
    const double bd = b;
    return bd < 1.6 ? (float)((3.535 * bd + 2.181 * bd * bd) / (1 + 2.276 * bd + 2.577 * bd * bd)) : 1.0f;
}

__device__ __forceinline__ float ggx_d(vec3 m, float ax, float ay) {  // include/optix/math/wrap.h:54-67
    if (m.z <= 0) return 0;
    const vec3 vh = v3(m.x / ax, m.y / ay, m.z);
    const float len2 = dot(vh, vh);
    return 1.f / (kPi * ax * ay * len2 * len2);
}
__device__ __forceinline__ vec3 ggx_sample(float sx, float sy, float ax, float ay) {  // wrap.h:70-77
    const double sz = sqrt((double)sx), sw = sqrt((double)(1 - sx));
    const float ang = 2 * kPi * sy;
    const float x = (float)((double)ax * sz / sw * cos((double)ang));
    const float y = (float)((double)ay * sz / sw * sin((double)ang));
    return normalized(v3(-x, -y, 1));
}
__device__ __forceinline__ float ggx_g1(vec3 v, float ax, float ay) {  // aniso_ggx.cu:9-15
    if (v.z <= 0.0f) return 0.0f;
    const vec3 vv = v3(v.x * ax, v.y * ay, v.z);
    return (float)(2.0 * (double)v.z / (double)(v.z + length(vv)));
}
__device__ __forceinline__ float schlick(float ldoth) {  // aniso_ggx.cu:53-58
    const float F0 = 0.04f;
    const double tmp = 1.0 - (double)ldoth;
    const float temp = fmin0(fmax0(0.0f, (float)tmp), 1.0f);  // std::min(std::max(0, x), 1)
    const double t = temp;
    return (float)((double)F0 + (1.0 - (double)F0) * t * t * t * t * t);
}
__device__ __forceinline__ float fresnel_conductor(float cosi, float eta, float k) {
    const float c2 = cosi * cosi;
    const float s2 = 1.0f - c2;
    const float e2 = eta * eta, k2 = k * k;
    const float t0 = e2 - k2 - s2;
    const float a2b2 = dsqrtf(t0 * t0 + 4.0f * e2 * k2);
    const float t1 = a2b2 + c2;
    const float a = dsqrtf(0.5f * (a2b2 + t0));
    const float t2 = 2.0f * cosi * a;
    const float rs = (t1 - t2) / (t1 + t2);
    const float t3 = c2 * a2b2 + s2 * s2;
    const float t4 = t2 * s2;
    const float rp = rs * (t3 - t4) / (t3 + t4);
    return 0.5f * (rp + rs);
}

__device__ __forceinline__ bool bsdf_is_diffuse(int type) {
    return type == DL_BSDF_DIFFUSE || type == DL_BSDF_MICROFACET || type == DL_BSDF_ANISO_GGX ||
           type == DL_BSDF_CONDUCTOR;
}

__device__ __forceinline__ float microfacet_ks(const float* p) { return 1 - cmax(pcol(p + 3)); }

__device__ inline vec3 bsdf_eval(int type, const float* p, const BRec& r) {
    switch (type) {
        case DL_BSDF_DIFFUSE:
            if (r.measure != kMeasureSolidAngle || r.wi.z <= 0 || r.wo.z <= 0) return v3(0, 0, 0);
            return scale(pcol(p), kInvPi);
        case DL_BSDF_MICROFACET: {
            const float alpha = p[0];
            const vec3 wh = normalized(add(r.wi, r.wo));
            const float D = sq_beckmann_pdf(wh, alpha);
            const float F = fresnel(dot(r.wi, wh), p[2], p[1]);
            const float G = mf_g1(r.wi, wh, alpha) * mf_g1(r.wo, wh, alpha);
            const float deno = 1 / (4 * r.wi.z * r.wo.z);
            return adds(scale(pcol(p + 3), kInvPi), microfacet_ks(p) * D * F * G * deno);
        }
        case DL_BSDF_ANISO_GGX: {
            if (r.wi.z <= 0 || r.wo.z <= 0) return v3(0, 0, 0);
            const float ax = p[6], ay = p[7];
            const vec3 wh = normalized(add(r.wi, r.wo));
            const float D = ggx_d(wh, ax, ay);
            const float F = schlick(dot(r.wi, wh));
            const float G = ggx_g1(r.wi, ax, ay) * ggx_g1(r.wo, ax, ay);
            const float deno = 1.f / (4.f * r.wi.z * r.wo.z);
            return add(scale(pcol(p), kInvPi), scale(scale(scale(scale(pcol(p + 3), D), F), G), deno));
        }
        case DL_BSDF_CONDUCTOR: {
            if (r.wi.z <= 0 || r.wo.z <= 0) return v3(0, 0, 0);
            const float a = p[0];
            const vec3 wh = normalized(add(r.wi, r.wo));
            const float D = ggx_d(wh, a, a);
            const float G = ggx_g1(r.wi, a, a) * ggx_g1(r.wo, a, a);
            const float ci = dot(r.wi, wh);
            const float deno = 1.f / (4.f * r.wi.z * r.wo.z);
            const float sc = (D * G) * deno;
            return v3(fresnel_conductor(ci, p[1], p[4]) * sc, fresnel_conductor(ci, p[2], p[5]) * sc,
                      fresnel_conductor(ci, p[3], p[6]) * sc);
        }
        default:
            return v3(0, 0, 0);
    }
}

__device__ inline float bsdf_pdf(int type, const float* p, const BRec& r) {
    switch (type) {
        case DL_BSDF_DIFFUSE:
            if (r.measure != kMeasureSolidAngle || r.wi.z <= 0 || r.wo.z <= 0) return 0.0f;
            return kInvPi * r.wo.z;
        case DL_BSDF_MICROFACET: {
            if (r.measure != kMeasureSolidAngle || r.wi.z <= 0 || r.wo.z <= 0) return 0.0f;
            const float ks = microfacet_ks(p);
            const vec3 wh = normalized(add(r.wi, r.wo));
            const float diff = (1 - ks) * sq_cos_hemi_pdf(r.wo);
            const float spec = ks * sq_beckmann_pdf(wh, p[0]) / 4 / dot(wh, r.wo);
            return diff + spec;
        }
        case DL_BSDF_ANISO_GGX: {
            if (r.measure != kMeasureSolidAngle || r.wi.z <= 0 || r.wo.z <= 0) return 0.f;
            const vec3 wh = normalized(add(r.wi, r.wo));
            const float diff = (1 - 0.5f) * sq_cos_hemi_pdf(r.wo);
            const float spec = 0.5f * ggx_d(wh, p[6], p[7]) * wh.z / 4.f / dot(wh, r.wo);
            return diff + spec;
        }
        case DL_BSDF_CONDUCTOR: {
            if (r.measure != kMeasureSolidAngle || r.wi.z <= 0 || r.wo.z <= 0) return 0.f;
            const vec3 wh = normalized(add(r.wi, r.wo));
            return ggx_d(wh, p[0], p[0]) * wh.z / 4.f / dot(wh, r.wo);
        }
        default:
            return 0.0f;
    }
}

__device__ inline vec3 bsdf_sample(int type, const float* p, BRec& r, float sx, float sy) {
    switch (type) {
        case DL_BSDF_DIFFUSE:
            if (r.wi.z <= 0) return v3(0, 0, 0);
            r.measure = kMeasureSolidAngle;
            r.wo = sq_cos_hemi(sx, sy);
            r.eta = 1.0f;
            return pcol(p);
        case DL_BSDF_MIRROR:
            if (r.wi.z <= 0) return v3(0, 0, 0);
            r.wo = v3(-r.wi.x, -r.wi.y, r.wi.z);
            r.measure = kMeasureDiscrete;
            r.eta = 1.0f;
            return v3(1, 1, 1);
        case DL_BSDF_DIELECTRIC: {
            r.measure = kMeasureDiscrete;
            const float c = r.wi.z;
            const float ri = fresnel(c, p[1], p[0]);
            if (sx < ri) {
                r.wo = v3(-r.wi.x, -r.wi.y, r.wi.z);
                r.eta = 1;
            } else {
                const float eta = p[1] / p[0];
                r.wo = normalized(refract(r.wi, eta));
                r.eta = c > 0.f ? eta : 1 / eta;
            }
            return v3(1.f, 1.f, 1.f);
        }
        case DL_BSDF_MICROFACET: {
            if (r.wi.z <= 0) return v3(0, 0, 0);
            r.measure = kMeasureSolidAngle;
            const float ks = microfacet_ks(p);
            if (sx > ks) {
                r.wo = sq_cos_hemi((sx - ks) / (1 - ks), sy);
            } else {
                const vec3 wh = sq_beckmann(sx / ks, sy, p[0]);
                r.wo = reflect(r.wi, wh);
            }
            if (r.wo.z <= 0) return v3(0, 0, 0);
            return divs(scale(bsdf_eval(type, p, r), r.wo.z), bsdf_pdf(type, p, r));
        }
        case DL_BSDF_ANISO_GGX: {
            if (r.wi.z <= 0) return v3(0, 0, 0);
            r.measure = kMeasureSolidAngle;
            if (sx > 0.5f) {
                r.wo = sq_cos_hemi((sx - 0.5f) / (1 - 0.5f), sy);
            } else {
                const vec3 wh = ggx_sample(sx / 0.5f, sy, p[6], p[7]);
                r.wo = reflect(r.wi, wh);
            }
            if (r.wo.z <= 0) return v3(0, 0, 0);
            return divs(scale(bsdf_eval(type, p, r), r.wo.z), bsdf_pdf(type, p, r));
        }
        case DL_BSDF_CONDUCTOR: {
            if (r.wi.z <= 0) return v3(0, 0, 0);
            r.measure = kMeasureSolidAngle;
            const vec3 wh = ggx_sample(sx, sy, p[0], p[0]);
            r.wo = reflect(r.wi, wh);
            if (r.wo.z <= 0) return v3(0, 0, 0);
            return divs(scale(bsdf_eval(type, p, r), r.wo.z), bsdf_pdf(type, p, r));
        }
    }
    return v3(0, 0, 0);
}

// ---- emitters -----------------------------------------------------------
struct DirSample {
    vec3 d, n;
    float dist, pdf;
    bool delta;
};

__device__ __forceinline__ DirSample dir_sample(vec3 o, vec3 dst, vec3 n) {  // mesh.h:89-104
    DirSample ds;
    ds.n = n;
    const vec3 vec = sub(dst, o);
    ds.d = normalized(vec);
    ds.dist = length(vec);
    ds.pdf = 0.0f;
    ds.delta = false;
    return ds;
}

// Wrap addressing: i mod n in [0, n).  Lookups at u, v in [0, 1] only ever
// need the in-range test; the 32-bit remainder covers every other index.
__device__ __forceinline__ int wrap_index(int i, int n) {
    if ((unsigned)i < (unsigned)n) return i;
    if (i == -1) return n - 1;
    if (i == n) return 0;
    const int r = i % n;
    return r < 0 ? r + n : r;
}

// Bilinear lookup with wrap addressing on the lat-long image (row 0 = top).
__device__ __forceinline__ vec3 env_lookup(const SceneView& s, const LightRec& e, float u, float v) {
    const int w = e.env_w, h = e.env_h;
    const float x = u * (float)w - 0.5f;
    const float y = v * (float)h - 0.5f;
    const float fx0 = floorf(x), fy0 = floorf(y);
    const float ax = x - fx0, ay = y - fy0;
    const int ix = (int)fx0, iy = (int)fy0;
    const int x0 = wrap_index(ix, w), x1 = wrap_index(ix + 1, w);
    const int y0 = wrap_index(iy, h), y1 = wrap_index(iy + 1, h);
    const float4* __restrict__ tx = s.env_rgb + e.env_base;
    const float4 c00 = tx[y0 * w + x0], c10 = tx[y0 * w + x1];
    const float4 c01 = tx[y1 * w + x0], c11 = tx[y1 * w + x1];
    const float w00 = (1.0f - ax) * (1.0f - ay), w10 = ax * (1.0f - ay);
    const float w01 = (1.0f - ax) * ay, w11 = ax * ay;
    vec3 r = add(add(add(scale(v3(c00.x, c00.y, c00.z), w00), scale(v3(c10.x, c10.y, c10.z), w10)),
                     scale(v3(c01.x, c01.y, c01.z), w01)),
                 scale(v3(c11.x, c11.y, c11.z), w11));
    return scale(r, e.env_scale);
}

// Envmap::environment_coords  include/optix/light/envmap.h:25-30
__device__ __forceinline__ void env_coords(vec3 d, float& u, float& v) {
    u = (float)atan2((double)d.z, (double)d.x) * (float)(0.5 / (double)kPi) + 0.5f;
    const float yc = fmax0(fmin0(d.y, 1.0f), -1.0f);
    v = (float)acos((double)yc) * kInvPi;
}
__device__ __forceinline__ vec3 env_eval(const SceneView& s, const LightRec& e, vec3 wi) {
    float u, v;
    env_coords(wi, u, v);
    return env_lookup(s, e, u, v);
}
__device__ __forceinline__ float env_pdf_uv(const SceneView& s, const LightRec& e, float u, float v) {
    const int x = min((int)(u * (float)e.env_w), e.env_w - 1);
    const int y = min((int)(v * (float)e.env_h), e.env_h - 1);
    return s.env_pdf[e.env_base + (size_t)y * e.env_w + x];
}
__device__ __forceinline__ float env_pdf(const SceneView& s, const LightRec& e, vec3 d) {
    float u, v;
    env_coords(d, u, v);
    return env_pdf_uv(s, e, u, v);
}
// Radiance and the emitter-sampling pdf (scene level, divided by the
// scene's emitter count nl as Scene::pdfEmitterDirection does) of direction
// d, from one evaluation of the lat-long coordinates.
__device__ __forceinline__ vec3 env_eval_pdf(const SceneView& s, const LightRec& e, int nl, vec3 d, float& pdf) {
    float u, v;
    env_coords(d, u, v);
    const float p = env_pdf_uv(s, e, u, v);
    pdf = nl == 1 ? p : p / (float)nl;
    return env_lookup(s, e, u, v);
}

// DiscretePDF::sampleReuse on a mesh CDF (include/core/math/dpdf.h:95-136).
// The lower_bound over cdf[0..n] searches only the range the CDF's bucket
// index gives for x (backend.h cdf_guide_bits): the same index as a search
// of all n + 1 entries, from a few dependent loads instead of log2(n).
__device__ __forceinline__ uint32_t cdf_sample_reuse(const float* __restrict__ cdf, uint32_t n,
                                                     const uint32_t* __restrict__ guide, uint32_t bits, float& x) {
    const uint32_t G = 1u << bits;
    const uint32_t g = min((uint32_t)(x * (float)G), G - 1u);  // x in [0, 1]; x * G is exact
    uint32_t lo = guide[g], count = guide[g + 1] - lo + 1;
    while (count > 0) {
        const uint32_t step = count / 2;
        const uint32_t it = lo + step;
        if (cdf[it] < x) {
            lo = it + 1;
            count -= step + 1;
        } else {
            count = step;
        }
    }
    uint32_t index = lo > 0 ? lo - 1 : 0;
    index = index < n - 1 ? index : n - 1;
    x = (x - cdf[index]) / (cdf[index + 1] - cdf[index]);
    return index;
}

// Emitter::sampleDirection: area (src/tracer/emitter/area.cpp:11-37), point
// (include/optix/light/point.h:15-31), environment (include/optix/light/envmap.h:32-79,
// third uniform drawn from the path's stream).
__device__ inline vec3 emitter_sample(const SceneView& s, const LightRec& e, float sx, float sy, const Its& its,
                                      DirSample& ds, Pcg32& rng) {
    if (e.type == DL_LIGHT_AREA) {
        const MeshRec& m = s.meshes[e.mesh];
        float x = sx;
        const uint32_t f = cdf_sample_reuse(s.light_cdf + m.cdf_offset, m.num_prims, s.light_guide + m.guide_offset,
                                            m.guide_bits, x) +
                           m.prim_offset;
        const uint32_t i0 = s.idx[3 * f], i1 = s.idx[3 * f + 1], i2 = s.idx[3 * f + 2];
        const vec3 p0 = ld3(s.pos, i0), p1 = ld3(s.pos, i1), p2 = ld3(s.pos, i2);
        const vec3 b = sq_triangle(x, sy);
        const vec3 position = add(add(scale(p0, b.x), scale(p1, b.y)), scale(p2, b.z));
        vec3 normal;
        if (m.has_normals)
            normal = add(add(scale(ld3(s.nrm, i0), b.x), scale(ld3(s.nrm, i1), b.y)), scale(ld3(s.nrm, i2), b.z));
        else
            normal = cross(sub(p1, p0), sub(p2, p1));
        normal = normalized(normal);
        ds = dir_sample(its.p, position, normal);
        const float pA = m.area_norm;
        const float dp = fmax0(dot(normal, neg(ds.d)), 0.0f);
        const float pw = dp != 0 ? pA * ds.dist * ds.dist / dp : 0;
        ds.pdf = pw;
        return pw != 0 ? divs(scale(pcol(e.radiance), fmax0(dot(its.sh.n, ds.d), 0.0f)), pw) : v3(0, 0, 0);
    }
    if (e.type == DL_LIGHT_POINT) {
        ds = dir_sample(its.p, pcol(e.position), v3(0, 0, 0));
        ds.pdf = 1;
        ds.delta = true;
        const float inv = 1.0f / ds.dist;
        const float c = fmax0(dot(its.sh.n, ds.d), 0.0f);
        return c > 0.f ? scale(scale(scale(pcol(e.radiance), c), inv), inv) : v3(0, 0, 0);
    }
    const float xz = rng.next_float();
    const uint32_t size = (uint32_t)e.env_w * (uint32_t)e.env_h;
    const uint32_t idx = min((uint32_t)(sx * (float)size), size - 1);
    uint32_t env_idx;
    float xy = sy;
    const float q = s.env_q[e.env_base + idx];
    if (xy < q) {
        env_idx = idx;
        xy /= q;
    } else {
        env_idx = s.env_alias[e.env_base + idx];
        xy = (xy - q) / (1.0f - q);
    }
    const uint32_t py = env_idx / (uint32_t)e.env_w, px = env_idx % (uint32_t)e.env_w;
    ds.pdf = s.env_pdf[e.env_base + env_idx];
    const float u = ((float)px + xy) / (float)e.env_w;
    const float phi = (float)((double)u * 2.0 * (double)kPi - (double)kPi);
    const float step = kPi / (float)e.env_h;
    const float th0 = (float)py * step;
    const float c0 = (float)cos((double)th0), c1 = (float)cos((double)(th0 + step));
    const float ct = c0 * (1.0f - xz) + c1 * xz;
    const float theta = (float)acos((double)ct);
    const float st = (float)sin((double)theta);
    const vec3 dir = v3((float)cos((double)phi) * st, ct, (float)sin((double)phi) * st);
    ds.d = dir;
    ds.dist = 1e16f;
    ds.delta = false;
    ds.n = neg(dir);
    const float v = theta * kInvPi;
    const vec3 val = env_lookup(s, e, u, v);
    const float c = fmax0(dot(its.sh.n, ds.d), 0.0f);
    return c > 0.f ? divs(scale(val, c), ds.pdf) : v3(0, 0, 0);
}

__device__ __forceinline__ float mis_weight(float a, float b) {  // path.cpp:104-108
    a *= a;
    b *= b;
    return a / (a + b);
}

// Scene::pdfEmitterDirection (scene.cpp:172-179) for emitter `l` of a scene
// with nl emitters.
__device__ __forceinline__ float pdf_emitter_direction(const SceneView& s, int l, int nl, const DirSample& ds) {
    const LightRec& e = s.lights[l];
    float p;
    if (e.type == DL_LIGHT_AREA) {
        const float pA = s.meshes[e.mesh].area_norm;
        const float dp = fmax0(dot(ds.n, neg(ds.d)), 0.0f);
        p = dp != 0 ? pA * ds.dist * ds.dist / dp : 0;
    } else if (e.type == DL_LIGHT_POINT) {
        p = 0;
    } else {
        p = env_pdf(s, e, ds.d);
    }
    if (nl == 1) return p;
    return p / (float)nl;
}

}  // namespace dl
