// Host-side scene preparation that the reference performs in its plugins'
// activate() methods, producing the tables the device kernels read:
//   camera basis         VirtualCamera ctor   src/tracer/camera/virtual.cpp:16-42
//   filter tabulation    ImageBlock ctor      src/core/bitmap/block.cpp:12-27
//                        filters              src/tracer/rfilter/rfilter.cpp:13-98
//   area CDF             Mesh::activate       src/tracer/mesh/mesh.cpp:15-28
//                        DiscretePDF          include/core/math/dpdf.h:37-84
//   tangents             Mesh::calTangent     src/tracer/mesh/mesh.cpp:144-177
//   environment tables   EnvironmentMap::activate / build_alias_map
//                                             src/tracer/emitter/envmap.cpp:59-141
// All arithmetic is binary32 in the reference's evaluation order, so the
// tables are bit-identical to what the reference computes.
#include "scene_prep.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "../../../include/dl_hip.h"
#include "mat4.h"

namespace {

struct F3 {
    float x, y, z;
};
inline F3 sub3(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float dot3(F3 a, F3 b) {
    float s = 0.0f;
    s += a.x * b.x;
    s += a.y * b.y;
    s += a.z * b.z;
    return s;
}
inline F3 cross3(F3 a, F3 b) { return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float len3(F3 a) { return std::sqrt(dot3(a, a)); }
inline F3 norm3(F3 a) {
    float l = len3(a);
    return F3{a.x / l, a.y / l, a.z / l};
}
inline F3 ld3(const float* p) { return F3{p[0], p[1], p[2]}; }

const float kPiF = 3.14159265358979323846f;

}  // namespace

namespace dl {

float area_cdf(const float* V, const uint32_t* F, uint32_t nf, std::vector<float>& cdf) {
    cdf.resize((size_t)nf + 1);
    cdf[0] = 0.0f;
    for (uint32_t f = 0; f < nf; f++) {
        F3 p0 = ld3(V + 3 * (size_t)F[3 * f]), p1 = ld3(V + 3 * (size_t)F[3 * f + 1]),
           p2 = ld3(V + 3 * (size_t)F[3 * f + 2]);
        float area = 0.5f * len3(cross3(sub3(p1, p0), sub3(p2, p0)));
        cdf[f + 1] = cdf[f] + area;
    }
    float sum = cdf[nf];
    if (sum > 0) {
        float norm = 1.0f / sum;
        for (size_t i = 1; i < cdf.size(); ++i) cdf[i] *= norm;
        cdf[nf] = 1.0f;
        return norm;
    }
    return 0.0f;
}

void env_tables(const float* rgb, int w, int h, std::vector<float>& q, std::vector<uint32_t>& alias,
                std::vector<float>& pdf) {
    const uint32_t size = (uint32_t)w * (uint32_t)h;
    std::vector<float> imp(size);
    float cos0 = 1.0f;
    const float step_phi = (float)(2.0 * (double)kPiF) / (float)w;
    const float step_theta = kPiF / (float)h;
    for (int y = 0; y < h; ++y) {
        float th1 = (float)(y + 1) * step_theta;
        float cos1 = (float)std::cos((double)th1);
        float area = (cos0 - cos1) * step_phi;
        cos0 = cos1;
        for (int x = 0; x < w; ++x) {
            const size_t i = (size_t)y * w + x;
            const float* p = rgb + 3 * i;
            imp[i] = area * std::max(p[0], std::max(p[1], p[2]));
        }
    }
    q.resize(size);
    alias.resize(size);
    pdf.resize(size);
    float sum = 0.0f;
    for (uint32_t i = 0; i < size; ++i) sum += imp[i];
    for (uint32_t i = 0; i < size; ++i) q[i] = (float)size * imp[i] / sum;
    // Two-ended partition: small entries from the front, large from the back,
    // then each small entry borrows from the current large one.
    std::vector<uint32_t> order(size);
    uint32_t small = 0u, large = size;
    for (uint32_t i = 0; i < size; ++i) {
        alias[i] = i;
        if (q[i] < 1.0f)
            order[small++] = i;
        else
            order[--large] = i;
    }
    for (small = 0; small < large && large < size; ++small) {
        const uint32_t j = order[small], k = order[large];
        alias[j] = k;
        q[k] += q[j] - 1.0f;
        if (q[k] < 1.0f) large++;
    }
    const float inv = 1.0f / sum;
    for (uint32_t i = 0; i < size; ++i) {
        const float* p = rgb + 3 * (size_t)i;
        pdf[i] = std::max(p[0], std::max(p[1], p[2])) * inv;
    }
}

}  // namespace dl

extern "C" {

void dl_camera_virtual(const float eye[3], const float lookat[3], const float up[3], float fov, int width, int height,
                       dl_camera_desc* out) {
    std::memset(out, 0, sizeof(*out));
    out->type = DL_CAMERA_VIRTUAL;
    out->width = width;
    out->height = height;
    float aspect = width / (float)height;
    F3 e = ld3(eye), la = ld3(lookat), u = ld3(up);
    F3 W = sub3(la, e);
    float wlen = len3(W);
    F3 U = norm3(cross3(W, u));
    F3 V = norm3(cross3(U, W));
    float vlen = wlen * tanf(0.5f * fov * kPiF / 180.f);
    V = F3{V.x * vlen, V.y * vlen, V.z * vlen};
    float ulen = vlen * aspect;
    U = F3{U.x * ulen, U.y * ulen, U.z * ulen};
    const F3 src[4] = {e, U, V, W};
    float* dst[4] = {out->eye, out->U, out->V, out->W};
    for (int i = 0; i < 4; i++) {
        dst[i][0] = src[i].x;
        dst[i][1] = src[i].y;
        dst[i][2] = src[i].z;
    }
}

// PerspectiveCamera (perspective.cpp:21-70): perspective projection onto
// z = 1 with the cotangent of half the field of view, shifted and scaled into
// [0, 1]^2 sample space (aspect applied to y), inverted once here.
dl_status dl_camera_perspective(const float to_world[16], float fov, float near_clip, float far_clip, int width,
                                int height, dl_camera_desc* out) {
    std::memset(out, 0, sizeof(*out));
    if (width <= 0 || height <= 0 || !(far_clip > near_clip)) return DL_ERR_INVALID;
    out->type = DL_CAMERA_PERSPECTIVE;
    out->width = width;
    out->height = height;
    out->inv_size[0] = 1.f / width;
    out->inv_size[1] = 1.f / height;
    out->near_clip = near_clip;
    out->far_clip = far_clip;
    const float aspect = width / (float)height;
    const float recip = 1.0f / (far_clip - near_clip);
    const float cot = 1.0f / std::tan(fov / 2.0f * (float)(M_PI / 180.0));
    float persp[4][4] = {{cot, 0, 0, 0}, {0, cot, 0, 0}, {0, 0, far_clip * recip, -near_clip * far_clip * recip}, {0, 0, 1, 0}};
    float scale[4][4] = {{-0.5f, 0, 0, 0}, {0, -0.5f * aspect, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    float trans[4][4] = {{1, 0, 0, -1.0f}, {0, 1, 0, -1.0f / aspect}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    float st[4][4], stp[4][4], inv[4][4];
    dlm::mat4_mul(scale, trans, st);
    dlm::mat4_mul(st, persp, stp);
    if (!dlm::mat4_inverse(stp, inv)) return DL_ERR_INVALID;
    for (int j = 0; j < 4; j++)
        for (int i = 0; i < 4; i++) {
            out->sample_to_camera[4 * j + i] = inv[j][i];
            out->camera_to_world[4 * j + i] = to_world ? to_world[4 * j + i] : (i == j ? 1.f : 0.f);
        }
    return DL_OK;
}

// OpencvCamera (opencv.cpp:9-24): the inverse of the extrinsic (world to
// camera) matrix, the eye at its translation (inverse * (0, 0, 0, 1)) and
// the intrinsics as given.
dl_status dl_camera_opencv(const float extrinsic[16], float cx, float cy, float fx, float fy, int width, int height,
                           dl_camera_desc* out) {
    std::memset(out, 0, sizeof(*out));
    if (!extrinsic || width <= 0 || height <= 0 || fx == 0.0f || fy == 0.0f) return DL_ERR_INVALID;
    float ext[4][4], inv[4][4];
    for (int j = 0; j < 4; j++)
        for (int i = 0; i < 4; i++) ext[j][i] = extrinsic[4 * j + i];
    if (!dlm::mat4_inverse(ext, inv)) return DL_ERR_INVALID;
    out->type = DL_CAMERA_OPENCV;
    out->width = width;
    out->height = height;
    for (int j = 0; j < 4; j++)
        for (int i = 0; i < 4; i++) out->camera_to_world[4 * j + i] = inv[j][i];
    for (int j = 0; j < 3; j++) out->eye[j] = inv[j][3];
    out->intrinsics[0] = cx;
    out->intrinsics[1] = cy;
    out->intrinsics[2] = fx;
    out->intrinsics[3] = fy;
    return DL_OK;
}

void dl_film_filter(int type, float radius, float p0, float p1, dl_film_desc* out) {
    std::memset(out, 0, sizeof(*out));
    if (type == 1) radius = 0.5f;  // BoxFilter
    if (type == 2) radius = 1.0f;  // TentFilter
    out->radius = radius;
    // (a radius that is not finite or is far beyond any filter's stays
    // unconverted: scene upload refuses it)
    out->border = std::isfinite(radius) && std::fabs(radius) < 1e6f ? (int)std::ceil(radius - 0.5f) : 0;
    for (int i = 0; i < 32; ++i) {
        float x = (radius * i) / 32;
        float val;
        if (type == 0) {
            float alpha = -1.0f / (2.0f * p0 * p0);
            val = std::max(0.0f, std::exp(alpha * x * x) - std::exp(alpha * radius * radius));
        } else if (type == 1) {
            val = 1.0f;
        } else if (type == 2) {
            val = std::max(0.0f, 1.0f - std::fabs(x));
        } else {
            float B = p0, C = p1;
            x = std::fabs(2.0f * x / radius);
            float x2 = x * x, x3 = x2 * x;
            if (x < 1)
                val = 1.0f / 6.0f * ((12 - 9 * B - 6 * C) * x3 + (-18 + 12 * B + 6 * C) * x2 + (6 - 2 * B));
            else if (x < 2)
                val = 1.0f / 6.0f * ((-B - 6 * C) * x3 + (6 * B + 30 * C) * x2 + (-12 * B - 48 * C) * x + (8 * B + 24 * C));
            else
                val = 0.0f;
        }
        out->table[i] = val;
    }
    out->table[32] = 0.0f;
    out->lookup = 32 / radius;
}

void dl_mesh_tangents(const float* V, uint32_t nv, const uint32_t* F, uint32_t nf, const float* UV, float* T) {
    std::memset(T, 0, sizeof(float) * 3 * (size_t)nv);
    for (uint32_t f = 0; f < nf; f++) {
        const uint32_t i0 = F[3 * f], i1 = F[3 * f + 1], i2 = F[3 * f + 2];
        F3 e1 = sub3(ld3(V + 3 * (size_t)i1), ld3(V + 3 * (size_t)i0));
        F3 e2 = sub3(ld3(V + 3 * (size_t)i2), ld3(V + 3 * (size_t)i0));
        float a = UV[2 * i1] - UV[2 * i0], b = UV[2 * i1 + 1] - UV[2 * i0 + 1];
        float c = UV[2 * i2] - UV[2 * i0], d = UV[2 * i2 + 1] - UV[2 * i0 + 1];
        float det = 1 / (a * d - b * c);
        // first row of (adjugate / det) * [e1; e2]
        float m00 = d / det, m01 = -b / det;
        float r[3];
        const float ex[3] = {e1.x, e1.y, e1.z}, ey[3] = {e2.x, e2.y, e2.z};
        for (int j = 0; j < 3; j++) {
            float acc = 0.0f;
            acc += m00 * ex[j];
            acc += m01 * ey[j];
            r[j] = acc;
        }
        F3 t = norm3(F3{r[0], r[1], r[2]});
        const uint32_t ids[3] = {i0, i1, i2};
        for (int k = 0; k < 3; k++) {
            T[3 * (size_t)ids[k]] = t.x;
            T[3 * (size_t)ids[k] + 1] = t.y;
            T[3 * (size_t)ids[k] + 2] = t.z;
        }
    }
}

}  // extern "C"
