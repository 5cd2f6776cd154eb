// Learned directional proposal: a mixture of 8 von Mises-Fisher lobes per
// shading point, predicted by a 96-64-64-64-64-48 ReLU MLP (bfloat16 weights
// and activations, float32 accumulation on the matrix cores) and trained
// online on the paths of each pass (KL divergence from the incident
// radiance).  Device helpers shared by the network kernels (guide.hip) and
// the guided shading kernels (integrator.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "../../../include/dl_hip.h"
#include "dl_math.h"

namespace dl {

constexpr int kGIn = 96, kGHid = 64, kGOut = 48, kGOutPad = 64, kGLobes = 8, kGFreq = 12;
constexpr int kGParams = 64 * 96 + 64 + 3 * (64 * 64 + 64) + 48 * 64 + 48;  // 21,808
constexpr int kGMix = kGLobes * 5;        // w, kappa, mu.xyz per lobe
constexpr int kGRowStride = kGIn + 8;     // bf16 activations per LDS row (16-byte pad)
constexpr int kGRecFloats = 24;           // training record
constexpr int kGRecDepth = 4;             // path vertices recorded per recorded path
constexpr float kGKappaLo = -3.0f, kGKappaHi = 7.0f;
constexpr float kGuideFraction = 0.5f;    // one-sample MIS: guide vs BSDF sampling

// Offsets of each layer's W[out][in] and b[out] in the parameter vector.
__host__ __device__ constexpr int glayer_in(int l) { return l == 0 ? 96 : 64; }
__host__ __device__ constexpr int glayer_out(int l) { return l == 4 ? 48 : 64; }
__host__ __device__ constexpr int glayer_w(int l) {
    return l == 0 ? 0 : l == 1 ? 6208 : l == 2 ? 6208 + 4160 : l == 3 ? 6208 + 2 * 4160 : 6208 + 3 * 4160;
}
__host__ __device__ constexpr int glayer_b(int l) { return glayer_w(l) + glayer_out(l) * glayer_in(l); }

// float32 -> bfloat16 bits, round to nearest even (finite inputs).
__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
#ifdef __HIP_DEVICE_COMPILE__
    uint32_t u = __float_as_uint(f);
#else
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
#endif
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// Training record layout (floats): position (normalised) 0..2, wo 3..5,
// shading normal 6..8, roughness 9, sampled direction 10..12, sampling pdf
// 13, throughput after the vertex 14..16, radiance accumulated before the
// continuation 17..19, path id 20 (bits), valid 21.
enum { kRecPos = 0, kRecWo = 3, kRecN = 6, kRecRough = 9, kRecDir = 10, kRecPdf = 13, kRecThr = 14, kRecRad = 17,
       kRecPath = 20, kRecValid = 21 };

// Network input of one shading point.
__device__ __forceinline__ void guide_features(const vec3& pn, const vec3& wo, const vec3& n, float rough,
                                               float* f /* 96 */) {
    const float pc[3] = {pn.x, pn.y, pn.z};
    for (int a = 0; a < 3; a++) {
        float scale = 3.14159265358979323846f;
        for (int k = 0; k < kGFreq; k++) {
            float s, c;
            sincosf(scale * pc[a], &s, &c);
            f[24 * a + 2 * k] = s;
            f[24 * a + 2 * k + 1] = c;
            scale *= 2.0f;
        }
    }
    f[72] = wo.x;
    f[73] = wo.y;
    f[74] = wo.z;
    f[75] = n.x;
    f[76] = n.y;
    f[77] = n.z;
    f[78] = rough;
    for (int i = 79; i < kGIn; i++) f[i] = 0.0f;
}

// Roughness input per material: 1 for diffuse, alpha for the microfacet
// models, sqrt(ax ay) for anisotropic GGX.
__device__ __forceinline__ float guide_roughness(int type, const float* p) {
    switch (type) {
        case DL_BSDF_DIFFUSE: return 1.0f;
        case DL_BSDF_MICROFACET: return p[0];
        case DL_BSDF_CONDUCTOR: return p[0];
        case DL_BSDF_ANISO_GGX: return sqrtf(p[6] * p[7]);
        default: return 0.0f;
    }
}

__device__ __forceinline__ bool guide_type(int type) {
    return type == DL_BSDF_DIFFUSE || type == DL_BSDF_MICROFACET || type == DL_BSDF_CONDUCTOR ||
           type == DL_BSDF_ANISO_GGX;
}

// Raw network outputs (48) -> mixture (w, kappa, mu) of 8 lobes.
__device__ __forceinline__ void guide_decode(const float* o, float* mix /* 40 */) {
    float mx = o[0];
    for (int j = 1; j < kGLobes; j++) mx = fmaxf(mx, o[5 * j]);
    float e[kGLobes], sum = 0.0f;
    for (int j = 0; j < kGLobes; j++) {
        e[j] = expf(o[5 * j] - mx);
        sum += e[j];
    }
    for (int j = 0; j < kGLobes; j++) {
        const float kr = fminf(fmaxf(o[5 * j + 1], kGKappaLo), kGKappaHi);
        const float mx_ = o[5 * j + 2], my = o[5 * j + 3], mz = o[5 * j + 4];
        const float nm = fmaxf(sqrtf(mx_ * mx_ + my * my + mz * mz), 1e-6f);
        mix[5 * j] = e[j] / sum;
        mix[5 * j + 1] = expf(kr);
        mix[5 * j + 2] = mx_ / nm;
        mix[5 * j + 3] = my / nm;
        mix[5 * j + 4] = mz / nm;
    }
}

__device__ __forceinline__ float vmf_pdf(float kappa, const vec3& mu, const vec3& d) {
    const float c = kappa / (6.28318530717958648f * -expm1f(-2.0f * kappa));
    return c * expf(kappa * (dot(mu, d) - 1.0f));
}

__device__ __forceinline__ float guide_pdf(const float* mix, const vec3& d) {
    float q = 0.0f;
    for (int j = 0; j < kGLobes; j++)
        q += mix[5 * j] * vmf_pdf(mix[5 * j + 1], v3(mix[5 * j + 2], mix[5 * j + 3], mix[5 * j + 4]), d);
    return q;
}

// Orthonormal basis completing a unit vector (Duff et al. 2017).
__device__ __forceinline__ void onb(const vec3& n, vec3& t, vec3& b) {
    const float sign = n.z >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (sign + n.z);
    const float bb = n.x * n.y * a;
    t = v3(1.0f + sign * n.x * n.x * a, sign * bb, -sign * n.x);
    b = v3(bb, sign + n.y * n.y * a, -n.y);
}

// Sample the mixture: u0 picks the lobe (and is reused as u1 within it),
// u2 the azimuth; Jakob 2012 for the polar angle.
__device__ __forceinline__ vec3 guide_sample(const float* mix, float u0, float u2) {
    int j = 0;
    float acc = 0.0f;
    for (; j < kGLobes - 1; j++) {
        const float w = mix[5 * j];
        if (u0 < acc + w) break;
        acc += w;
    }
    const float w = mix[5 * j];
    float u1 = w > 0.0f ? (u0 - acc) / w : 0.5f;
    u1 = fminf(fmaxf(u1, 0.0f), 0.99999994f);
    const float kappa = mix[5 * j + 1];
    const vec3 mu = v3(mix[5 * j + 2], mix[5 * j + 3], mix[5 * j + 4]);
    float c = 1.0f + logf(u1 + (1.0f - u1) * expf(-2.0f * kappa)) / kappa;
    c = fminf(fmaxf(c, -1.0f), 1.0f);
    const float s = sqrtf(fmaxf(0.0f, 1.0f - c * c));
    float sp, cp;
    sincosf(6.28318530717958648f * u2, &sp, &cp);
    vec3 t, b;
    onb(mu, t, b);
    return normalized(add(add(scale(t, s * cp), scale(b, s * sp)), scale(mu, c)));
}

}  // namespace dl
