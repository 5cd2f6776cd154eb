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
constexpr float kGKappaLo = -3.0f, kGKappaHi = 7.0f;  // log kappa = lo + (hi - lo) sigmoid(raw)
constexpr float kGuideFraction = 0.5f;    // one-sample MIS: guide vs BSDF sampling
constexpr uint32_t kGuideBatch = 1u << 18;  // records per Adam step
// Adam steps after each pass of a training render: kGuideWarmStepsPerPass
// until the network has taken kGuideWarmSteps steps, kGuideStepsPerPass after
constexpr int kGuideStepsPerPass = 4, kGuideWarmStepsPerPass = 16, kGuideWarmSteps = 256;
constexpr float kGuideMaxShare = 1.0f / 1024.0f;  // largest share of a batch one record may take

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

// Network inputs 8q..8q+7 of one shading point (q = 0..11): for q < 9 the
// pairs sin(2^k pi x_a), cos(2^k pi x_a) with a = q / 3 and k = 4 (q % 3)
// .. +3; q = 9 holds wo, the shading normal and the roughness; 10 and 11
// are zero padding.
__device__ __forceinline__ void guide_feature_chunk(const vec3& pn, const vec3& wo, const vec3& n, float rough, int q,
                                                    float* f /* 8 */) {
    if (q < 9) {
        // sin, cos of 2^k pi x for k = k0 .. k0 + 3: one sincosf, then the
        // double-angle identities (each step at most doubles the absolute
        // error, so k0 + 3 stays within 8 ulp of the direct evaluation --
        // far below the bf16 rounding of the network's inputs)
        const int a = q / 3, k0 = 4 * (q % 3);
        const float x = a == 0 ? pn.x : a == 1 ? pn.y : pn.z;
        float sn, cs;
        sincosf(ldexpf(3.14159265358979323846f, k0) * x, &sn, &cs);
        for (int j = 0; j < 4; j++) {
            f[2 * j] = sn;
            f[2 * j + 1] = cs;
            const float s2 = 2.0f * sn * cs, c2 = (cs - sn) * (cs + sn);
            sn = s2;
            cs = c2;
        }
    } else if (q == 9) {
        f[0] = wo.x;
        f[1] = wo.y;
        f[2] = wo.z;
        f[3] = n.x;
        f[4] = n.y;
        f[5] = n.z;
        f[6] = rough;
        f[7] = 0.0f;
    } else {
        for (int j = 0; j < 8; j++) f[j] = 0.0f;
    }
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

// Materials sampled with the proposal: those with a diffuse part or a GGX
// roughness of at least kGuideMinRoughness.  For glossier lobes, BSDF
// sampling combined with emitter sampling already follows the integrand
// closely, and the proposal's half of the samples would mostly be wasted.
constexpr float kGuideMinRoughness = 0.3f;
__host__ __device__ inline bool guide_material(int type, const float* p) {
    switch (type) {
        case DL_BSDF_DIFFUSE: return true;
        case DL_BSDF_MICROFACET: return p[3] > 0.0f || p[4] > 0.0f || p[5] > 0.0f || p[0] >= kGuideMinRoughness;
        case DL_BSDF_CONDUCTOR: return p[0] >= kGuideMinRoughness;
        case DL_BSDF_ANISO_GGX:
            return p[0] > 0.0f || p[1] > 0.0f || p[2] > 0.0f || sqrtf(p[6] * p[7]) >= kGuideMinRoughness;
        default: return false;
    }
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
        const float sg = 1.0f / (1.0f + expf(-o[5 * j + 1]));
        const float mx_ = o[5 * j + 2], my = o[5 * j + 3], mz = o[5 * j + 4];
        const float nm = fmaxf(sqrtf(mx_ * mx_ + my * my + mz * mz), 1e-6f);
        mix[5 * j] = e[j] / sum;
        mix[5 * j + 1] = expf(kGKappaLo + (kGKappaHi - kGKappaLo) * sg);
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

// ---- network evaluation on the matrix cores --------------------------------
typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kGWaves = 4;            // waves per block of the network kernels
constexpr int kGRowsPerWave = 64;
constexpr int kGWSlot = 64 * 96;      // bf16 elements reserved per layer (W or W^T)

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 pack8(const float* v) {
    bf16x8 r;
#pragma unroll
    for (int j = 0; j < 8; j++) r[j] = (short)f2bf(v[j]);
    return r;
}

struct Net {
    const uint16_t* w;   // W[l]: [64][in(l)] bf16 (layer 4 rows 48..63 zero), slot kGWSlot
    const uint16_t* wt;  // W^T[l]: [in(l)][64] bf16
    const float* bias;   // [5][64] (layer 4 entries 48..63 zero)
};

// One layer for the wave's 64 LDS rows (stride kGRowStride): out = act W^T
// + b.  Hidden layers write ReLU(out) back in place as bfloat16; the last
// layer calls emit(row, col, value) for col < 48.
template <typename Emit>
__device__ __forceinline__ void layer(const Net& net, int l, uint16_t* s, Emit emit) {
    const int lane = threadIdx.x & 63;
    const int K = glayer_in(l);
    const uint16_t* W = net.w + l * kGWSlot;
    const float* bias = net.bias + l * 64;
#pragma unroll 1
    for (int rb = 0; rb < 2; rb++) {
        f32x16 acc0 = {0}, acc1 = {0};
        const uint16_t* arow = s + (rb * 32 + (lane & 31)) * kGRowStride + (lane >> 5) * 8;
        const uint16_t* b0 = W + (lane & 31) * K + (lane >> 5) * 8;
        const uint16_t* b1 = W + (32 + (lane & 31)) * K + (lane >> 5) * 8;
        for (int kt = 0; kt < K / 16; kt++) {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(arow + kt * 16);
            acc0 = mfma(a, *reinterpret_cast<const bf16x8*>(b0 + kt * 16), acc0);
            acc1 = mfma(a, *reinterpret_cast<const bf16x8*>(b1 + kt * 16), acc1);
        }
        const int c0 = lane & 31, c1 = 32 + (lane & 31);
        const float bb0 = bias[c0], bb1 = bias[c1];
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int row = rb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const float v0 = acc0[r] + bb0, v1 = acc1[r] + bb1;
            if (l < 4) {
                s[row * kGRowStride + c0] = f2bf(fmaxf(v0, 0.0f));
                s[row * kGRowStride + c1] = f2bf(fmaxf(v1, 0.0f));
            } else {
                emit(row, c0, v0);
                if (c1 < kGOut) emit(row, c1, v1);
            }
        }
    }
    // the next layer reads rows written by other lanes of this wave (LDS
    // operations of one wave complete in order; this keeps the compiler
    // from reordering them)
    __builtin_amdgcn_wave_barrier();
}

// Forward pass of rows [row0, row0 + 64) whose bfloat16 inputs are already
// in the wave's LDS rows.  keep: copy each layer's input rows to act[l].
template <typename Emit>
__device__ __forceinline__ void forward_rows(const Net& net, uint16_t* s, uint32_t row0, uint32_t n, bool keep,
                                             uint16_t* const* act, Emit emit) {
    const int lane = threadIdx.x & 63;
    __builtin_amdgcn_wave_barrier();  // input rows were written by their own lanes
    for (int l = 0; l < 5; l++) {
        if (keep) {
            const int K = glayer_in(l);
            const uint32_t row = row0 + lane;
            if (row < n) {
                const uint4* src = reinterpret_cast<const uint4*>(s + lane * kGRowStride);
                uint4* dst = reinterpret_cast<uint4*>(act[l] + (size_t)row * K);
                for (int q = 0; q < K / 8; q++) dst[q] = src[q];
            }
        }
        layer(net, l, s, emit);
    }
}


}  // namespace dl
