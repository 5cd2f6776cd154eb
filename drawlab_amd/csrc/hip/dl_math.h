// Device arithmetic for the gfx950 kernels.
//
// The backend must reproduce the reference CPU path bit-for-bit, so every
// vector operation spells out its binary32 evaluation order (the file is
// compiled with -ffp-contract=off and IEEE division/sqrt):
//   dot      ((0 + x*x') + y*y') + z*z'      include/core/math/vector.h:329-334
//   cross    (y z' - z y', z x' - x z', x y' - y x')   vector.h:341-346
//   length   sqrt of the zero-started sum; normalized() divides by it
//   s * v    v[i] * s                        vector.h friend operator*
// Transcendentals that the reference evaluates through binary64 libm calls
// (src/core/math/wrap.cpp) are evaluated in binary64 here too and rounded
// once to binary32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dl {

struct vec3 {
    float x, y, z;
};
struct vec2 {
    float x, y;
};

__device__ __forceinline__ vec3 v3(float x, float y, float z) { return vec3{x, y, z}; }
__device__ __forceinline__ vec3 add(vec3 a, vec3 b) { return vec3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ vec3 sub(vec3 a, vec3 b) { return vec3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ vec3 neg(vec3 a) { return vec3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ vec3 scale(vec3 a, float s) { return vec3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ vec3 divs(vec3 a, float s) { return vec3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ vec3 mul(vec3 a, vec3 b) { return vec3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ vec3 adds(vec3 a, float s) { return vec3{a.x + s, a.y + s, a.z + s}; }
__device__ __forceinline__ float dot(vec3 a, vec3 b) {
    float s = 0.0f;
    s = s + a.x * b.x;
    s = s + a.y * b.y;
    s = s + a.z * b.z;
    return s;
}
__device__ __forceinline__ vec3 cross(vec3 a, vec3 b) {
    return vec3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float length(vec3 a) { return sqrtf(dot(a, a)); }  // IEEE sqrt (not __fsqrt_rn: native)
__device__ __forceinline__ vec3 normalized(vec3 a) {
    float l = length(a);
    return vec3{a.x / l, a.y / l, a.z / l};
}
__device__ __forceinline__ float csum(vec3 c) { return (c.x + c.y) + c.z; }
__device__ __forceinline__ float cmax(vec3 c) {
    float m = (c.x < c.y) ? c.y : c.x;
    return (m < c.z) ? c.z : m;
}
__device__ __forceinline__ bool is_zero(vec3 c) { return c.x == 0.0f && c.y == 0.0f && c.z == 0.0f; }
__device__ __forceinline__ float fmax0(float a, float b) { return (a < b) ? b : a; }  // std::max(a, b)
__device__ __forceinline__ float fmin0(float a, float b) { return (b < a) ? b : a; }  // std::min(a, b)

constexpr float kEpsilon = 0.0001f;
constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kInvTwoPi = 0.15915494309189533577f;

// Orthonormal shading frame (include/core/math/frame.h:28-48).
struct Frame {
    vec3 t, b, n;
};
__device__ __forceinline__ Frame frame_from_normal(vec3 n) {
    vec3 t;
    if (fabs((double)n.z - 1.0) > (double)kEpsilon && fabs((double)n.z + 1.0) > (double)kEpsilon)
        t = v3(0.0f, 0.0f, 1.0f);
    else
        t = v3(1.0f, 0.0f, 0.0f);
    vec3 b = cross(n, t);
    float l = length(b);
    b = vec3{b.x / l, b.y / l, b.z / l};
    t = cross(b, n);
    return Frame{t, b, n};
}
__device__ __forceinline__ vec3 to_local(const Frame& f, vec3 v) { return v3(dot(v, f.t), dot(v, f.b), dot(f.n, v)); }
__device__ __forceinline__ vec3 to_world(const Frame& f, vec3 v) {
    return add(add(scale(f.t, v.x), scale(f.b, v.y)), scale(f.n, v.z));
}

// pcg32 (PCG-XSH-RR, O'Neill 2014) with (initstate, initseq) seeding and the
// 23-bit mantissa float conversion used by the reference sampler.
struct Pcg32 {
    uint64_t state, inc;
    __device__ __forceinline__ uint32_t next_uint() {
        uint64_t old = state;
        state = old * 0x5851f42d4c957f2dULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31u));
    }
    __device__ __forceinline__ void seed(uint64_t initstate, uint64_t initseq) {
        state = 0u;
        inc = (initseq << 1u) | 1u;
        next_uint();
        state += initstate;
        next_uint();
    }
    __device__ __forceinline__ float next_float() {
        uint32_t u = (next_uint() >> 9) | 0x3f800000u;
        return __uint_as_float(u) - 1.0f;
    }
};

}  // namespace dl
