// CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// This header restates the binary32 arithmetic of the DrawLab reference
// (xuxmin/DrawLab) exactly as its C++ evaluates it, so that the oracle's
// results can be compared bit-for-bit with the HIP backend.  Nothing in the
// product (drawlab_amd/) may include, link or call anything under oracle/;
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg do.
//
// Evaluation-order facts restated here (reference file:line):
//  * TVector::dot starts from a zero accumulator and adds the products in
//    index order: ((0 + x*x') + y*y') + z*z'    include/core/math/vector.h:329-334
//  * cross product component formulas            include/core/math/vector.h:341-346
//  * length = sqrt(squaredLength()) with the same zero-started sum, and
//    normalized() divides each component by it   include/core/math/vector.h:291-318
//  * cwiseInverse is 1 / d per component         include/core/math/vector.h:78-83
//  * scalar * vector multiplies rhs[i] * scalar  include/core/math/vector.h (friend operator*)
//  * Color3f::sum is m0 + m1 + m2 (no zero start) include/core/bitmap/color.h:30
//  * transcendental calls on float arguments resolve to the binary64 libm
//    routines (the float overloads are not visible in namespace drawlab), so
//    they are evaluated in binary64 and rounded once to binary32
//    src/core/math/wrap.cpp:33-97 (checked against oracle/_ref).
#pragma once
#include <cmath>
#include <cstdint>

namespace dlo {

static const float kEpsilon = 0.0001f;                     // math.h:14
static const float kPi = 3.14159265358979323846f;          // math.h:17
static const float kInvPi = 0.31830988618379067154f;       // math.h:18
static const float kInvTwoPi = 0.15915494309189533577f;    // math.h:19
static const float kInvFourPi = 0.07957747154594766788f;   // math.h:20

struct V3 {
    float x, y, z;
};
struct V2 {
    float x, y;
};

static inline V3 mk(float x, float y, float z) { return V3{x, y, z}; }
static inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
// vector * scalar and scalar * vector both compute v[i] * s.
static inline V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
static inline V3 operator*(float s, V3 a) { return V3{a.x * s, a.y * s, a.z * s}; }
static inline V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
static inline V3 cmul(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
static inline V3 cdiv(V3 a, V3 b) { return V3{a.x / b.x, a.y / b.y, a.z / b.z}; }
static inline V3 add_scalar(V3 a, float s) { return V3{a.x + s, a.y + s, a.z + s}; }

static inline float dot(V3 a, V3 b) {
    float s = 0.0f;
    s += a.x * b.x;
    s += a.y * b.y;
    s += a.z * b.z;
    return s;
}
static inline V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
static inline float length(V3 a) { return (float)std::sqrt((double)dot(a, a)); }
static inline V3 normalized(V3 a) {
    float l = length(a);
    return V3{a.x / l, a.y / l, a.z / l};
}
static inline float csum(V3 c) { return c.x + c.y + c.z; }
static inline float cmax(V3 c) {
    // TVector::maxCoeff: m = std::max(m, c[i]), and std::max(a, b) is (a < b) ? b : a
    float m = c.x;
    m = (m < c.y) ? c.y : m;
    m = (m < c.z) ? c.z : m;
    return m;
}
static inline bool is_zero(V3 c) { return c.x == 0.0f && c.y == 0.0f && c.z == 0.0f; }

// binary64 libm evaluation rounded to binary32 (see header comment).
static inline float dsqrt(float x) { return (float)std::sqrt((double)x); }
static inline double dcos(float x) { return std::cos((double)x); }
static inline double dsin(float x) { return std::sin((double)x); }

// Shading frame, include/core/math/frame.h:28-48.
struct Frame {
    V3 t, b, n;
};
static inline Frame frame_from_normal(V3 n) {
    V3 t;
    if (std::fabs((double)n.z - 1.0) > (double)kEpsilon &&
        std::fabs((double)n.z + 1.0) > (double)kEpsilon)
        t = mk(0.0f, 0.0f, 1.0f);
    else
        t = mk(1.0f, 0.0f, 0.0f);
    V3 b = cross(n, t);
    float l = length(b);
    b = V3{b.x / l, b.y / l, b.z / l};
    t = cross(b, n);
    return Frame{t, b, n};
}
// toLocal: (v.t, v.b, n.v)    frame.h:41-43
static inline V3 to_local(const Frame& f, V3 v) {
    return V3{dot(v, f.t), dot(v, f.b), dot(f.n, v)};
}
// toWorld: t*v.x + b*v.y + n*v.z   frame.h:46-48
static inline V3 to_world(const Frame& f, V3 v) {
    return (f.t * v.x + f.b * v.y) + f.n * v.z;
}

// pcg32, the PCG-XSH-RR generator of M. E. O'Neill (pcg-random.org, 2014)
// with the seeding procedure of the reference pcg32 header (initstate,
// initseq) and the 23-bit mantissa float conversion.  The reference uses
// this generator through src/tracer/sampler/independent.cpp:30-46.
struct Pcg32 {
    uint64_t state, inc;
    static const uint64_t kMult = 0x5851f42d4c957f2dULL;
    uint32_t next_uint() {
        uint64_t old = state;
        state = old * kMult + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31u));
    }
    void seed(uint64_t initstate, uint64_t initseq) {
        state = 0u;
        inc = (initseq << 1u) | 1u;
        next_uint();
        state += initstate;
        next_uint();
    }
    float next_float() {
        union {
            uint32_t u;
            float f;
        } x;
        x.u = (next_uint() >> 9) | 0x3f800000u;
        return x.f - 1.0f;
    }
    V2 next2d() {
        float a = next_float();
        float b = next_float();
        return V2{a, b};
    }
};

}  // namespace dlo
