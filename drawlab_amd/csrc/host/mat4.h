// Row-major 4x4 float matrix helpers shared by the camera set-up and the
// plugin layer.  Products accumulate each entry as
// ((((0 + a0 b0) + a1 b1) + a2 b2) + a3 b3), the reference's row.dot(col)
// order (include/core/math/matrix.h:216-224).
#pragma once
#include <cmath>

namespace dlm {

inline void mat4_mul(const float a[4][4], const float b[4][4], float out[4][4]) {
    float r[4][4];
    for (int j = 0; j < 4; j++)
        for (int i = 0; i < 4; i++) {
            float s = 0.f;
            for (int k = 0; k < 4; k++) s += a[j][k] * b[k][i];
            r[j][i] = s;
        }
    for (int j = 0; j < 4; j++)
        for (int i = 0; i < 4; i++) out[j][i] = r[j][i];
}

// Inverse by the adjugate: the 2x2 minors of the upper and lower row pairs
// give every 3x3 cofactor (Laplace expansion along complementary rows).
// Returns false for a singular or non-finite determinant.
inline bool mat4_inverse(const float a[4][4], float out[4][4]) {
    const float s0 = a[0][0] * a[1][1] - a[1][0] * a[0][1];
    const float s1 = a[0][0] * a[1][2] - a[1][0] * a[0][2];
    const float s2 = a[0][0] * a[1][3] - a[1][0] * a[0][3];
    const float s3 = a[0][1] * a[1][2] - a[1][1] * a[0][2];
    const float s4 = a[0][1] * a[1][3] - a[1][1] * a[0][3];
    const float s5 = a[0][2] * a[1][3] - a[1][2] * a[0][3];
    const float c5 = a[2][2] * a[3][3] - a[3][2] * a[2][3];
    const float c4 = a[2][1] * a[3][3] - a[3][1] * a[2][3];
    const float c3 = a[2][1] * a[3][2] - a[3][1] * a[2][2];
    const float c2 = a[2][0] * a[3][3] - a[3][0] * a[2][3];
    const float c1 = a[2][0] * a[3][2] - a[3][0] * a[2][2];
    const float c0 = a[2][0] * a[3][1] - a[3][0] * a[2][1];
    const float det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
    if (det == 0.f || !std::isfinite(det)) return false;
    const float inv = 1.f / det;
    float r[4][4];
    r[0][0] = (a[1][1] * c5 - a[1][2] * c4 + a[1][3] * c3) * inv;
    r[0][1] = (-a[0][1] * c5 + a[0][2] * c4 - a[0][3] * c3) * inv;
    r[0][2] = (a[3][1] * s5 - a[3][2] * s4 + a[3][3] * s3) * inv;
    r[0][3] = (-a[2][1] * s5 + a[2][2] * s4 - a[2][3] * s3) * inv;
    r[1][0] = (-a[1][0] * c5 + a[1][2] * c2 - a[1][3] * c1) * inv;
    r[1][1] = (a[0][0] * c5 - a[0][2] * c2 + a[0][3] * c1) * inv;
    r[1][2] = (-a[3][0] * s5 + a[3][2] * s2 - a[3][3] * s1) * inv;
    r[1][3] = (a[2][0] * s5 - a[2][2] * s2 + a[2][3] * s1) * inv;
    r[2][0] = (a[1][0] * c4 - a[1][1] * c2 + a[1][3] * c0) * inv;
    r[2][1] = (-a[0][0] * c4 + a[0][1] * c2 - a[0][3] * c0) * inv;
    r[2][2] = (a[3][0] * s4 - a[3][1] * s2 + a[3][3] * s0) * inv;
    r[2][3] = (-a[2][0] * s4 + a[2][1] * s2 - a[2][3] * s0) * inv;
    r[3][0] = (-a[1][0] * c3 + a[1][1] * c1 - a[1][2] * c0) * inv;
    r[3][1] = (a[0][0] * c3 - a[0][1] * c1 + a[0][2] * c0) * inv;
    r[3][2] = (-a[3][0] * s3 + a[3][1] * s1 - a[3][2] * s0) * inv;
    r[3][3] = (a[2][0] * s3 - a[2][1] * s1 + a[2][2] * s0) * inv;
    for (int j = 0; j < 4; j++)
        for (int i = 0; i < 4; i++) out[j][i] = r[j][i];
    return true;
}

}  // namespace dlm
