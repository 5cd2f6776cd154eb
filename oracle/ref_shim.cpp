// C entry points over the reference's own math routines, for checking the
// oracle restatement.  Built only by oracle/ref.mk against /root/reference.
#include "core/math/math.h"
#include "core/math/point.h"
#include "core/math/vector.h"
#include "core/math/wrap.h"
#include <cstdint>

using namespace drawlab;

extern "C" {
// which: 0 disk 1 cosine hemisphere 2 beckmann 3 triangle 4 beckmann pdf
//        5 cosine pdf
void ref_warp(int which, const float* in, uint32_t n, float a0, float a1, float* out) {
    (void)a1;
    for (uint32_t i = 0; i < n; i++) {
        Point2f s(in[3 * i], in[3 * i + 1]);
        Vector3f v(in[3 * i], in[3 * i + 1], in[3 * i + 2]);
        float r[3] = {0, 0, 0};
        if (which == 0) {
            Point2f d = Warp::squareToUniformDisk(s);
            r[0] = d.x(); r[1] = d.y();
        } else if (which == 1) {
            Vector3f d = Warp::squareToCosineHemisphere(s);
            r[0] = d.x(); r[1] = d.y(); r[2] = d.z();
        } else if (which == 2) {
            Vector3f d = Warp::squareToBeckmann(s, a0);
            r[0] = d.x(); r[1] = d.y(); r[2] = d.z();
        } else if (which == 3) {
            Vector3f d = Warp::squareToUniformTriangle(s);
            r[0] = d.x(); r[1] = d.y(); r[2] = d.z();
        } else if (which == 4) {
            r[0] = Warp::squareToBeckmannPdf(v, a0);
        } else if (which == 5) {
            r[0] = Warp::squareToCosineHemispherePdf(v);
        }
        out[3 * i] = r[0]; out[3 * i + 1] = r[1]; out[3 * i + 2] = r[2];
    }
}
// which: 0 fresnel(cos, ext, int) 1 refract(wi, eta) 2 reflect(wi, n)
void ref_optics(int which, const float* in, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; i++) {
        const float* a = in + 6 * i;
        float r[3] = {0, 0, 0};
        if (which == 0) {
            r[0] = fresnel(a[0], a[1], a[2]);
        } else if (which == 1) {
            Vector3f d = refract(Vector3f(a[0], a[1], a[2]), a[3]);
            r[0] = d.x(); r[1] = d.y(); r[2] = d.z();
        } else if (which == 2) {
            Vector3f d = reflect(Vector3f(a[0], a[1], a[2]), Vector3f(a[3], a[4], a[5]));
            r[0] = d.x(); r[1] = d.y(); r[2] = d.z();
        }
        out[3 * i] = r[0]; out[3 * i + 1] = r[1]; out[3 * i + 2] = r[2];
    }
}
// Vector primitives: 0 dot 1 cross 2 normalized 3 length
void ref_vec(int which, const float* in, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; i++) {
        const float* a = in + 6 * i;
        Vector3f x(a[0], a[1], a[2]), y(a[3], a[4], a[5]);
        float r[3] = {0, 0, 0};
        if (which == 0) r[0] = x.dot(y);
        else if (which == 1) { Vector3f c = x.cross(y); r[0] = c.x(); r[1] = c.y(); r[2] = c.z(); }
        else if (which == 2) { Vector3f c = x.normalized(); r[0] = c.x(); r[1] = c.y(); r[2] = c.z(); }
        else r[0] = x.length();
        out[3 * i] = r[0]; out[3 * i + 1] = r[1]; out[3 * i + 2] = r[2];
    }
}
}
