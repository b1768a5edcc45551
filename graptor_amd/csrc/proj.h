// Pointwise projections shared by the stencil kernels: the reference's
// off-diagonal weight, its closed-form 2x2 nuclear route and a cyclic Jacobi
// eigen-decomposition of symmetric 3x3 matrices.
#pragma once

#include "common.h"

namespace gr {

constexpr float W_OFF = 1.4142f;  // the reference's float literal for sqrt(2)

// ---- closed-form projection helpers -------------------------------------

// 2x2 Gram route of update_NormV_Nucl (2D:124-204): M = sum_k S_k e_k e_k^T
// with S_k = min(1, 1/(alphainv sigma_k)); G treated as diagonal when its
// off-diagonal is negligible against the smaller diagonal entry.  The route's
// branches and formulas are the reference's, evaluated in FP64 from the FP32
// Gram entries: in FP32 its sigma_2 = sqrt((d - f)/2) and second eigenvector
// (h a + b^2, h b + b c), h = (e - f)/2, cancel catastrophically when the
// Gram matrix is near rank one, which moves a whole FP32 iteration by 1e-2
// relative after 150 steps (two FP32 compilations of the reference text differ
// from each other by as much); in FP64 the route equals the exact spectral
// projection to 1e-7 and the iteration stays within 1e-6 of FP64 throughout.
__device__ __forceinline__ void nucl2d(float af, float bf, float cf, float ainvf, float& m00,
                                       float& m01, float& m11)
{
    const double a = af, b = bf, c = cf, ainv = ainvf;
    const double eps = 1e-5;
    const double o = fmin(a, c);
    if ((b * b) / (o * o) > eps && o * o > 0.0) {
        const double d = a + c, e = a - c;
        double f = 4.0 * b * b + e * e;
        f = f < 0.0 ? 0.0 : sqrt(f);
        const double s2 = d - f > 0.0 ? sqrt((d - f) * 0.5) : 0.0;
        const double s1 = d + f > 0.0 ? sqrt((d + f) * 0.5) : 0.0;
        const double g = (e + f) * 0.5, hh = (e - f) * 0.5;
        double v0 = g * a + b * b, v1 = g * b + c * b;
        double v2 = hh * a + b * b, v3 = hh * b + b * c;
        const double n1 = sqrt(v0 * v0 + v1 * v1), n2 = sqrt(v2 * v2 + v3 * v3);
        if (n1 > 0.0) { v0 /= n1; v1 /= n1; }
        if (n2 > 0.0) { v2 /= n2; v3 /= n2; }
        const double S1 = s1 * ainv > 1.0 ? 1.0 / (s1 * ainv) : 1.0;
        const double S2 = s2 * ainv > 1.0 ? 1.0 / (s2 * ainv) : 1.0;
        m00 = (float)(v0 * v0 * S1 + v2 * v2 * S2);
        m11 = (float)(v1 * v1 * S1 + v3 * v3 * S2);
        m01 = (float)(S1 * v0 * v1 + S2 * v2 * v3);
    } else {
        const double s1 = sqrt(a), s2 = sqrt(c);
        m00 = (float)(ainv * s1 > 1.0 ? 1.0 / (ainv * s1) : 1.0);
        m11 = (float)(ainv * s2 > 1.0 ? 1.0 / (ainv * s2) : 1.0);
        m01 = 0.f;
    }
}

// Cyclic Jacobi eigen-decomposition of a symmetric 3x3 matrix
// A = (a00, a11, a22, a01, a02, a12): ev[k], eigenvector k in column k of V.
template <typename T>
__device__ __forceinline__ void jacobi3(const T (&A)[6], T (&ev)[3], T (&V)[3][3])
{
    T a[3][3] = {{A[0], A[3], A[4]}, {A[3], A[1], A[5]}, {A[4], A[5], A[2]}};
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) V[r][k] = (r == k) ? T(1) : T(0);
#pragma unroll
    for (int sweep = 0; sweep < 6; ++sweep) {
#pragma unroll
        for (int pq = 0; pq < 3; ++pq) {
            const int P = (pq == 2) ? 1 : 0, Q = (pq == 0) ? 1 : 2;
            const T apq = a[P][Q];
            if (apq == T(0)) continue;
            const T theta = (a[Q][Q] - a[P][P]) / (T(2) * apq);
            const T t = copysign(T(1), theta) / (fabs(theta) + sqrt(theta * theta + T(1)));
            const T cs = T(1) / sqrt(t * t + T(1)), sn = t * cs;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const T akp = a[k][P], akq = a[k][Q];
                a[k][P] = cs * akp - sn * akq;
                a[k][Q] = sn * akp + cs * akq;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const T apk = a[P][k], aqk = a[Q][k];
                a[P][k] = cs * apk - sn * aqk;
                a[Q][k] = sn * apk + cs * aqk;
            }
            a[P][Q] = a[Q][P] = T(0);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const T vkp = V[k][P], vkq = V[k][Q];
                V[k][P] = cs * vkp - sn * vkq;
                V[k][Q] = sn * vkp + cs * vkq;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) ev[k] = a[k][k];
}

}  // namespace gr
