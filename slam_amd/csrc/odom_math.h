// odom_math.h -- small fixed-size float/double linear algebra used by the
// device-side solve (the last workgroup of each step) and by the host class.
// Every routine is written out operation by operation; the library is built
// with -ffp-contract=off, so host and device evaluate them identically.
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <math.h>

namespace slam {

#define ODOM_HD __host__ __device__ __forceinline__

// exp(x) for x <= 0: Cody-Waite reduction by ln2 (high part with 9 trailing
// zero bits) and a degree-7 Taylor polynomial in Horner form; 0 below -87.
ODOM_HD float bexpf(float x)
{
    if (!(x >= -87.0f)) return 0.0f;
    if (x > 0.0f) x = 0.0f;
    const float kf = __builtin_rintf(x * 1.44269504f);
    float r = x - kf * 0.693145751953125f;
    r = r - kf * 1.42860677e-06f;
    float p = 1.98412698e-04f;
    p = p * r + 1.38888889e-03f;
    p = p * r + 8.33333333e-03f;
    p = p * r + 4.16666667e-02f;
    p = p * r + 1.66666667e-01f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    const int kk = (int)kf;
    const unsigned int bits = (unsigned int)(kk + 127) << 23;
    return p * __builtin_bit_cast(float, bits);
}

// __float2int_rn: round half to even, saturating, NaN -> 0.  The saturation
// bound only has to exceed any image dimension.
ODOM_HD int f2i_rn(float f)
{
    float r = __builtin_rintf(f);
    if (!(r == r)) r = 0.0f;
    if (r > 1.0e9f) r = 1.0e9f;
    if (r < -1.0e9f) r = -1.0e9f;
    return (int)r;
}

ODOM_HD void mm3d(const double* a, const double* b, double* o)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            o[i * 3 + j] = a[i * 3] * b[j] + a[i * 3 + 1] * b[3 + j] + a[i * 3 + 2] * b[6 + j];
}

ODOM_HD void mm3f(const float* a, const float* b, float* o)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            o[i * 3 + j] = a[i * 3] * b[j] + a[i * 3 + 1] * b[3 + j] + a[i * 3 + 2] * b[6 + j];
}

ODOM_HD void mm4d(const double* a, const double* b, double* o)
{
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            o[i * 4 + j] = a[i * 4] * b[j] + a[i * 4 + 1] * b[4 + j] + a[i * 4 + 2] * b[8 + j] +
                           a[i * 4 + 3] * b[12 + j];
}

// 3x3 inverse by cofactors in float (Rprev.inverse(), RGBDOdometryef.cpp:386)
ODOM_HD void inv3f(const float* m, float* o)
{
    const float c0 = m[4] * m[8] - m[5] * m[7];
    const float c1 = m[5] * m[6] - m[3] * m[8];
    const float c2 = m[3] * m[7] - m[4] * m[6];
    const float det = m[0] * c0 + m[1] * c1 + m[2] * c2;
    const float id = 1.0f / det;
    o[0] = c0 * id;
    o[1] = (m[2] * m[7] - m[1] * m[8]) * id;
    o[2] = (m[1] * m[5] - m[2] * m[4]) * id;
    o[3] = c1 * id;
    o[4] = (m[0] * m[8] - m[2] * m[6]) * id;
    o[5] = (m[2] * m[3] - m[0] * m[5]) * id;
    o[6] = c2 * id;
    o[7] = (m[1] * m[6] - m[0] * m[7]) * id;
    o[8] = (m[0] * m[4] - m[1] * m[3]) * id;
}

// K and its closed-form inverse in double, from float level intrinsics
ODOM_HD void kmat(float fx, float fy, float cx, float cy, double* K, double* Ki)
{
    const double dfx = fx, dfy = fy, dcx = cx, dcy = cy;
    K[0] = dfx; K[1] = 0; K[2] = dcx; K[3] = 0; K[4] = dfy; K[5] = dcy; K[6] = 0; K[7] = 0; K[8] = 1;
    Ki[0] = 1.0 / dfx; Ki[1] = 0; Ki[2] = -dcx / dfx;
    Ki[3] = 0; Ki[4] = 1.0 / dfy; Ki[5] = -dcy / dfy;
    Ki[6] = 0; Ki[7] = 0; Ki[8] = 1;
}

// Symmetric solve, LDL^T with diagonal pivoting on the largest remaining
// |diagonal| and a pseudo-inverse of D (|d| <= DBL_MIN gives a zero
// component): the semantics of Eigen's LDLT::solve used by the reference
// (RGBDOdometryef.cpp:366, 544).  Fully unrolled for a fixed N, with the
// pivot swaps done by compile-time-indexed selects, so the arrays stay in
// registers on the device.  Returns the number of zero pivots.
template <int N>
ODOM_HD int ldlt_solve(const double* Ain, const double* b, double* x)
{
    double A[N][N];
    int perm[N];
    double y[N];
    int zeros = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        perm[i] = i;
#pragma unroll
        for (int j = 0; j < N; ++j) A[i][j] = Ain[i * N + j];
    }
#pragma unroll
    for (int k = 0; k < N; ++k) {
        int p = k;
        double best = fabs(A[k][k]);
#pragma unroll
        for (int i = k + 1; i < N; ++i) {
            const bool g = fabs(A[i][i]) > best;
            best = g ? fabs(A[i][i]) : best;
            p = g ? i : p;
        }
#pragma unroll
        for (int q = k + 1; q < N; ++q) {
            if (p == q) {
#pragma unroll
                for (int j = 0; j < N; ++j) { const double t = A[k][j]; A[k][j] = A[q][j]; A[q][j] = t; }
#pragma unroll
                for (int i = 0; i < N; ++i) { const double t = A[i][k]; A[i][k] = A[i][q]; A[i][q] = t; }
                const int t = perm[k]; perm[k] = perm[q]; perm[q] = t;
            }
        }
        const double d = A[k][k];
        if (fabs(d) <= DBL_MIN) {
#pragma unroll
            for (int i = k + 1; i < N; ++i) A[i][k] = 0.0;
        } else {
#pragma unroll
            for (int i = k + 1; i < N; ++i) A[i][k] = A[i][k] / d;
#pragma unroll
            for (int i = k + 1; i < N; ++i)
#pragma unroll
                for (int j = k + 1; j <= i; ++j) {
                    A[i][j] = A[i][j] - A[i][k] * A[j][k] * d;
                    A[j][i] = A[i][j];
                }
        }
    }
    // y = P b
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double v = 0.0;
#pragma unroll
        for (int q = 0; q < N; ++q) v = (perm[i] == q) ? b[q] : v;
        y[i] = v;
    }
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < i; ++j) y[i] = y[i] - A[i][j] * y[j];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const double d = A[i][i];
        if (fabs(d) > DBL_MIN) y[i] = y[i] / d;
        else { y[i] = 0.0; ++zeros; }
    }
#pragma unroll
    for (int i = N - 1; i >= 0; --i)
#pragma unroll
        for (int j = i + 1; j < N; ++j) y[i] = y[i] - A[j][i] * y[j];
    // x = P^T y
#pragma unroll
    for (int q = 0; q < N; ++q) {
        double v = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) v = (perm[i] == q) ? y[i] : v;
        x[q] = v;
    }
    return zeros;
}

// The same solve for a well-conditioned symmetric positive definite A: LDL^T
// without pivoting (in exact arithmetic the same x as ldlt_solve; ~6x fewer
// operations, no row swaps).  Returns false, leaving x unspecified, when a
// pivot is not above 1e-9 of the largest diagonal entry: the caller then uses
// ldlt_solve, whose pivoting and zero-pivot rule decide singular systems.
// 1/d for the positive pivots of ldlt_spd.  On the device: v_rcp_f64 and two
// Newton steps (each at least doubles the correct bits, so the result is
// within an ulp of the quotient) instead of the full division sequence.
ODOM_HD double recip_pos(double d)
{
#ifdef __HIP_DEVICE_COMPILE__
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    return fma(r, e, r);
#else
    return 1.0 / d;
#endif
}

template <int N>
ODOM_HD bool ldlt_spd(const double* A, const double* b, double* x)
{
    double L[N][N], d[N], id[N], y[N];
    double amax = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) amax = fmax(amax, fabs(A[i * N + i]));
    const double tol = 1e-9 * amax;
    bool ok = amax > 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        double dj = A[j * N + j];
#pragma unroll
        for (int k = 0; k < j; ++k) dj = dj - L[j][k] * L[j][k] * d[k];
        ok = ok && (dj > tol);
        d[j] = dj;
        id[j] = recip_pos(dj);
#pragma unroll
        for (int i = j + 1; i < N; ++i) {
            double v = A[i * N + j];
#pragma unroll
            for (int k = 0; k < j; ++k) v = v - L[i][k] * L[j][k] * d[k];
            L[i][j] = v * id[j];
        }
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double v = b[i];
#pragma unroll
        for (int k = 0; k < i; ++k) v = v - L[i][k] * y[k];
        y[i] = v;
    }
#pragma unroll
    for (int i = N - 1; i >= 0; --i) {
        double v = y[i] * id[i];
#pragma unroll
        for (int k = i + 1; k < N; ++k) v = v - L[k][i] * x[k];
        x[i] = v;
    }
    return ok;
}

// OdometryProvider::rodrigues, utils.h:16-52
ODOM_HD void rodrigues(const double* src, double* R)
{
    for (int i = 0; i < 9; ++i) R[i] = (i % 4 == 0) ? 1.0 : 0.0;
    double rx = src[0], ry = src[1], rz = src[2];
    const double theta = sqrt(rx * rx + ry * ry + rz * rz);
    if (theta >= DBL_EPSILON) {
#ifdef __HIP_DEVICE_COMPILE__
        // the increments of a Gauss-Newton step are small: below 0.25 rad the
        // Taylor series to theta^13 (sin) and theta^14 (cos) are exact to
        // double precision (next terms below 1e-21); above, the library
        double s, c;
        if (theta <= 0.25) {
            const double t2 = theta * theta;
            s = theta * (1.0 + t2 * (-1.0 / 6 + t2 * (1.0 / 120 + t2 * (-1.0 / 5040 + t2 * (1.0 / 362880 +
                t2 * (-1.0 / 39916800 + t2 * (1.0 / 6227020800.0)))))));
            c = 1.0 + t2 * (-0.5 + t2 * (1.0 / 24 + t2 * (-1.0 / 720 + t2 * (1.0 / 40320 + t2 * (-1.0 / 3628800 +
                t2 * (1.0 / 479001600.0 + t2 * (-1.0 / 87178291200.0)))))));
        } else {
            sincos(theta, &s, &c);
        }
#else
        const double c = cos(theta), s = sin(theta);
#endif
        const double c1 = 1. - c;
        const double itheta = theta ? 1. / theta : 0.;
        rx *= itheta; ry *= itheta; rz *= itheta;
        const double rrt[9] = {rx * rx, rx * ry, rx * rz, rx * ry, ry * ry, ry * rz, rx * rz, ry * rz, rz * rz};
        const double rx_[9] = {0, -rz, ry, rz, 0, -rx, -ry, rx, 0};
        const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        for (int k = 0; k < 9; ++k) R[k] = c * I[k] + c1 * rrt[k] + s * rx_[k];
    }
}

// Unpack of the 29 packed sums into symmetric A and b (reduce.cu:475-489)
ODOM_HD void unpack_se3(const double* in, double* A, double* b)
{
    int s = 0;
    for (int i = 0; i < 6; ++i)
        for (int j = i; j < 7; ++j) {
            const double v = in[s++];
            if (j == 6) b[i] = v;
            else { A[j * 6 + i] = v; A[i * 6 + j] = v; }
        }
}

// reduce.cu:1125-1139
ODOM_HD void unpack_so3(const double* in, double* A, double* b)
{
    int s = 0;
    for (int i = 0; i < 3; ++i)
        for (int j = i; j < 4; ++j) {
            const double v = in[s++];
            if (j == 3) b[i] = v;
            else { A[j * 3 + i] = v; A[i * 3 + j] = v; }
        }
}

}  // namespace slam
