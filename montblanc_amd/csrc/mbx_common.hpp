// Shared device/host helpers of the gfx950 RIME library.
//
// Complex numbers are (re, im) pairs of T, stored interleaved exactly like
// numpy/torch complex64/complex128. A Jones matrix is four complex numbers,
// row-major [XX, XY, YX, YY]; its product rules are those of
// montblanc/src/montblanc/include/jones.cuh:36-115 (A.B and A.B^H), written
// here as plain per-thread 2x2 arithmetic: one thread owns all four
// correlations, so no cross-lane shuffles are involved (the reference's
// 4-lane shuffle layout assumes 32-wide warps).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "montblanc_amd.h"

namespace mbx {

template <typename T> struct cx { T re, im; };

template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// Jones matrix, row-major.
template <typename T> struct J4 { cx<T> v[4]; };

template <typename T>
__device__ __forceinline__ J4<T> jzero() {
    J4<T> r;
#pragma unroll
    for (int i = 0; i < 4; ++i) r.v[i] = {T(0), T(0)};
    return r;
}

// r = A.B
template <typename T>
__device__ __forceinline__ J4<T> jmul(const J4<T> &A, const J4<T> &B) {
    J4<T> r;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            cx<T> a0 = A.v[2 * i], a1 = A.v[2 * i + 1];
            cx<T> b0 = B.v[j], b1 = B.v[2 + j];
            r.v[2 * i + j].re = a0.re * b0.re - a0.im * b0.im + a1.re * b1.re - a1.im * b1.im;
            r.v[2 * i + j].im = a0.re * b0.im + a0.im * b0.re + a1.re * b1.im + a1.im * b1.re;
        }
    return r;
}

// r = A.B^H
template <typename T>
__device__ __forceinline__ J4<T> jmul_h(const J4<T> &A, const J4<T> &B) {
    J4<T> r;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            cx<T> a0 = A.v[2 * i], a1 = A.v[2 * i + 1];
            cx<T> b0 = B.v[2 * j], b1 = B.v[2 * j + 1];
            // a0*conj(b0) + a1*conj(b1)
            r.v[2 * i + j].re = a0.re * b0.re + a0.im * b0.im + a1.re * b1.re + a1.im * b1.im;
            r.v[2 * i + j].im = a0.im * b0.re - a0.re * b0.im + a1.im * b1.re - a1.re * b1.im;
        }
    return r;
}

// acc += f * A.B^H  (the coherency accumulation, 32 FMAs + scaling of A)
template <typename T>
__device__ __forceinline__ void jacc_h(J4<T> &acc, const J4<T> &A, const J4<T> &B) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            cx<T> a0 = A.v[2 * i], a1 = A.v[2 * i + 1];
            cx<T> b0 = B.v[2 * j], b1 = B.v[2 * j + 1];
            cx<T> &r = acc.v[2 * i + j];
            r.re = fma(a0.re, b0.re, r.re);
            r.re = fma(a0.im, b0.im, r.re);
            r.re = fma(a1.re, b1.re, r.re);
            r.re = fma(a1.im, b1.im, r.re);
            r.im = fma(a0.im, b0.re, r.im);
            r.im = fma(-a0.re, b0.im, r.im);
            r.im = fma(a1.im, b1.re, r.im);
            r.im = fma(-a1.re, b1.im, r.im);
        }
}

template <typename T>
__device__ __forceinline__ J4<T> jscale(const J4<T> &A, T f) {
    J4<T> r;
#pragma unroll
    for (int i = 0; i < 4; ++i) r.v[i] = {A.v[i].re * f, A.v[i].im * f};
    return r;
}

// Vector loads/stores of a Jones matrix from interleaved complex storage.
__device__ __forceinline__ J4<float> jload(const cx<float> *p) {
    const float4 *q = reinterpret_cast<const float4 *>(p);
    float4 a = q[0], b = q[1];
    J4<float> r;
    r.v[0] = {a.x, a.y}; r.v[1] = {a.z, a.w};
    r.v[2] = {b.x, b.y}; r.v[3] = {b.z, b.w};
    return r;
}
__device__ __forceinline__ J4<double> jload(const cx<double> *p) {
    const double2 *q = reinterpret_cast<const double2 *>(p);
    J4<double> r;
#pragma unroll
    for (int i = 0; i < 4; ++i) { double2 a = q[i]; r.v[i] = {a.x, a.y}; }
    return r;
}
__device__ __forceinline__ void jstore(cx<float> *p, const J4<float> &J) {
    float4 *q = reinterpret_cast<float4 *>(p);
    q[0] = make_float4(J.v[0].re, J.v[0].im, J.v[1].re, J.v[1].im);
    q[1] = make_float4(J.v[2].re, J.v[2].im, J.v[3].re, J.v[3].im);
}
__device__ __forceinline__ void jstore(cx<double> *p, const J4<double> &J) {
    double2 *q = reinterpret_cast<double2 *>(p);
#pragma unroll
    for (int i = 0; i < 4; ++i) q[i] = make_double2(J.v[i].re, J.v[i].im);
}

// Active extent of a dimension (upper - lower), the reference's DEXT().
__host__ __device__ __forceinline__ uint32_t dext(const mbx_dim &x) {
    return x.upper_extent - x.lower_extent;
}

// Brightness square root of one (source, time, channel), in T.
// B = [[I+Q, U+iV], [U-iV, I-Q]] * (f/rf)^alpha; sqrt(B) = (B + s 1)/t with
// s = sqrt(det B), t = sqrt(trace B + 2 s); the matrix is left undivided
// when t == 0 (then B == 0). CPUSolver.py:235-317, brightness.cuh:108-164.
// 1 / sqrt(x) for x > 0: the hardware reciprocal square root, refined by
// Newton steps y <- y (3 - x y^2) / 2 (one in float, two in float64, each
// doubling the correct bits), cheaper than a divide by a square root.
__device__ __forceinline__ float rsqrt_acc(float x) {
    const float y = __builtin_amdgcn_rsqf(x);
    return y * fmaf(-0.5f * x * y, y, 1.5f);
}
__device__ __forceinline__ double rsqrt_acc(double x) {
    double y = __builtin_amdgcn_rsq(x);
    y = y * fma(-0.5 * x * y, y, 1.5);
    return y * fma(-0.5 * x * y, y, 1.5);
}

template <typename T>
__device__ __forceinline__ J4<T> brightness_sqrt_scaled(T I, T Q, T U, T V, T p) {
    I *= p; Q *= p; U *= p; V *= p;
    T s = sqrt(I * I - Q * Q - U * U - V * V);
    T t = sqrt(T(2) * I + T(2) * s);
    J4<T> B;
    B.v[0] = {I + Q + s, T(0)};
    B.v[1] = {U, V};
    B.v[2] = {U, -V};
    B.v[3] = {I - Q + s, T(0)};
    if (t != T(0)) {
        T r = T(1) / t;
        B = jscale(B, r);
    }
    return B;
}

template <typename T>
__device__ __forceinline__ J4<T> brightness_sqrt(T I, T Q, T U, T V, T alpha, T f, T rf) {
    return brightness_sqrt_scaled<T>(I, Q, U, V, pow(f / rf, alpha));
}

// Scalar phase term exp(i*phi), phi = -2 pi f / C (u l + v m + w (n - 1)),
// evaluated in float64 whatever T is: |uvw| f / C reaches ~1e4 cycles, far
// beyond what a float32 argument resolves. RimeEKBSqrt.py:150-161.
__device__ __forceinline__ void phase_sincos(double u, double v, double w, double l,
                                             double m, double nm1, double f,
                                             double two_pi_over_c, double *s, double *c) {
    double p = w * nm1 + v * m + u * l;
    double phi = -two_pi_over_c * f * p;
    sincos(phi, s, c);
}

// Host-side validation shared by the entry points.
int set_error(int code, const char *fmt, ...);
int check_dims(const mbx_dims *d, bool need_beam);
int hip_status(hipError_t e, const char *where);

}  // namespace mbx
