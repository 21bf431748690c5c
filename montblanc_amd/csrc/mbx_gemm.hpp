// Point-source coherency sums as a Hermitian matrix product on the matrix
// cores (mbx_gemm.hip).
#pragma once

#include "mbx_tile.hpp"

namespace mbx {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

// v_mfma_*_16x16x4: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16].
// Output element r of lane l is, in float, D[4 (l / 16) + r][l % 16] and,
// in float64, D[4 r + l / 16][l % 16]. For float64 the A rows are loaded
// permuted (lane i of a group supplies the tile's row arow(i)) so that
// both types end with the same layout: element r of lane l is tile row
// 4 (l / 16) + r, i.e. two antennas (rows 2p, 2p + 1) per lane.
template <typename T> struct Mfma;
template <> struct Mfma<float> {
    typedef f32x4 acc_t;
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int arow(int i) { return i; }
};
template <> struct Mfma<double> {
    typedef f64x4 acc_t;
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int arow(int i) { return 4 * (i & 3) + (i >> 2); }
};
// tile row of output element r of a lane (both types, see above)
__device__ __forceinline__ int out_row(int lane, int r) { return 4 * (lane >> 4) + r; }


// LDS a workgroup may take on the current device (160 KB on gfx950).
constexpr size_t kGemmLdsMax = 160 * 1024;
inline size_t gemm_lds_budget() {
    int dev = 0, bytes = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        bytes <= 0)
        return 64 * 1024;
    return bytes < int(kGemmLdsMax) ? size_t(bytes) : kGemmLdsMax;
}

// Source parameters (l, m, n - 1) and sqrt(B) (f/f_ref)^alpha of source s
// at time t for the workgroup's channel record Lch.
template <typename T>
__device__ __forceinline__ J4<T> source_sqrtb(const RimeArgs<T> &A, const double *Lch, int t, int s,
                                             double &l0, double &m0, double &nm1) {
    const uint64_t NT = A.d.ntime.local_size;
    l0 = A.lm[2 * uint64_t(s)]; m0 = A.lm[2 * uint64_t(s) + 1];
    nm1 = sqrt(1.0 - l0 * l0 - m0 * m0) - 1.0;
    const T *S = A.stokes + (uint64_t(s) * NT + t) * 4;
    const T power = exp_t<T>(double(A.alpha[uint64_t(s) * NT + t]) * Lch[4]);
    return brightness_sqrt_scaled<T>(S[0], S[1], S[2], S[3], power);
}

// Per-source record of a source tile, computed once per source (by one
// thread, a tile ahead) and read by every producer thread that builds one
// of the source's per-antenna Jones: l, m, n - 1, the shape-coordinate
// coefficients (x = cu u + cv v, y = du u + dv v; zero for point sources)
// and the kind (0 point, 1 gaussian, 2 sersic) in float64, then sqrt(B) of
// the workgroup's channel c.
template <typename T> struct alignas(16) SrcRec {
    double l0, m0, nm1, cu, cv, du, dv, kind;
    J4<T> B;
};

// The record of source s at time t and channel c. gauss_pre scales the
// gaussian coordinates (sqrt(log2 e) where exp2 evaluates the factor).
// Reads the channel's frequencies from global memory (usable before the
// channel record is staged in LDS); the arithmetic is that of
// source_sqrtb with tile_setup's channel record.
template <typename T>
__device__ __forceinline__ void source_rec(const RimeArgs<T> &A, int t, int c, int s,
                                           double gauss_pre, SrcRec<T> &r) {
    const mbx_dims &d = A.d;
    const uint64_t NT = d.ntime.local_size;
    const int np = dext(d.npsrc), ng = dext(d.ngsrc);
    const uint64_t NGS = d.ngsrc.local_size, NSS = d.nssrc.local_size;
    const double f = A.freq[c];
    const double lf = log(f / double(A.ref_freq[c]));
    const double l0 = A.lm[2 * uint64_t(s)], m0 = A.lm[2 * uint64_t(s) + 1];
    r.l0 = l0; r.m0 = m0;
    r.nm1 = sqrt(1.0 - l0 * l0 - m0 * m0) - 1.0;
    const T *S = A.stokes + (uint64_t(s) * NT + t) * 4;
    const T power = exp_t<T>(double(A.alpha[uint64_t(s) * NT + t]) * lf);
    r.B = brightness_sqrt_scaled<T>(S[0], S[1], S[2], S[3], power);
    r.cu = r.cv = r.du = r.dv = 0.0;
    r.kind = 0.0;
    if (s >= np && s < np + ng) {
        const int gi = s - np;
        const double el = A.gauss[gi], em = A.gauss[gi + NGS], R = A.gauss[gi + 2 * NGS];
        const double k = A.gauss_scale * f * gauss_pre;
        r.cu = k * R * em; r.cv = -k * R * el; r.du = k * el; r.dv = k * em;
        r.kind = 1.0;
    } else if (s >= np + ng) {
        const int si = s - np - ng;
        const double e1 = A.sersic[si], e2 = A.sersic[si + NSS];
        const double Rp = A.sersic[si + 2 * NSS] / (1.0 - e1 * e1 - e2 * e2);
        const double k = A.two_pi_over_c * f * Rp;
        r.cu = k * (1.0 + e1); r.cv = k * e2; r.du = k * e2; r.dv = k * (1.0 - e1);
        r.kind = 2.0;
    }
}

// J = E K sqrt(B) of one (source, antenna a) for the workgroup's channel c
// from LDS-staged per-antenna records: uvw [3], parallactic sin/cos [2],
// pointing error and scaling [4], channel record Lch.
template <typename T, int UT>
__device__ __forceinline__ J4<T> staged_jones(const RimeArgs<T> &A, const double *Luvw,
                                              const double *Lpa, const double *Lant,
                                              const double *Lch, int c, int a, double l0,
                                              double m0, double nm1, const J4<T> &B) {
    const double *an = Lant + 4 * a;
    const double *x = Luvw + 3 * a;
    if constexpr (UT == 4) {
        // the four beam cells are loaded first and the phase evaluated
        // while they are in flight
        const BeamTapAddr<T> ta =
            beam_taps<T>(A, Lpa[2 * a], Lpa[2 * a + 1], an[0], an[1], an[2], an[3], c, l0, m0);
        BeamCell<T> e[4];
#pragma unroll
        for (int tap = 0; tap < 4; ++tap) e[tap] = ta.cell(tap);
        T sn, cs;
        PhaseCS<T>::eval(-A.two_pi_over_c * Lch[3] * (x[2] * nm1 + x[1] * m0 + x[0] * l0), &sn, &cs);
        const cx<T> K = {cs, sn};
        J4<T> KB;
#pragma unroll
        for (int q = 0; q < 4; ++q) KB.v[q] = cmul(K, B.v[q]);
        cx<T> acc[4];
        T amp[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) { acc[p] = {T(0), T(0)}; amp[p] = T(0); }
#pragma unroll
        for (int tap = 0; tap < 4; ++tap) beam_accumulate(e[tap], tap, ta.ld, ta.md, acc, amp);
        return jmul(beam_normalise(acc, amp), KB);
    } else {
        const J4<T> E = beam_jones_at<T, UT>(A, Lpa[2 * a], Lpa[2 * a + 1], an[0], an[1], an[2], an[3],
                                             c, l0, m0);
        T sn, cs;
        PhaseCS<T>::eval(-A.two_pi_over_c * Lch[3] * (x[2] * nm1 + x[1] * m0 + x[0] * l0), &sn, &cs);
        const cx<T> K = {cs, sn};
        J4<T> KB;
#pragma unroll
        for (int q = 0; q < 4; ++q) KB.v[q] = cmul(K, B.v[q]);
        return jmul(E, KB);
    }
}

// Per-antenna pointing error and scaling of time t, channel c into LDS.
template <typename T>
__device__ __forceinline__ void stage_antennas(const RimeArgs<T> &A, double *Lant, int t, int c,
                                               int tid, int nth) {
    const mbx_dims &d = A.d;
    const int NA = d.na.local_size;
    const uint64_t NC = d.nchan.local_size;
    for (int a = tid; a < NA; a += nth) {
        const T *pet = A.pe + ((uint64_t(t) * NA + a) * NC + c) * 2;
        const T *asc = A.ascale + (uint64_t(a) * NC + c) * 2;
        Lant[4 * a] = double(pet[0]); Lant[4 * a + 1] = double(pet[1]);
        Lant[4 * a + 2] = double(asc[0]); Lant[4 * a + 3] = double(asc[1]);
    }
}

template <typename T> struct GemmArgs {
    const int32_t *bmap;   // (na x na) pair -> baseline map, device
    cx<T> *dst;            // running source sum, (ntime, nbl, nchan, 4)
    const cx<T> *vin;      // sum to add to (may alias dst) or NULL
    int rows, st, ntp, tpw;
};

// True when the antenna count fits the kernel's tile assignment.
template <typename T> bool point_gemm_supported(const mbx_dims &d);

// Whether the gaussian and sersic sources are also worth the matrix cores.
template <typename T> bool shaped_gemm_pays(const mbx_dims &d);

// dst[t, b, c] = (vin ? vin[t, b, c] : 0) + sum over the point sources
// [0, npsrc extent) of J_p J_q^H -- and with `shaped` over the gaussian
// and sersic sources too, f_pq J_p J_q^H -- for every baseline of the map.
//
// With `epilogue_wgs` set, a float shaped pass (k_shaped_mx, or with point
// sources both passes in one kernel, k_forward_mx, whose point sums stay
// in LDS and never reach dst) also applies the epilogue (gains, flags,
// model or residual, chi^2 terms; A's output arrays) and writes one chi^2
// partial per workgroup to A.partial; the number of workgroups
// (gemm_workgroups) is stored there, 0 when no pass applied the epilogue.
template <typename T>
int launch_point_gemm(const RimeArgs<T> &A, const int32_t *bmap, cx<T> *dst, const cx<T> *vin,
                      bool shaped, hipStream_t stream, uint64_t *epilogue_wgs = nullptr);

// Workgroups of the matrix-core passes (one per time and channel, padded
// to the XCD order).
uint64_t gemm_workgroups(const mbx_dims &d);

// Forward of nrep samples (float; A.span set with wgs =
// gemm_workgroups(A.d)): one launch of the fused kernel where it fits
// (else the point sources of every sample into its slot of buf -- nrep
// slots of A.span.vstride complex values -- then the gaussian and sersic
// sources on top), with the epilogue: chi^2 partials per slot, A's model
// / chi^2 term arrays where set, and the pre-gain sums where A.ekb is set
// (to buf, slot by slot). any_point: some sample has point sources.
int launch_batched_mx(const RimeArgs<float> &A, const int32_t *bmap, cx<float> *buf, bool any_point,
                      unsigned nrep, hipStream_t stream);
bool batched_mx_supported(const mbx_dims &d);

}  // namespace mbx
