// Point-source coherency sums on the matrix cores (gfx950 MFMA).
//
// For one time t and channel c the point sources' part of every
// baseline's coherency sum (shape factor 1; CPUSolver.py:562-634,
// RimeSumCoherencies.py:236-305) is one Hermitian matrix product. Stack
// the per-antenna Jones of the S point sources as the 2na x 2S complex
// matrix X, X[2p+i][2s+j] = J_ps[i][j]; then
//     (X X^H)[2p+i][2q+k] = sum_s (J_ps J_qs^H)[i][k] = V_pq[i][k].
// In real arithmetic, with A = [Re X | Im X] and A' = [Im X | -Re X]
// (2na x 4S each):  Re V = A A^T,  Im V = A' A^T.
//
// Kernel k_point_gemm: one workgroup per (t, c), eight waves.
//   producer waves (4-7): per-antenna Jones of the next tile of `st`
//     sources into one of two LDS buffers, laid out X[s][row] with
//     row = 2p + i and the four reals (Re J_i0, Re J_i1, Im J_i0, Im J_i1);
//   consumer waves (0-3): own 16x16 output tiles (R, C), R <= C, of both
//     Re V and Im V ("tile pairs") in MFMA accumulators and step through
//     the other buffer four sources at a time with
//     v_mfma_{f32,f64}_16x16x4: lane group g = lane / 16 supplies source
//     s0 + g, its float4 of reals being the four k-slices of four MFMAs
//     (the order of the k sum is free, so this keeps every LDS read a
//     single 16/32-byte access).
// Only the upper triangle of tiles is computed (V_qp = V_pq^H). The
// epilogue scatters each (p <= q) element to its baseline through the
// host-built pair -> baseline map (reversed baselines get the conjugate
// transpose) into the running-sum buffer the fused kernel then finishes
// (other source types, gains, flags, chi^2).

#include <algorithm>

#include "mbx_gemm.hpp"

namespace mbx {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
template <> struct Mfma<float> {
    typedef f32x4 acc_t;
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // output element r of a lane: row 4 * (lane / 16) + r, column lane % 16
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};
template <> struct Mfma<double> {
    typedef f64x4 acc_t;
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // output element r of a lane: row 4 * r + lane / 16, column lane % 16
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * r + (lane >> 4); }
};

// One row of X for one source: Re J_i0, Re J_i1, Im J_i0, Im J_i1.
template <typename T> struct alignas(4 * sizeof(T)) XRow { T v[4]; };

// Four consumer waves (one per SIMD, each with its matrix core) and the
// producer waves the register file leaves room for: 12 in float (1024
// threads, 128 VGPRs each), 4 in float64 (512 threads, whose consumers
// hold 144 VGPRs of accumulators). The Jones producers are bound by the
// latency of the beam gathers, so more of them overlap more.
template <typename T> constexpr int gemm_threads() { return sizeof(T) == 4 ? 1024 : 512; }
constexpr int kGemmConsumers = 256;
constexpr int kGemmMaxTp = 9;         // tile pairs per consumer wave
constexpr int kGemmMaxRowBlocks = 8;  // 4 * 9 >= 8 * 9 / 2: na <= 64
constexpr size_t kGemmLdsMax = 160 * 1024;
// beam taps a producer keeps in flight
template <typename T> constexpr int kGemmTaps() { return sizeof(T) == 4 ? 4 : 2; }

static inline int gemm_rows(int na) { return ((2 * na + 15) / 16) * 16; }

// LDS: per antenna uvw [3], parallactic sin/cos [2], pointing error and
// scaling [4]; the channel record [CH]; then the two X buffers (32-byte
// aligned).
static __host__ __device__ inline size_t gemm_x_offset(int na) {
    return ((sizeof(double) * (9 * size_t(na) + CH)) + 31) & ~size_t(31);
}

template <typename T>
static inline size_t gemm_lds_bytes(int na, int st) {
    return gemm_x_offset(na) + 2 * size_t(st) * gemm_rows(na) * sizeof(XRow<T>);
}

template <typename T>
static inline int gemm_st(int na, size_t budget) {
    int st = 32;
    while (st > 4 && gemm_lds_bytes<T>(na, st) > budget) st -= 4;
    return st;
}

// LDS a workgroup may take on the current device (160 KB on gfx950).
static size_t gemm_lds_budget() {
    int dev = 0, bytes = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        bytes <= 0)
        return 64 * 1024;
    return std::min<size_t>(size_t(bytes), kGemmLdsMax);
}

template <typename T>
__global__ void __launch_bounds__(gemm_threads<T>())
k_point_gemm(RimeArgs<T> A, GemmArgs<T> G) {
    constexpr int kGemmThreads = gemm_threads<T>();
    extern __shared__ __align__(32) unsigned char smem_raw[];
    const mbx_dims &d = A.d;
    const int NA = d.na.local_size;
    const int nt_ext = dext(d.ntime), nc_ext = dext(d.nchan);
    const int tid = threadIdx.x;
    // Time fastest: the dispatcher deals consecutive workgroups to the
    // eight XCDs in turn, so the workgroups one XCD runs together share a
    // channel and with it the two beam-cube frequency planes they gather
    // from (L2 per XCD).
    const uint64_t w = blockIdx.x;
    if (w >= uint64_t(nt_ext) * nc_ext) return;
    const int t = int(w % nt_ext), c = int(w / nt_ext);

    const int rows = G.rows, st = G.st;
    double *Luvw = reinterpret_cast<double *>(smem_raw);
    double *Lpa = Luvw + 3 * NA;
    double *Lant = Lpa + 2 * NA;
    double *Lch = Lant + 4 * NA;
    XRow<T> *X = reinterpret_cast<XRow<T> *>(smem_raw + gemm_x_offset(NA));
    const int xstride = st * rows;   // one buffer

    tile_setup<T>(A, Luvw, Lpa, Lch, t, c, tid, kGemmThreads);
    {
        const uint64_t NC = d.nchan.local_size;
        const int cc = min(c, nc_ext - 1);
        for (int a = tid; a < NA; a += kGemmThreads) {
            const T *pet = A.pe + ((uint64_t(t) * NA + a) * NC + cc) * 2;
            const T *asc = A.ascale + (uint64_t(a) * NC + cc) * 2;
            Lant[4 * a] = double(pet[0]); Lant[4 * a + 1] = double(pet[1]);
            Lant[4 * a + 2] = double(asc[0]); Lant[4 * a + 3] = double(asc[1]);
        }
    }
    // rows past 2 na stay zero in both buffers
    for (int i = tid; i < 2 * st * (rows - 2 * NA); i += kGemmThreads) {
        const int per = rows - 2 * NA;
        const int bs = i / per;           // buffer * st + source
        X[bs * rows + 2 * NA + i % per] = XRow<T>{{T(0), T(0), T(0), T(0)}};
    }
    __syncthreads();

    const int np = dext(d.npsrc);
    const int ntiles = (np + st - 1) / st;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    if (wave < kGemmConsumers / 64) {
        typedef typename Mfma<T>::acc_t acc_t;
        const int lane = tid & 63, i16 = lane & 15, g = lane >> 4;
        // this wave's tile pairs: [wave * tpw, wave * tpw + tpw) in the
        // row-major order of (R, C >= R)
        int tpR[kGemmMaxTp], tpC[kGemmMaxTp];
        const int nR = rows / 16;
        {
            int idx = wave * G.tpw;
            int R = 0;
            while (R < nR && idx >= nR - R) { idx -= nR - R; ++R; }
            int C = R + idx;
#pragma unroll
            for (int j = 0; j < kGemmMaxTp; ++j) {
                tpR[j] = R; tpC[j] = C;
                if (++C >= nR) { ++R; C = R; }
            }
        }
        const int ntp_mine = max(0, min(G.tpw, G.ntp - wave * G.tpw));
        acc_t are[kGemmMaxTp], aim[kGemmMaxTp];
#pragma unroll
        for (int j = 0; j < kGemmMaxTp; ++j) {
            are[j] = acc_t{T(0), T(0), T(0), T(0)};
            aim[j] = acc_t{T(0), T(0), T(0), T(0)};
        }
        __syncthreads();                      // tile 0 produced
        for (int it = 0; it < ntiles; ++it) {
            const XRow<T> *Xb = X + (it & 1) * xstride;
            const int nst = min(st, np - it * st);
            for (int s0 = 0; s0 < nst; s0 += 4) {
                const XRow<T> *Xs = Xb + (s0 + g) * rows + i16;
#pragma unroll
                for (int j = 0; j < kGemmMaxTp; ++j) {
                    if (j < ntp_mine) {
                        const XRow<T> a = Xs[16 * tpR[j]];
                        const XRow<T> b = Xs[16 * tpC[j]];
                        const T ai[4] = {a.v[2], a.v[3], -a.v[0], -a.v[1]};
#pragma unroll
                        for (int m = 0; m < 4; ++m) {
                            are[j] = Mfma<T>::mma(a.v[m], b.v[m], are[j]);
                            aim[j] = Mfma<T>::mma(ai[m], b.v[m], aim[j]);
                        }
                    }
                }
            }
            __syncthreads();
        }
        // epilogue: scatter the upper-triangle elements to their baselines
        const uint64_t NC = d.nchan.local_size, NBL = d.nbl.local_size;
#pragma unroll
        for (int j = 0; j < kGemmMaxTp; ++j) {
            if (j >= ntp_mine) continue;
            const int col = 16 * tpC[j] + i16;
            const int q = col >> 1, kq = col & 1;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * tpR[j] + Mfma<T>::row(lane, r);
                const int p = row >> 1, ip = row & 1;
                if (p >= NA || q >= NA || p > q) continue;
                const int e = G.bmap[p * NA + q];
                if (e < 0) continue;
                const bool rev = (e & MBX_BL_REVERSED) != 0;
                const int b = e & ~(MBX_BL_REVERSED | MBX_BL_AUTO);
                if (uint64_t(b) >= NBL) continue;   // a map for another layout
                cx<T> v = {are[j][r], rev ? -aim[j][r] : aim[j][r]};
                const int corr = rev ? 2 * kq + ip : 2 * ip + kq;
                const uint64_t vi = ((uint64_t(t) * NBL + b) * NC + c) * 4 + corr;
                if (G.vin) { const cx<T> o = G.vin[vi]; v.re += o.re; v.im += o.im; }
                G.dst[vi] = v;
            }
        }
    } else {
        const int ptid = tid - kGemmConsumers, pth = kGemmThreads - kGemmConsumers;
        const uint64_t NT = d.ntime.local_size;
        // A producer thread owns one source of the tile and a stride of
        // antennas: the source's parameters and sqrt(B) are loaded once,
        // and each Jones then needs only the beam taps from memory.
        auto produce = [&](int buf, int s0, int nst) {
            XRow<T> *Xb = X + buf * xstride;
            const int astep = pth / nst;
            const int si = ptid % nst, a0 = ptid / nst;
            if (a0 < astep) {
                const int s = s0 + si;
                const double l0 = A.lm[2 * uint64_t(s)], m0 = A.lm[2 * uint64_t(s) + 1];
                const double nm1 = sqrt(1.0 - l0 * l0 - m0 * m0) - 1.0;
                const T *S = A.stokes + (uint64_t(s) * NT + t) * 4;
                const T power = exp_t<T>(double(A.alpha[uint64_t(s) * NT + t]) * Lch[4]);
                const J4<T> B = brightness_sqrt_scaled<T>(S[0], S[1], S[2], S[3], power);
                const double kf = -A.two_pi_over_c * Lch[3];
                for (int a = a0; a < NA; a += astep) {
                    const double *an = Lant + 4 * a;
                    const J4<T> E = beam_jones_at<T, kGemmTaps<T>()>(A, Lpa[2 * a], Lpa[2 * a + 1], an[0],
                                                                    an[1], an[2], an[3], c, l0, m0);
                    const double *x = Luvw + 3 * a;
                    T sn, cs;
                    PhaseCS<T>::eval(kf * (x[2] * nm1 + x[1] * m0 + x[0] * l0), &sn, &cs);
                    const cx<T> K = {cs, sn};
                    J4<T> KB;
#pragma unroll
                    for (int q = 0; q < 4; ++q) KB.v[q] = cmul(K, B.v[q]);
                    const J4<T> J = jmul(E, KB);
                    Xb[si * rows + 2 * a] = XRow<T>{{J.v[0].re, J.v[1].re, J.v[0].im, J.v[1].im}};
                    Xb[si * rows + 2 * a + 1] = XRow<T>{{J.v[2].re, J.v[3].re, J.v[2].im, J.v[3].im}};
                }
            }
            // sources past the tile's end up to a multiple of four read as zero
            const int pad = ((nst + 3) & ~3) - nst;
            for (int i = ptid; i < pad * rows; i += pth)
                Xb[(nst + i / rows) * rows + i % rows] = XRow<T>{{T(0), T(0), T(0), T(0)}};
        };
        if (ntiles > 0) produce(0, 0, min(st, np));
        __syncthreads();
        for (int it = 0; it < ntiles; ++it) {
            const int s1 = (it + 1) * st;
            if (it + 1 < ntiles) produce((it + 1) & 1, s1, min(st, np - s1));
            __syncthreads();
        }
    }
}

template <typename T>
bool point_gemm_supported(const mbx_dims &d) {
    const int na = d.na.local_size;
    return dext(d.npsrc) > 0 && na > 0 && gemm_rows(na) / 16 <= kGemmMaxRowBlocks;
}

template <typename T>
int launch_point_gemm(const RimeArgs<T> &A, const int32_t *bmap, cx<T> *dst, const cx<T> *vin,
                      hipStream_t stream) {
    const mbx_dims &d = A.d;
    const int na = d.na.local_size;
    GemmArgs<T> G;
    G.bmap = bmap; G.dst = dst; G.vin = vin;
    G.rows = gemm_rows(na);
    const int nR = G.rows / 16;
    G.ntp = nR * (nR + 1) / 2;
    G.tpw = (G.ntp + kGemmConsumers / 64 - 1) / (kGemmConsumers / 64);
    if (G.tpw > kGemmMaxTp) return set_error(MBX_E_UNSUPPORTED, "point gemm: %d antennas", na);
    static size_t budget = gemm_lds_budget();
    G.st = gemm_st<T>(na, budget);
    const uint64_t nwg = uint64_t(dext(d.ntime)) * dext(d.nchan);
    RimeArgs<T> Ag = A;
    Ag.ct = 1;
    const size_t lds = gemm_lds_bytes<T>(na, G.st);
    hipLaunchKernelGGL(k_point_gemm<T>, dim3(unsigned(nwg)), dim3(gemm_threads<T>()), lds, stream,
                       Ag, G);
    return hip_status(hipGetLastError(), "point gemm launch");
}

template bool point_gemm_supported<float>(const mbx_dims &);
template bool point_gemm_supported<double>(const mbx_dims &);
template int launch_point_gemm<float>(const RimeArgs<float> &, const int32_t *, cx<float> *,
                                      const cx<float> *, hipStream_t);
template int launch_point_gemm<double>(const RimeArgs<double> &, const int32_t *, cx<double> *,
                                       const cx<double> *, hipStream_t);

}  // namespace mbx
