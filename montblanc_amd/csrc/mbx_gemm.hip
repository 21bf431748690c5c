// Coherency sums on the matrix cores (gfx950 MFMA): point sources in
// k_point_gemm<T, false>; gaussian and sersic sources in float64 in
// k_point_gemm<T, true> (the 16x16 products scaled per element by the
// shape factor) and in float in k_shaped_mx (4x4-block products with the
// factor on the B operand; it also applies the epilogue), both described
// at their kernels. In float k_forward_mx runs the point pass and the
// shaped pass one after the other in one workgroup, the point sums handed
// over in LDS. Every kernel has a BATCHED instance that covers several
// samples' grids in one launch (SampleSpan, mbx_tile.hpp).
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
// Kernel k_point_gemm: one workgroup per (t, c); four consumer waves and
// twelve (float) or four (float64) producer waves.
//   producer waves: per-antenna Jones of the next tile of `st`
//     sources into one of two LDS buffers, laid out X[s][row] with
//     row = 2p + i and the four reals (Re J_i0, Re J_i1, Im J_i0, Im J_i1);
//   consumer waves (waves 0-3): own 16x16 output tiles (R, C), R <= C, of both
//     Re V and Im V ("tile pairs") in MFMA accumulators and step through
//     the other buffer four sources at a time with
//     v_mfma_{f32,f64}_16x16x4: lane group g = lane / 16 supplies source
//     s0 + g, its float4 of reals being the four k-slices of four MFMAs
//     (the order of the k sum is free, so this keeps every LDS read a
//     single 16/32-byte access).
// Only the upper triangle of tiles is computed (V_qp = V_pq^H). The
// epilogue scatters each (p <= q) element to its baseline through the
// host-built pair -> baseline map (reversed baselines get the conjugate
// transpose) into the running-sum buffer that a later pass finishes (other
// source types, gains, flags, chi^2) -- or, inside k_forward_mx, into LDS.

#include <algorithm>

#include "mbx_fmath.hpp"
#include "mbx_gemm.hpp"

namespace mbx {

// One row of X for one source: Re J_i0, Re J_i1, Im J_i0, Im J_i1.
template <typename T> struct alignas(4 * sizeof(T)) XRow { T v[4]; };

// Wave layout per pass. Point pass: four consumer waves (one per SIMD and
// its matrix core) own up to nine tile pairs each; producers fill the
// workgroup -- 12 in float (1024 threads, 128 VGPRs each), 4 in float64
// (512 threads, 256 VGPRs: the consumers hold 144 of accumulators).
// Shaped pass, whose consumers also evaluate a shape factor per element:
// in float eight consumer waves with up to five tile pairs and 8 producers;
// in float64 twelve consumer waves with up to three (three per SIMD, so
// that the latency of each one's factor arithmetic hides behind the
// others') and 4 producers; 1024 threads, 128 VGPRs, in both.
template <typename T, bool SHAPED> constexpr int gemm_threads() {
    return sizeof(T) == 4 || SHAPED ? 1024 : 512;
}
template <typename T, bool SHAPED> constexpr int gemm_consumers() {
    return SHAPED ? (sizeof(T) == 8 ? 768 : 512) : 256;
}
template <typename T, bool SHAPED> constexpr int gemm_max_tp() {
    return SHAPED ? (sizeof(T) == 8 ? 3 : 5) : 9;
}
constexpr int kGemmMaxRowBlocks = 8;  // 4 * 9, 8 * 5 and 12 * 3 >= 8 * 9 / 2: na <= 64
// beam taps a producer keeps in flight
template <typename T, bool SHAPED> constexpr int kGemmTaps() { return sizeof(T) == 4 || !SHAPED ? 4 : 2; }

static inline int gemm_rows(int na) { return ((2 * na + 15) / 16) * 16; }

// LDS: per antenna uvw [3], parallactic sin/cos [2], pointing error and
// scaling [4]; the channel record [CH]; then the two X buffers (32-byte
// aligned).
static __host__ __device__ inline size_t gemm_x_offset(int na) {
    return ((sizeof(double) * (9 * size_t(na) + CH)) + 31) & ~size_t(31);
}

// Then the two source-record buffers, and in the shaped pass the shape
// coordinates and kinds of both tiles (and in float64 the exp_neg table).
template <typename T, bool SHAPED>
static inline size_t gemm_lds_bytes(int na, int st) {
    return gemm_x_offset(na) + 2 * size_t(st) * gemm_rows(na) * sizeof(XRow<T>) +
           2 * size_t(st) * sizeof(SrcRec<T>) +
           (SHAPED ? 2 * 2 * size_t(na) * st * sizeof(T) + 2 * size_t(st) * sizeof(int) : 0) +
           (SHAPED && sizeof(T) == 8 ? kExpTab * sizeof(double) : 0);
}

template <typename T, bool SHAPED>
static inline int gemm_st(int na, size_t budget) {
    int st = 32;
    while (st > 4 && gemm_lds_bytes<T, SHAPED>(na, st) > budget) st -= 4;
    return st;
}

// Shape factor of one (p, q) element from the per-antenna shape
// coordinates (x, y): gaussian exp(-(dx^2 + dy^2)), the coordinates
// pre-scaled by sqrt(log2 e) in float so that exp2 applies; sersic
// (1 + dx^2 + dy^2)^(-3/2).
template <typename T> __device__ __forceinline__ T shape_f(bool gauss, T dx, T dy);
template <> __device__ __forceinline__ float shape_f<float>(bool gauss, float dx, float dy) {
    const float r2 = fmaf(dx, dx, dy * dy);
    if (gauss) return __builtin_amdgcn_exp2f(-r2);
    const float rs = __builtin_amdgcn_rsqf(1.0f + r2);
    return rs * rs * rs;
}
template <> __device__ __forceinline__ double shape_f<double>(bool gauss, double dx, double dy) {
    const double r2 = fma(dx, dx, dy * dy);
    if (gauss) return exp(-r2);
    const double den = 1.0 + r2;
    return 1.0 / (den * sqrt(den));
}

// x with its sign bit flipped where m = 0x80000000 (m = 0: x).
__device__ __forceinline__ float flip_sign(float x, unsigned m) {
    return __builtin_bit_cast(float, __builtin_bit_cast(unsigned, x) ^ m);
}
__device__ __forceinline__ double flip_sign(double x, unsigned m) {
    return __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, x) ^
                                          (static_cast<unsigned long long>(m) << 32));
}
template <typename T> struct Vec2;
template <> struct Vec2<float> { typedef float2 type; };
template <> struct Vec2<double> { typedef double2 type; };

// The value of lane 4 (l / 4) + perm[l % 4] (DPP quad permutation, the
// four 2-bit lane selects packed in CTRL; no LDS).
template <int CTRL> __device__ __forceinline__ float quad_perm(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
template <int CTRL> __device__ __forceinline__ double quad_perm(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_mov_dpp(int(b), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(int(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// Index of the pair p <= q in a packed upper triangle of na antennas.
__host__ __device__ inline int stage_pair(int p, int q, int na) { return p * na - (p * (p - 1)) / 2 + (q - p); }

// Bytes of the staged point sums of every pair p <= q (2 x 2 complex).
template <typename T> static inline size_t stage_bytes(int na) {
    return size_t(na) * (na + 1) / 2 * 4 * sizeof(cx<T>);
}

// The point sources (SHAPED = false) or, in float64, the gaussian and
// sersic sources (true; float takes k_shaped_mx) of one (t, c), with the
// pass's wave roles over kGemmThreads threads, from the LDS set-up on.
// STAGE: the sums of every pair p <= q go to `stage` in LDS (packed upper
// triangle, stage_pair; the X buffers are dead by then, so the staging
// area may overlap them) instead of the baselines' running sums in G.dst.
template <typename T, bool SHAPED, bool STAGE>
__device__ __forceinline__ void point_phase(const RimeArgs<T> &A, const GemmArgs<T> &G,
                                            unsigned char *smem_raw, int t, int c, cx<T> *stage) {
    constexpr int kGemmThreads = gemm_threads<T, SHAPED>();
    constexpr int kGemmConsumers = gemm_consumers<T, SHAPED>();
    constexpr int kGemmMaxTp = gemm_max_tp<T, SHAPED>();
    const mbx_dims &d = A.d;
    const int NA = d.na.local_size;
    const int tid = threadIdx.x;

    const int rows = G.rows, st = G.st;
    double *Luvw = reinterpret_cast<double *>(smem_raw);
    double *Lpa = Luvw + 3 * NA;
    double *Lant = Lpa + 2 * NA;
    double *Lch = Lant + 4 * NA;
    // two source-tile buffers: Jones X[s][row] (four reals per row); two
    // source-record buffers; in the shaped pass the shape coordinates
    // XY[s][a] = (x, y) and source kinds of both tiles
    T *Xbase = reinterpret_cast<T *>(smem_raw + gemm_x_offset(NA));
    const int xstride = 4 * st * rows;       // T per Jones buffer
    SrcRec<T> *Pbase = reinterpret_cast<SrcRec<T> *>(Xbase + 2 * xstride);
    T *XYbase = reinterpret_cast<T *>(Pbase + 2 * st);
    const int xystride = 2 * NA * st;        // T per coordinate buffer
    int *Kbase = reinterpret_cast<int *>(XYbase + 2 * xystride);
    double *Ltab = reinterpret_cast<double *>(Kbase + 2 * st);   // st is a multiple of 4

    const int np = dext(d.npsrc), ns = dext(d.nsrc);
    const int sbeg = SHAPED ? np : 0, send = SHAPED ? ns : np;
    const int ntiles = (send - sbeg + st - 1) / st;
    auto tile_range = [&](int it, int &s0, int &nst) {
        s0 = sbeg + it * st;
        nst = min(st, send - s0);
    };
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    // source records of tile it into buffer it & 1, one producer thread per
    // source (they read no LDS: tiles 0 and 1 go before the set-up barrier)
    auto records = [&](int it) {
        const int ptid = tid - kGemmConsumers;
        int s0, nst;
        tile_range(it, s0, nst);
        if (ptid >= 0 && ptid < nst)
            source_rec<T>(A, t, c, s0 + ptid, sizeof(T) == 4 ? 1.2011224087864498 : 1.0,
                          Pbase[(it & 1) * st + ptid]);
    };
    if (wave >= kGemmConsumers / 64) {
        if (ntiles > 0) records(0);
        if (ntiles > 1) records(1);
    }

    tile_setup<T>(A, Luvw, Lpa, Lch, t, c, tid, kGemmThreads);
    if constexpr (SHAPED && sizeof(T) == 8) exp_table_fill(Ltab, tid, kGemmThreads);
    stage_antennas<T>(A, Lant, t, c, tid, kGemmThreads);
    // Jones rows past 2 na read as zero: clear both buffers once (point
    // and shaped tiles write rows < 2 na only)
    for (int i = tid; i < 2 * xstride; i += kGemmThreads) Xbase[i] = T(0);
    __syncthreads();

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
        // float64 point pass: three real products per complex one (Gauss),
        // are = P1 (Re Re), aim = P2 (Im Im), a3 = P3 ((Re + Im)(Re - Im)^T),
        // combined before the epilogue: Re V = P1 + P2, Im V = P3 - P1 + P2
        constexpr bool kGauss3 = sizeof(T) == 8 && !SHAPED;
        acc_t a3[kGauss3 ? kGemmMaxTp : 1];
#pragma unroll
        for (int j = 0; j < kGemmMaxTp; ++j) {
            are[j] = acc_t{T(0), T(0), T(0), T(0)};
            aim[j] = acc_t{T(0), T(0), T(0), T(0)};
            if constexpr (kGauss3) a3[j] = acc_t{T(0), T(0), T(0), T(0)};
        }
        // Im part of a shaped tile's product: lane group g supplies
        // component (g + 2) mod 4 of A' = [Im X | -Re X]
        const acc_t zero = acc_t{T(0), T(0), T(0), T(0)};
        // shaped pass, per tile pair j: byte offsets of this lane's A and B
        // reals in a source's rows (A: row 16 R + arow(i16), B: row 16 C +
        // i16, real g of each) and of the shape coordinates of its column
        // antenna q and row antenna pm. The lane's rows are antennas p0 =
        // 8 R + 2 g and p0 + 1 (rows past 2 na: any finite factor,
        // discarded); its neighbour (lane ^ 1, the other column of antenna
        // q) needs the same two factors: each evaluates one (pm = p0 + the
        // lane's parity) and takes the other across.
        int offA[SHAPED ? kGemmMaxTp : 1], offB[SHAPED ? kGemmMaxTp : 1];
        int offQ[SHAPED ? kGemmMaxTp : 1], offP[SHAPED ? kGemmMaxTp : 1];
        const int dA2 = (((g + 2) & 3) - g) * int(sizeof(T));
        const unsigned sgn = g < 2 ? 0u : 0x80000000u;
        // point pass, per tile pair j: byte offsets of this lane's A and B
        // rows (source g of a group of four)
        int offRa[SHAPED ? 1 : kGemmMaxTp], offCb[SHAPED ? 1 : kGemmMaxTp];
        if constexpr (!SHAPED) {
#pragma unroll
            for (int j = 0; j < kGemmMaxTp; ++j) {
                offRa[j] = (g * rows + 16 * tpR[j] + Mfma<T>::arow(i16)) * int(sizeof(XRow<T>));
                offCb[j] = (g * rows + 16 * tpC[j] + i16) * int(sizeof(XRow<T>));
            }
        }
        if constexpr (SHAPED) {
#pragma unroll
            for (int j = 0; j < kGemmMaxTp; ++j) {
                const int ra = 16 * tpR[j] + Mfma<T>::arow(i16), cb = 16 * tpC[j] + i16;
                offA[j] = (ra * 4 + g) * int(sizeof(T));
                offB[j] = (cb * 4 + g) * int(sizeof(T));
                offQ[j] = min(cb >> 1, NA - 1) * 2 * int(sizeof(T));
                offP[j] = min(8 * tpR[j] + 2 * g + (lane & 1), NA - 1) * 2 * int(sizeof(T));
            }
        }

        __syncthreads();                      // tile 0 produced
        for (int it = 0; it < ntiles; ++it) {
            const int buf = it & 1;
            int s0t, nst;
            tile_range(it, s0t, nst);
            if (!SHAPED) {
                const char *Xb = reinterpret_cast<const char *>(Xbase + buf * xstride);
                for (int s0 = 0; s0 < nst; s0 += 4) {
                    // the group's byte offset, opaque to the compiler (else
                    // each tile pair's two addresses become induction
                    // variables of their own, stepped per group)
                    int xo = s0 * rows * int(sizeof(XRow<T>));
                    __asm__ volatile("" : "+s"(xo));
                    const char *Xs = Xb + xo;
                    // float: A rows read once per row block of the wave's
                    // tile pairs (consecutive in its row-major share); float64
                    // reads them per tile pair (measured faster there)
                    XRow<T> a;
#pragma unroll
                    for (int j = 0; j < kGemmMaxTp; ++j) {
                        if (j < ntp_mine) {
                            if (sizeof(T) == 8 || j == 0 || tpR[j] != tpR[j - 1])
                                a = *reinterpret_cast<const XRow<T> *>(Xs + offRa[j]);
                            const XRow<T> b = *reinterpret_cast<const XRow<T> *>(Xs + offCb[j]);
                            if constexpr (kGauss3) {
#pragma unroll
                                for (int m = 0; m < 2; ++m) {
                                    are[j] = Mfma<T>::mma(a.v[m], b.v[m], are[j]);
                                    aim[j] = Mfma<T>::mma(a.v[m + 2], b.v[m + 2], aim[j]);
                                    a3[j] = Mfma<T>::mma(a.v[m] + a.v[m + 2], b.v[m] - b.v[m + 2], a3[j]);
                                }
                            } else {
                                const T ai[4] = {a.v[2], a.v[3], -a.v[0], -a.v[1]};
#pragma unroll
                                for (int m = 0; m < 4; ++m) {
                                    are[j] = Mfma<T>::mma(a.v[m], b.v[m], are[j]);
                                    aim[j] = Mfma<T>::mma(ai[m], b.v[m], aim[j]);
                                }
                            }
                            // float64 operands take twice the registers: one
                            // tile pair's loads in flight at a time
                            if constexpr (sizeof(T) == 8) __builtin_amdgcn_sched_barrier(0);
                        }
                    }
                }
            } else {
                // one source per MFMA (k = the four reals of its Jones row), each
                // product scaled per element by f_pq. Same source-major rows as
                // the point pass, so a lane's read of one real of a row is
                // conflict-free; shape coordinates (x, y) interleaved per
                // (source, antenna). The lane's byte offsets within a source's
                // rows and coordinates are fixed per tile pair (set up before
                // the tile loop); per source only the source's base is added.
                const char *Xg = reinterpret_cast<const char *>(Xbase + buf * xstride);
                const char *XY = reinterpret_cast<const char *>(XYbase + buf * xystride);
                // the tile's source kinds, one per lane (st <= 32), read per
                // source with v_readlane instead of an LDS round trip
                const int kinds = lane < nst ? Kbase[buf * st + lane] : 0;
                for (int s = 0; s < nst; ++s) {
                    // the source's byte offsets, opaque to the compiler: it
                    // would otherwise step every per-tile-pair address of
                    // the loop as its own induction variable (an addition
                    // per address per source on top of the one per read)
                    int xo = s * rows * 4 * int(sizeof(T)), yo = s * NA * 2 * int(sizeof(T));
                    __asm__ volatile("" : "+s"(xo), "+s"(yo));
                    const char *Xs = Xg + xo;
                    const char *XYs = XY + yo;
                    // the kind is the same for the whole wave (a uniform branch)
                    const bool gauss = __builtin_amdgcn_readlane(kinds, s) == 1;
                    // the A operands and row antenna coordinates: the same for
                    // the tile pairs of one row block (consecutive in the
                    // wave's row-major share), read once per row block
                    T av, a2f;
                    typename Vec2<T>::type cp;
#pragma unroll
                    for (int j = 0; j < kGemmMaxTp; ++j) {
                        if (j < ntp_mine) {
                            if (j == 0 || tpR[j] != tpR[j - 1]) {
                                av = *reinterpret_cast<const T *>(Xs + offA[j]);
                                a2f = flip_sign(*reinterpret_cast<const T *>(Xs + offA[j] + dA2), sgn);
                                cp = *reinterpret_cast<const typename Vec2<T>::type *>(XYs + offP[j]);
                            }
                            const T bv = *reinterpret_cast<const T *>(Xs + offB[j]);
                            const acc_t P = Mfma<T>::mma(av, bv, zero);
                            const acc_t Q = Mfma<T>::mma(a2f, bv, zero);
                            const typename Vec2<T>::type cq =
                                *reinterpret_cast<const typename Vec2<T>::type *>(XYs + offQ[j]);
                            const T dx = cq.x - cp.x, dy = cq.y - cp.y;
                            T fm;
                            if constexpr (sizeof(T) == 8) {
                                const T r2 = fma(dx, dx, dy * dy);
                                fm = gauss ? exp_neg(r2, Ltab) : rpow32(1.0 + r2);
                            } else {
                                fm = shape_f<T>(gauss, dx, dy);
                            }
                            // rows of p0 from the even lane, of p0 + 1 from the odd one
                            const T f0 = quad_perm<0xA0>(fm), f1 = quad_perm<0xF5>(fm);
                            are[j][0] = fma(f0, P[0], are[j][0]); aim[j][0] = fma(f0, Q[0], aim[j][0]);
                            are[j][1] = fma(f0, P[1], are[j][1]); aim[j][1] = fma(f0, Q[1], aim[j][1]);
                            are[j][2] = fma(f1, P[2], are[j][2]); aim[j][2] = fma(f1, Q[2], aim[j][2]);
                            are[j][3] = fma(f1, P[3], are[j][3]); aim[j][3] = fma(f1, Q[3], aim[j][3]);
                            // one tile pair at a time: bounds the live registers
                            __builtin_amdgcn_sched_barrier(0);
                        }
                    }
                }
            }
            __syncthreads();
        }
        if constexpr (kGauss3) {
#pragma unroll
            for (int j = 0; j < kGemmMaxTp; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const T p1 = are[j][r], p2 = aim[j][r];
                    are[j][r] = p1 + p2;
                    aim[j][r] = a3[j][r] - p1 + p2;
                }
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
                const int row = 16 * tpR[j] + out_row(lane, r);
                const int p = row >> 1, ip = row & 1;
                if (p >= NA || q >= NA || p > q) continue;
                if constexpr (STAGE) {
                    stage[stage_pair(p, q, NA) * 4 + 2 * ip + kq] = cx<T>{are[j][r], aim[j][r]};
                    continue;
                }
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
        // tile it into buffer it & 1 from its source records: one Jones per
        // (source, antenna) item, items taken in order by the producer
        // threads (consecutive threads take consecutive antennas of a
        // source: conflict-free 16/32-byte row writes)
        auto produce = [&](int it) {
            const int buf = it & 1;
            int s0, nst;
            tile_range(it, s0, nst);
            const int nst4 = (nst + 3) & ~3;
            XRow<T> *Xb = reinterpret_cast<XRow<T> *>(Xbase + buf * xstride);
            const SrcRec<T> *P = Pbase + buf * st;
            T *XY = XYbase + buf * xystride;
            int *kind = Kbase + buf * st;
            // item i = si NA + a; the stride pth advances (si, a) by (dsi, da)
            const int dsi = pth / NA, da = pth - dsi * NA;
            int si = ptid / NA, a = ptid - si * NA;
            for (; si < nst; si += dsi, a += da) {
                if (a >= NA) { a -= NA; ++si; }
                if (si >= nst) break;
                const SrcRec<T> &r = P[si];
                const J4<T> J = staged_jones<T, kGemmTaps<T, SHAPED>()>(A, Luvw, Lpa, Lant, Lch, c, a, r.l0,
                                                                r.m0, r.nm1, r.B);
                Xb[si * rows + 2 * a] = XRow<T>{{J.v[0].re, J.v[1].re, J.v[0].im, J.v[1].im}};
                Xb[si * rows + 2 * a + 1] = XRow<T>{{J.v[2].re, J.v[3].re, J.v[2].im, J.v[3].im}};
                if (SHAPED) {
                    // shape coordinates: linear in the antenna's (u, v)
                    const double u = Luvw[3 * a], v = Luvw[3 * a + 1];
                    XY[(si * NA + a) * 2] = T(r.cu * u + r.cv * v);
                    XY[(si * NA + a) * 2 + 1] = T(r.du * u + r.dv * v);
                    if (a == 0) kind[si] = int(r.kind);
                }
            }
            // point pass: sources past the tile's end up to a multiple of
            // four read as zero
            if (!SHAPED)
                for (int i = ptid; i < (nst4 - nst) * rows; i += pth)
                    Xb[(nst + i / rows) * rows + i % rows] = XRow<T>{{T(0), T(0), T(0), T(0)}};
        };
        // records of tile it + 2 go to buffer it & 1 while tile it + 1 is
        // produced: tile it's records were last read before the barrier
        // that ended the previous step
        if (ntiles > 0) produce(0);
        __syncthreads();
        for (int it = 0; it < ntiles; ++it) {
            if (it + 1 < ntiles) produce(it + 1);
            if (it + 2 < ntiles) records(it + 2);
            __syncthreads();
        }
    }
}

template <typename T, bool SHAPED, bool BATCHED = false>
__global__ void __launch_bounds__((gemm_threads<T, SHAPED>()))
k_point_gemm(RimeArgs<T> A, GemmArgs<T> G) {
    extern __shared__ __align__(32) unsigned char smem_raw[];
    // in a launch over several samples: this workgroup's sample and slot
    const SampleAt at = enter_sample_if<BATCHED>(A);
    if constexpr (BATCHED) {
        G.dst += uint64_t(at.j) * A.span.vstride;
        if (G.vin) G.vin += uint64_t(at.j) * A.span.vstride;
    }
    const int nt_ext = dext(A.d.ntime), nc_ext = dext(A.d.nchan);
    // XCD order of kChanGroup (mbx_tile.hpp)
    int t, c;
    if (!tc_coords(tc_per_xcd(nt_ext, nc_ext), nt_ext, nc_ext, t, c, at.bx)) return;
    point_phase<T, SHAPED, false>(A, G, smem_raw, t, c, nullptr);
}

// ---------------------------------------------------------------------
// Gaussian and sersic sources in float on the matrix cores, 4x4 blocks.
//
// The shape factor f_pq(s) differs per element of the coherency matrix,
// so a 16x16 product over the Jones reals has to be scaled element by
// element after every source. v_mfma_f32_4x4x1f32 instead runs sixteen
// independent 4x4 rank-1 updates: lane 4b + i supplies row i of block b's
// A operand, lane 4b + j column j of its B operand, and output register r
// of lane 4b + j is element (r, j) of block b. Block b of a consumer wave
// owns, for each of SETS sets g, the pairs (p_g, q) of one antenna p_g
// with an antenna quad q = 4Q + j. With the four reals of a Jones row and
// its rotation,
//     X row ia = (Re J_ia0, Re J_ia1, Im J_ia0, Im J_ia1),
//     R row ia = (Im J_ia0, Im J_ia1, -Re J_ia0, -Re J_ia1),
// lane i supplies (i < 2 ? X row i : R row i - 2)[m] of antenna p_g and
// lane j supplies f_pq (X row b')[m] of antenna q; summed over m = 0..3,
//     D_{g,b'}[i][j] = f_pq (Re V[i][b'] for i < 2, Im V[i - 2][b'] else),
// V = J_p J_q^H. The factor rides on the B operand (along a block's
// columns only q varies) and the accumulators sum over the sources inside
// the matrix core: the vector units evaluate f and scale eight values per
// pair, with no pass over the products afterwards.
//
// Strips: the pairs p <= q are covered by (p, quad Q) strips, p in
// [0, min(4Q + 4, na)); consecutive runs of SETS strips of one quad form
// a block's group. A host plan picks SETS and the number of consumer waves
// (4, 8 or 12; producers fill the rest of 1024 threads), at most four
// sets (eight accumulators each).

constexpr int kMxThreads = 1024;
constexpr int kMxMaxSets = 4;   // 1024 threads: 128 VGPRs; five sets spill

// antennas p of the strips of quad m
__host__ __device__ inline int mx_quad_span(int na, int m) { return min(4 * m + 4, na); }

static inline int mx_groups(int na, int sets) {
    int n = 0;
    for (int m = 0; 4 * m < na; ++m) n += (mx_quad_span(na, m) + sets - 1) / sets;
    return n;
}

// Per buffer: X and R rows of st sources x na antennas plus kMxMaxSets
// zeroed pad antennas (a block's last set and a partial last quad read
// up to kMxMaxSets - 1 antennas past the end), the same for the shape
// coordinates; then one chi^2 partial per wave.
__host__ __device__ inline size_t mx_buf_ants(int na, int st) { return size_t(st) * na + kMxMaxSets; }
static inline size_t mx_lds_bytes(int na, int st) {
    return gemm_x_offset(na) + 2 * mx_buf_ants(na, st) * (4 * sizeof(XRow<float>) + sizeof(float2)) +
           (kMxThreads / 64) * sizeof(double) +  // per-wave chi^2 partials
           2 * size_t(st) * sizeof(SrcRec<float>);   // source records of two tiles
}

struct MxPlan {
    int sets, waves, st;
    size_t lds;
};

// Fewest sets per SIMD (consumer waves per SIMD x sets) over the
// layouts whose groups fit 16 blocks per consumer wave; the source tile
// is the largest multiple of four up to 32 that fits the LDS budget.
static bool mx_plan(int na, size_t budget, MxPlan &p) {
    int best = 1 << 30;
    p.sets = p.waves = 0;
    for (int w : {8, 4, 12}) {
        for (int g = 1; g <= kMxMaxSets; ++g) {
            if (mx_groups(na, g) > 16 * w) continue;
            const int cost = ((w + 3) / 4) * g;
            if (cost < best) { best = cost; p.sets = g; p.waves = w; }
            break;
        }
    }
    if (!p.sets) return false;
    p.st = 32;
    while (p.st > 4 && mx_lds_bytes(na, p.st) > budget) p.st -= 4;
    p.lds = mx_lds_bytes(na, p.st);
    return p.lds <= budget && p.st <= kMxThreads - 64 * p.waves;
}

// Sources [s0, s1) of one kind into a block's SETS groups of pairs (see
// above). Xp / XYp address antenna p0 of source s0 (set g reads antenna
// p0 + g at a fixed offset), Xq / XYq antenna q; one source is na antenna
// records further on.
template <int SETS, bool GAUSS>
__device__ __forceinline__ void mx_sources(const XRow<float> *Xp, const XRow<float> *Xq, const float2 *XYp,
                                           const float2 *XYq, int na, int s0, int s1, f32x4 (&D)[SETS][2]) {
    for (int s = s0; s < s1; ++s) {
        const XRow<float> b0 = Xq[0], b1 = Xq[1];
        const float2 cq = XYq[0];
        // the source's scaled B operands first, each in its own register,
        // then the products: no vector write sits right before the
        // matrix instruction that reads it
        XRow<float> a[SETS];
        float fb[SETS][2][4];
#pragma unroll
        for (int g = 0; g < SETS; ++g) {
            a[g] = Xp[4 * g];
            const float2 cp = XYp[g];
            const float f = shape_f<float>(GAUSS, cq.x - cp.x, cq.y - cp.y);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                fb[g][0][m] = f * b0.v[m];
                fb[g][1][m] = f * b1.v[m];
            }
        }
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int g = 0; g < SETS; ++g) {
                D[g][0] = __builtin_amdgcn_mfma_f32_4x4x1f32(a[g].v[m], fb[g][0][m], D[g][0], 0, 0, 0);
                D[g][1] = __builtin_amdgcn_mfma_f32_4x4x1f32(a[g].v[m], fb[g][1][m], D[g][1], 0, 0, 0);
            }
        // schedule: every vector instruction of the source, then its
        // 8 SETS matrix instructions
        __builtin_amdgcn_sched_group_barrier(0x002, 64, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 8 * SETS, 0);
        Xp += 4 * na; Xq += 4 * na; XYp += na; XYq += na;
    }
}

// epi = 0: the sums go to G.dst (plus G.vin) as in the point pass; epi =
// 1: the epilogue is applied here (vis_out) and a chi^2 partial per
// workgroup written to A.partial.
// FROM_STAGE: the LDS set-up (antenna and channel records) is in place and
// the point sources' sums of every pair p <= q wait in `stage`
// (point_phase); the blocks start from them. bx: the workgroup's block
// index (its chi^2 partial's slot).
template <int SETS, bool FROM_STAGE>
__device__ __forceinline__ void shaped_phase(const RimeArgs<float> &A, const GemmArgs<float> &G, int ncw,
                                             int epi, unsigned char *smem_raw, int t, int c, uint32_t bx,
                                             const cx<float> *stage) {
    const mbx_dims &d = A.d;
    const int NA = d.na.local_size;
    const int tid = threadIdx.x;

    const int st = G.st;
    double *Luvw = reinterpret_cast<double *>(smem_raw);
    double *Lpa = Luvw + 3 * NA;
    double *Lant = Lpa + 2 * NA;
    double *Lch = Lant + 4 * NA;
    // two source-tile buffers: X and R rows per (source, antenna) as
    // [X row 0, X row 1, R row 0, R row 1], then shape coordinates (x, y)
    // per (source, antenna)
    XRow<float> *Xbase = reinterpret_cast<XRow<float> *>(smem_raw + gemm_x_offset(NA));
    const int bufa = int(mx_buf_ants(NA, st));
    const int xstride = 4 * bufa;            // rows per buffer
    float2 *XYbase = reinterpret_cast<float2 *>(Xbase + 2 * xstride);
    double *Lpart = reinterpret_cast<double *>(XYbase + 2 * bufa);
    SrcRec<float> *Pbase = reinterpret_cast<SrcRec<float> *>(Lpart + kMxThreads / 64);
    double term_sum = 0.0;

    const int np = dext(d.npsrc), ns = dext(d.nsrc), ng = dext(d.ngsrc);
    const int ntiles = (ns - np + st - 1) / st;
    auto tile_range = [&](int it, int &s0, int &nst) {
        s0 = np + it * st;
        nst = min(st, ns - s0);
    };
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    // source records of tile it into buffer it & 1, one producer thread per
    // source (no LDS reads: tiles 0 and 1 go before the set-up barrier)
    auto records = [&](int it) {
        const int ptid = tid - 64 * ncw;
        int s0, nst;
        tile_range(it, s0, nst);
        if (ptid >= 0 && ptid < nst)
            source_rec<float>(A, t, c, s0 + ptid, 1.2011224087864498, Pbase[(it & 1) * st + ptid]);
    };
    if (wave >= ncw) {
        if (ntiles > 0) records(0);
        if (ntiles > 1) records(1);
    }
    // pad antennas of both buffers: finite values for the reads past the end
    auto pad_clear = [&]() {
        for (int k = tid; k < 2 * kMxMaxSets; k += kMxThreads) {
            const int b = k / kMxMaxSets, a = st * NA + k % kMxMaxSets;
            for (int r = 0; r < 4; ++r) Xbase[b * xstride + 4 * a + r] = XRow<float>{{0.f, 0.f, 0.f, 0.f}};
            XYbase[b * bufa + a] = float2{0.f, 0.f};
        }
    };
    if constexpr (!FROM_STAGE) {
        tile_setup<float>(A, Luvw, Lpa, Lch, t, c, tid, kMxThreads);
        stage_antennas<float>(A, Lant, t, c, tid, kMxThreads);
        pad_clear();
        __syncthreads();
    }

    if (wave < ncw) {
        const int lane = tid & 63, blk = lane >> 2, i = lane & 3;
        // this block's group: quad Q, antennas p0 .. p0 + SETS - 1 (a
        // block past the last group computes quad 0 and stores nothing)
        int gi = wave * 16 + blk, Q = -1, p0 = 0;
        for (int m = 0; 4 * m < NA; ++m) {
            const int n = (mx_quad_span(NA, m) + SETS - 1) / SETS;
            if (gi < n) { Q = m; p0 = gi * SETS; break; }
            gi -= n;
        }
        const bool active = Q >= 0;
        if (!active) Q = 0;
        const int q = 4 * Q + i;                 // < na + 3: inside the pad
        const bool qok = active && q < NA;
        const int span = active ? mx_quad_span(NA, Q) : 0;
        // groups end inside the quad's span, so p0 + SETS - 1 < na + SETS - 1
        f32x4 D[SETS][2];
#pragma unroll
        for (int g = 0; g < SETS; ++g) D[g][0] = D[g][1] = f32x4{0.f, 0.f, 0.f, 0.f};
        if constexpr (FROM_STAGE) {
            // the point sums of this lane's pairs (p_g, q), p_g <= q
#pragma unroll
            for (int g = 0; g < SETS; ++g) {
                const int p = p0 + g;
                if (!qok || p >= span || p > q) continue;
                const cx<float> *v = stage + stage_pair(p, q, NA) * 4;
#pragma unroll
                for (int ip = 0; ip < 2; ++ip)
#pragma unroll
                    for (int kq = 0; kq < 2; ++kq) {
                        D[g][kq][ip] = v[2 * ip + kq].re;
                        D[g][kq][2 + ip] = v[2 * ip + kq].im;
                    }
            }
            __syncthreads();                  // staged sums read: the buffers are free
            pad_clear();
        }
        __syncthreads();                      // tile 0 produced
        for (int it = 0; it < ntiles; ++it) {
            const int buf = it & 1;
            int s0t, nst;
            tile_range(it, s0t, nst);
            const XRow<float> *Xb = Xbase + buf * xstride;
            const float2 *XY = XYbase + buf * bufa;
            // gaussian sources precede the sersic ones: one loop per kind
            const int ngt = min(max(np + ng - s0t, 0), nst);
            mx_sources<SETS, true>(Xb + 4 * p0 + i, Xb + 4 * q, XY + p0, XY + q, NA, 0, ngt, D);
            mx_sources<SETS, false>(Xb + 4 * (ngt * NA + p0) + i, Xb + 4 * (ngt * NA + q),
                                    XY + ngt * NA + p0, XY + ngt * NA + q, NA, ngt, nst, D);
            __syncthreads();
        }
        // the (p <= q) pairs to their baselines (reversed baselines get the
        // conjugate transpose): running sums, or with epi the epilogue
        const uint64_t NC = d.nchan.local_size, NBL = d.nbl.local_size;
        const bool final_batch = d.nsrc.upper_extent >= d.nsrc.global_size;
#pragma unroll
        for (int g = 0; g < SETS; ++g) {
            const int p = p0 + g;
            if (!qok || p >= span || p > q) continue;
            const int e = G.bmap[p * NA + q];
            if (e < 0) continue;
            const bool rev = (e & MBX_BL_REVERSED) != 0;
            const int b = e & ~(MBX_BL_REVERSED | MBX_BL_AUTO);
            if (uint64_t(b) >= NBL) continue;   // a map for another layout
            J4<float> V;
#pragma unroll
            for (int ip = 0; ip < 2; ++ip)
#pragma unroll
                for (int kq = 0; kq < 2; ++kq) {
                    const float re = D[g][kq][ip], im = D[g][kq][2 + ip];
                    V.v[rev ? 2 * kq + ip : 2 * ip + kq] = cx<float>{re, rev ? -im : im};
                }
            const uint64_t vi = ((uint64_t(t) * NBL + b) * NC + c) * 4;
            if (G.vin) {
                const J4<float> o = jload(G.vin + vi);
#pragma unroll
                for (int k = 0; k < 4; ++k) { V.v[k].re += o.v[k].re; V.v[k].im += o.v[k].im; }
            }
            if (epi)   // the baseline's own antenna order: (q, p) when reversed
                term_sum += vis_out(A, t, b, c, rev ? q : p, rev ? p : q, V, final_batch);
            else
                jstore(G.dst + vi, V);
        }
    } else {
        const int ptid = tid - 64 * ncw, pth = kMxThreads - 64 * ncw;
        // tile it into buffer it & 1 from its source records: one Jones per
        // (source, antenna) item, items taken in order by the producer
        // threads
        auto produce = [&](int it) {
            const int buf = it & 1;
            int s0, nst;
            tile_range(it, s0, nst);
            XRow<float> *Xb = Xbase + buf * xstride;
            float2 *XY = XYbase + buf * bufa;
            const SrcRec<float> *P = Pbase + buf * st;
            // item i = si NA + a; the stride pth advances (si, a) by (dsi, da)
            const int dsi = pth / NA, da = pth - dsi * NA;
            int si = ptid / NA, a = ptid - si * NA;
            for (; si < nst; si += dsi, a += da) {
                if (a >= NA) { a -= NA; ++si; }
                if (si >= nst) break;
                const SrcRec<float> &sr = P[si];
                const J4<float> J = staged_jones<float, kGemmTaps<float, true>()>(A, Luvw, Lpa, Lant, Lch, c, a,
                                                                            sr.l0, sr.m0, sr.nm1, sr.B);
                XRow<float> *r = Xb + 4 * (si * NA + a);
                r[0] = XRow<float>{{J.v[0].re, J.v[1].re, J.v[0].im, J.v[1].im}};
                r[1] = XRow<float>{{J.v[2].re, J.v[3].re, J.v[2].im, J.v[3].im}};
                r[2] = XRow<float>{{J.v[0].im, J.v[1].im, -J.v[0].re, -J.v[1].re}};
                r[3] = XRow<float>{{J.v[2].im, J.v[3].im, -J.v[2].re, -J.v[3].re}};
                // shape coordinates: linear in the antenna's (u, v)
                const double u = Luvw[3 * a], v = Luvw[3 * a + 1];
                XY[si * NA + a] = float2{float(sr.cu * u + sr.cv * v), float(sr.du * u + sr.dv * v)};
            }
        };
        // records of tile it + 2 go to buffer it & 1 while tile it + 1 is
        // produced (tile it's records were last read before the barrier
        // that ended the previous step)
        if constexpr (FROM_STAGE) {
            __syncthreads();                  // staged sums read: the buffers are free
            pad_clear();
        }
        if (ntiles > 0) produce(0);
        __syncthreads();
        for (int it = 0; it < ntiles; ++it) {
            if (it + 1 < ntiles) produce(it + 1);
            if (it + 2 < ntiles) records(it + 2);
            __syncthreads();
        }
    }
    if (epi) wg_partial(term_sum, Lpart, A.partial + bx, tid, kMxThreads);
}

template <int SETS, bool BATCHED = false>
__global__ void __launch_bounds__(kMxThreads)
k_shaped_mx(RimeArgs<float> A, GemmArgs<float> G, int ncw, int epi) {
    extern __shared__ __align__(32) unsigned char smem_raw[];
    // in a launch over several samples: this workgroup's sample and slot
    const SampleAt at = enter_sample_if<BATCHED>(A);
    if constexpr (BATCHED) {
        G.dst += uint64_t(at.j) * A.span.vstride;
        if (G.vin) G.vin += uint64_t(at.j) * A.span.vstride;
    }
    const int nt_ext = dext(A.d.ntime), nc_ext = dext(A.d.nchan);
    int t, c;
    if (!tc_coords(tc_per_xcd(nt_ext, nc_ext), nt_ext, nc_ext, t, c, at.bx)) {
        if (epi && threadIdx.x == 0) A.partial[at.bx] = 0.0;   // padding workgroup
        return;
    }
    shaped_phase<SETS, false>(A, G, ncw, epi, smem_raw, t, c, at.bx, nullptr);
}

// The whole float forward of one (t, c) in one workgroup: the point
// sources with the point pass's wave roles (Gp), their sums of every pair
// p <= q staged in LDS over the dead X buffers, then the gaussian and
// sersic sources on 4x4 blocks starting from those sums (Gs, ncw consumer
// waves), and the epilogue. No running sums pass through device memory.
static_assert(gemm_threads<float, false>() == kMxThreads, "both phases use the same workgroup");
template <int SETS, bool BATCHED = false>
__global__ void __launch_bounds__(kMxThreads)
k_forward_mx(RimeArgs<float> A, GemmArgs<float> Gp, GemmArgs<float> Gs, int ncw) {
    extern __shared__ __align__(32) unsigned char smem_raw[];
    // in a launch over several samples: this workgroup's sample and slot
    const SampleAt at = enter_sample_if<BATCHED>(A);
    if constexpr (BATCHED) {
        if (Gs.vin) Gs.vin += uint64_t(at.j) * A.span.vstride;
    }
    const int nt_ext = dext(A.d.ntime), nc_ext = dext(A.d.nchan);
    int t, c;
    if (!tc_coords(tc_per_xcd(nt_ext, nc_ext), nt_ext, nc_ext, t, c, at.bx)) {
        if (threadIdx.x == 0) A.partial[at.bx] = 0.0;   // padding workgroup
        return;
    }
    cx<float> *stage = reinterpret_cast<cx<float> *>(smem_raw + gemm_x_offset(A.d.na.local_size));
    point_phase<float, false, true>(A, Gp, smem_raw, t, c, stage);
    __syncthreads();                           // staged sums complete
    shaped_phase<SETS, true>(A, Gs, ncw, 1, smem_raw, t, c, at.bx, stage);
}

static int launch_shaped_mx(const RimeArgs<float> &A, GemmArgs<float> G, uint64_t nwg, int epi,
                            hipStream_t stream) {
    const int na = A.d.na.local_size;
    static size_t budget = gemm_lds_budget();
    MxPlan p;
    if (!mx_plan(na, budget, p)) return set_error(MBX_E_UNSUPPORTED, "shaped matrix pass: %d antennas", na);
    G.st = p.st;
    const dim3 grid{unsigned(nwg)}, block{unsigned(kMxThreads)};
    switch (p.sets) {
    case 1: hipLaunchKernelGGL((k_shaped_mx<1>), grid, block, p.lds, stream, A, G, p.waves, epi); break;
    case 2: hipLaunchKernelGGL((k_shaped_mx<2>), grid, block, p.lds, stream, A, G, p.waves, epi); break;
    case 3: hipLaunchKernelGGL((k_shaped_mx<3>), grid, block, p.lds, stream, A, G, p.waves, epi); break;
    default: hipLaunchKernelGGL((k_shaped_mx<4>), grid, block, p.lds, stream, A, G, p.waves, epi); break;
    }
    return hip_status(hipGetLastError(), "shaped matrix pass launch");
}

// The fused float forward (k_forward_mx): the point phase's source tile,
// the shaped phase's plan and the LDS of both; false where the staged sums
// would reach the shaped phase's source records (written while the sums
// are read) or the LDS does not hold both phases.
struct FwdMxPlan {
    MxPlan mx;
    int stp;
    size_t lds;
};

static bool forward_mx_plan(int na, size_t budget, FwdMxPlan &f) {
    if (!mx_plan(na, budget, f.mx)) return false;
    f.stp = gemm_st<float, false>(na, budget);
    const size_t xy_bytes = 2 * mx_buf_ants(na, f.mx.st) * (4 * sizeof(XRow<float>) + sizeof(float2));
    if (stage_bytes<float>(na) > xy_bytes) return false;
    f.lds = std::max(gemm_lds_bytes<float, false>(na, f.stp), f.mx.lds);
    return f.lds <= budget;
}

// k_forward_mx over `grid` workgroups; vin: the running sums of earlier
// source batches, or NULL.
template <bool BATCHED>
static int launch_forward_mx(const RimeArgs<float> &A, const int32_t *bmap, const cx<float> *vin,
                             const FwdMxPlan &f, uint64_t grid, hipStream_t stream) {
    const int na = A.d.na.local_size;
    GemmArgs<float> Gp;
    Gp.bmap = bmap; Gp.dst = nullptr; Gp.vin = nullptr;
    Gp.rows = gemm_rows(na);
    const int nR = Gp.rows / 16;
    Gp.ntp = nR * (nR + 1) / 2;
    Gp.st = f.stp;
    constexpr int cw = gemm_consumers<float, false>() / 64;
    Gp.tpw = (Gp.ntp + cw - 1) / cw;
    if (Gp.tpw > gemm_max_tp<float, false>()) return set_error(MBX_E_UNSUPPORTED, "fused forward: %d antennas", na);
    GemmArgs<float> Gs = Gp;
    Gs.vin = vin;
    Gs.st = f.mx.st;
    RimeArgs<float> Ag = A;
    Ag.ct = 1;
    const dim3 g{unsigned(grid)}, block{unsigned(kMxThreads)};
    const int w = f.mx.waves;
    switch (f.mx.sets) {
    case 1: hipLaunchKernelGGL((k_forward_mx<1, BATCHED>), g, block, f.lds, stream, Ag, Gp, Gs, w); break;
    case 2: hipLaunchKernelGGL((k_forward_mx<2, BATCHED>), g, block, f.lds, stream, Ag, Gp, Gs, w); break;
    case 3: hipLaunchKernelGGL((k_forward_mx<3, BATCHED>), g, block, f.lds, stream, Ag, Gp, Gs, w); break;
    default: hipLaunchKernelGGL((k_forward_mx<4, BATCHED>), g, block, f.lds, stream, Ag, Gp, Gs, w); break;
    }
    return hip_status(hipGetLastError(), "fused forward launch");
}

template <typename T>
bool point_gemm_supported(const mbx_dims &d) {
    const int na = d.na.local_size;
    return dext(d.nsrc) > 0 && na > 0 && gemm_rows(na) / 16 <= kGemmMaxRowBlocks;
}

template <typename T>
bool shaped_gemm_pays(const mbx_dims &d) {
    // Measured on MI355X at 64 antennas with 600 gaussian + sersic
    // sources: float, the 4x4-block pass (k_shaped_mx) 11.6 ms against
    // 14.9 ms for the vector-unit kernel; float64, the 16x16 pass with
    // per-element shape factors ahead of the vector-unit kernel. They are
    // used from 40 (float) and 48 (float64) antennas up: fewer antennas
    // give fewer pairs per source to fill the matrix cores.
    if (sizeof(T) == 4) {
        MxPlan p;
        return d.na.local_size >= 40 && mx_plan(d.na.local_size, gemm_lds_budget(), p);
    }
    return d.na.local_size >= 48;
}

uint64_t gemm_workgroups(const mbx_dims &d) {
    return 8 * uint64_t(tc_per_xcd(dext(d.ntime), dext(d.nchan)));
}

template <typename T>
int launch_point_gemm(const RimeArgs<T> &A, const int32_t *bmap, cx<T> *dst, const cx<T> *vin,
                      bool shaped, hipStream_t stream, uint64_t *epilogue_wgs) {
    const mbx_dims &d = A.d;
    if (epilogue_wgs) *epilogue_wgs = 0;
    const int na = d.na.local_size;
    GemmArgs<T> G;
    G.bmap = bmap; G.dst = dst; G.vin = vin;
    G.rows = gemm_rows(na);
    const int nR = G.rows / 16;
    G.ntp = nR * (nR + 1) / 2;
    static size_t budget = gemm_lds_budget();
    const uint64_t nwg = gemm_workgroups(d);
    RimeArgs<T> Ag = A;
    Ag.ct = 1;
    const int np = dext(d.npsrc), ns = dext(d.nsrc);
    if constexpr (sizeof(T) == 4) {
        // float with the epilogue: both passes in one kernel where it fits
        FwdMxPlan f;
        if (shaped && epilogue_wgs && np > 0 && forward_mx_plan(na, budget, f)) {
            const int r = launch_forward_mx<false>(Ag, bmap, vin, f, nwg, stream);
            if (!r) *epilogue_wgs = nwg;
            return r;
        }
    }
    if (np > 0) {
        G.st = gemm_st<T, false>(na, budget);
        const size_t lds = gemm_lds_bytes<T, false>(na, G.st);
        constexpr int cw = gemm_consumers<T, false>() / 64;
        G.tpw = (G.ntp + cw - 1) / cw;
        if (G.tpw > gemm_max_tp<T, false>()) return set_error(MBX_E_UNSUPPORTED, "point gemm: %d antennas", na);
        hipLaunchKernelGGL((k_point_gemm<T, false>), dim3(unsigned(nwg)), dim3(gemm_threads<T, false>()),
                           lds, stream, Ag, G);
        int r = hip_status(hipGetLastError(), "point gemm launch");
        if (r) return r;
        G.vin = dst;   // the shaped pass continues from the point sums
    }
    if (shaped && ns > np) {
        if constexpr (sizeof(T) == 4) {
            const int r = launch_shaped_mx(Ag, G, nwg, epilogue_wgs ? 1 : 0, stream);
            if (!r && epilogue_wgs) *epilogue_wgs = nwg;
            return r;
        }
        constexpr int cw = gemm_consumers<T, true>() / 64;
        G.tpw = (G.ntp + cw - 1) / cw;
        if (G.tpw > gemm_max_tp<T, true>()) return set_error(MBX_E_UNSUPPORTED, "shaped gemm: %d antennas", na);
        G.st = gemm_st<T, true>(na, budget);
        const size_t lds = gemm_lds_bytes<T, true>(na, G.st);
        hipLaunchKernelGGL((k_point_gemm<T, true>), dim3(unsigned(nwg)), dim3(gemm_threads<T, true>()),
                           lds, stream, Ag, G);
        return hip_status(hipGetLastError(), "shaped gemm launch");
    }
    return MBX_OK;
}

bool batched_mx_supported(const mbx_dims &d) {
    MxPlan p;
    return point_gemm_supported<float>(d) && mx_plan(d.na.local_size, gemm_lds_budget(), p);
}

int launch_batched_mx(const RimeArgs<float> &A, const int32_t *bmap, cx<float> *buf, bool any_point,
                      unsigned nrep, hipStream_t stream) {
    const mbx_dims &d = A.d;
    const int na = d.na.local_size;
    if (!batched_mx_supported(d)) return set_error(MBX_E_UNSUPPORTED, "batched matrix passes: %d antennas", na);
    const uint64_t nwg = gemm_workgroups(d);
    if (A.span.wgs != nwg) return set_error(MBX_E_DIMS, "batched matrix passes: span of %u workgroups, "
                                            "the passes need %llu", A.span.wgs, (unsigned long long)nwg);
    const dim3 grid{unsigned(nwg * nrep)};
    GemmArgs<float> G;
    G.bmap = bmap; G.dst = buf; G.vin = nullptr;
    G.rows = gemm_rows(na);
    const int nR = G.rows / 16;
    G.ntp = nR * (nR + 1) / 2;
    static size_t budget = gemm_lds_budget();
    RimeArgs<float> Ag = A;
    Ag.ct = 1;
    FwdMxPlan f;
    if (any_point && forward_mx_plan(na, budget, f))   // one launch: both passes in one kernel
        return launch_forward_mx<true>(Ag, bmap, nullptr, f, nwg * nrep, stream);
    if (any_point) {
        G.st = gemm_st<float, false>(na, budget);
        const size_t lds = gemm_lds_bytes<float, false>(na, G.st);
        constexpr int cw = gemm_consumers<float, false>() / 64;
        G.tpw = (G.ntp + cw - 1) / cw;
        if (G.tpw > gemm_max_tp<float, false>()) return set_error(MBX_E_UNSUPPORTED, "point gemm: %d antennas", na);
        hipLaunchKernelGGL((k_point_gemm<float, false, true>), grid, dim3(gemm_threads<float, false>()), lds,
                           stream, Ag, G);
        const int r = hip_status(hipGetLastError(), "batched point gemm launch");
        if (r) return r;
        G.vin = buf;
    }
    MxPlan p;
    mx_plan(na, budget, p);
    G.st = p.st;
    const dim3 block{unsigned(kMxThreads)};
    switch (p.sets) {
    case 1: hipLaunchKernelGGL((k_shaped_mx<1, true>), grid, block, p.lds, stream, Ag, G, p.waves, 1); break;
    case 2: hipLaunchKernelGGL((k_shaped_mx<2, true>), grid, block, p.lds, stream, Ag, G, p.waves, 1); break;
    case 3: hipLaunchKernelGGL((k_shaped_mx<3, true>), grid, block, p.lds, stream, Ag, G, p.waves, 1); break;
    default: hipLaunchKernelGGL((k_shaped_mx<4, true>), grid, block, p.lds, stream, Ag, G, p.waves, 1); break;
    }
    return hip_status(hipGetLastError(), "batched shaped pass launch");
}

template bool point_gemm_supported<float>(const mbx_dims &);
template bool point_gemm_supported<double>(const mbx_dims &);
template bool shaped_gemm_pays<float>(const mbx_dims &);
template bool shaped_gemm_pays<double>(const mbx_dims &);
template int launch_point_gemm<float>(const RimeArgs<float> &, const int32_t *, cx<float> *,
                                      const cx<float> *, bool, hipStream_t, uint64_t *);
template int launch_point_gemm<double>(const RimeArgs<double> &, const int32_t *, cx<double> *,
                                       const cx<double> *, bool, hipStream_t, uint64_t *);

}  // namespace mbx
