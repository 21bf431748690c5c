// Source stage of the adjoint, one workgroup per (time t, channel c). No
// reference counterpart.
//
// With Vb the visibility adjoint (from k_adj_vis), the Jones adjoint of
// antenna p for source s is
//     Jb_ps = sum_q f_pq,s W_pq J_qs,
// where W is the 2na x 2na Hermitian completion of Vb: W_pq = Vb_b and
// W_qp = Vb_b^H for baseline b = (p, q), W_pp = Vb_b + Vb_b^H for an
// auto-correlation.
//
// Point sources (f = 1): Jb = W X is a matrix product, X[2q+k][2s+j] =
// J_qs[k][j]; in real arithmetic Re Jb = [Re W | -Im W] [Re X; Im X] and
// Im Jb = [Im W | Re W] [Re X; Im X], on v_mfma_*_16x16x4 with 16x16
// output tiles (antenna rows x (source, column)). k_adj_point pipelines
// it over source tiles with W in the matrix waves' registers;
// k_adj_gemm<T, false> (W in LDS, one phase per tile) is the fallback.
// Shaped sources (gaussian, sersic), k_adj_gemm<T, true, WIDE>: one thread
// per (source, antenna p) walks q on the vector units, with W and J from
// LDS, and also forms the shape-factor adjoint d chi^2/d f_pq =
// Re <J_p, W_pq J_q> (summed over both orders of a pair, this is
// 2 Re <J_p, Vb J_q> per baseline).
// Every pass then runs the per-(source, antenna) backward of J = E K
// sqrt(B) and reduces each source's gradient in LDS before one set of
// float64 atomics per source and window of source tiles. BATCHED
// instances cover several samples' grids in one launch (SampleSpan,
// mbx_tile.hpp; enter_sample for AdjArgs, mbx_backward.hpp).

#include <type_traits>

#include "mbx_backward.hpp"
#include "mbx_fmath.hpp"
#include "mbx_gemm.hpp"

namespace mbx {

constexpr int kAdjThreads = 512;
// Every wave takes one 16x16 output tile of the product (nR * st / 8 <= 8
// tiles): two waves per SIMD hide each other's dependent-MFMA latency.
constexpr int kAdjMfmaWaves = 8;
constexpr int kAdjMaxTiles = 1;

struct AdjGemmShape {
    int rows;    // 2 na rounded up to 16
    int st;      // sources per tile (multiple of 8, st * na <= threads)
    int sw;      // sources per gradient flush window (multiple of st)
};

static __host__ __device__ inline size_t adjg_setup_bytes(int na) {
    return ((sizeof(double) * (9 * size_t(na) + CH)) + 31) & ~size_t(31);
}

// Shaped pass records (elements). W is kept as its upper triangle: block
// (a, b), a <= b, at adjs_wrow(a) + 8 (b - a), rows padded by four values
// so that the rows of the antennas one wave reads at once start in
// different LDS banks; W_pq for q < p is the conjugate transpose of block
// (q, p). X record (s, a) at s XPS + 10 a, XPS / 2 odd for the same reason.
__host__ __device__ inline int adjs_wrow(int a, int na) { return a * (8 * na + 4) - 4 * a * (a - 1); }
__host__ __device__ inline int adjs_welems(int na) { return 4 * na * (na + 2); }
__host__ __device__ inline int adjs_xps(int na) { return 10 * na + ((na & 1) ? 4 : 2); }

template <typename T>
static inline size_t adjg_lds_bytes(int na, int rows, int st, int sw, bool shaped) {
    const size_t w = shaped ? size_t(adjs_welems(na)) : 2 * size_t(rows) * rows;      // W
    const size_t x = shaped ? size_t(st) * adjs_xps(na) : 2 * size_t(2 * st) * rows; // X (then Jb)
    return adjg_setup_bytes(na) + (w + x) * sizeof(T)
           + size_t(sw) * 10 * sizeof(double)                           // gradient sums
           + 2 * size_t(st) * 4 * sizeof(double)                        // l, m, n - 1, 1 / n (2 tiles)
           + 2 * size_t(st) * 24 * sizeof(T)                            // shape, I Q U V P, sqrt(B)
           + (shaped ? 2 * size_t(na) : 0) * sizeof(T)                  // u, v
           + (shaped && sizeof(T) == 8 ? kExpTab * sizeof(double) : 0); // exp_neg table
}

// Gaussian shape coordinates are kept scaled by sqrt(log2 e) in float, so
// that a pair's exp(-r2) is one exp2 of the scaled r2 (gauss_of_scaled).
template <typename T> constexpr double adj_gauss_pre() { return sizeof(T) == 4 ? 1.2011224087864498 : 1.0; }

// A Jones matrix from eight consecutive values (re, im interleaved) at an
// address aligned to two values.
template <typename T>
__device__ __forceinline__ J4<T> jload_pairs(const T *p) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 *q = reinterpret_cast<const v2 *>(p);
    J4<T> r;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const v2 a = q[i];
        r.v[i] = {a.x, a.y};
    }
    return r;
}

// Sersic factor of a pair from r2 = dx^2 + dy^2: (1 + r2)^(-3/2), with
// id = 1 / (1 + r2) (the gaussian one is gauss_of, mbx_tile.hpp).
static __device__ __forceinline__ float sersic_factor(float r2, float &id) {
    const float rs = __builtin_amdgcn_rsqf(1.0f + r2);
    id = rs * rs;
    return id * rs;
}
static __device__ __forceinline__ double sersic_factor(double r2, double &id) {
    const double rs = rsq_refined(1.0 + r2);
    id = rs * rs;
    return id * rs;
}

// WIDE (shaped pass only): the items' forward state is not held across the
// pair loop but recomputed for the backward, which leaves registers for
// more waves: 1024 threads, one (source, antenna p) item each, and the
// item's thread runs its backward right after its pair loop. Float only
// (in float64 the recomputed backward needs more registers than sixteen,
// or twelve, waves leave; measured slower than keeping the state).
template <typename T, bool WIDE> constexpr int adjg_threads() { return WIDE ? 1024 : kAdjThreads; }

template <typename T, bool SHAPED, bool WIDE = false, bool BATCHED = false>
__global__ void __launch_bounds__((adjg_threads<T, WIDE>()))
k_adj_gemm(AdjArgs<T> A, AdjGemmShape S) {
    static_assert(SHAPED || !WIDE, "the wide layout is for the shaped pass");
    static_assert(sizeof(T) == 4 || !WIDE, "the wide layout is for float");
    constexpr int kThreads = adjg_threads<T, WIDE>();
    extern __shared__ __align__(32) unsigned char smem_raw[];
    // in a launch over several samples: this workgroup's sample and slot
    const SampleAt at = enter_sample_if<BATCHED>(A);
    const RimeArgs<T> &R = A.R;
    const mbx_dims &d = R.d;
    const int NA = d.na.local_size;
    const uint64_t NT = d.ntime.local_size, NC = d.nchan.local_size, NBL = d.nbl.local_size;
    const int nt_ext = dext(d.ntime), nc_ext = dext(d.nchan), nb_ext = dext(d.nbl);
    const int tid = threadIdx.x;
    // XCD order of kChanGroup (mbx_tile.hpp)
    int t, c;
    if (!tc_coords(tc_per_xcd(nt_ext, nc_ext), nt_ext, nc_ext, t, c, at.bx)) return;
    const int rows = S.rows, st = S.st;

    double *Luvw = reinterpret_cast<double *>(smem_raw);
    double *Lpa = Luvw + 3 * NA;
    double *Lant = Lpa + 2 * NA;
    double *Lch = Lant + 4 * NA;
    // Point pass: W and X as separate re and im planes; X is [n][row] (MFMA
    // operand reads along the rows), n = 2 s + j, row = 2 a + k.
    // Shaped pass: complex interleaved records, walked by pointers over q:
    //   W block (a, b), a <= b, at Wr + adjs_wrow(a) + 8 (b - a): W_ab[i][k]
    //   at 2 (2 i + k) (re, im);
    //   X record (s, a) at Xr + s XPS + 10 a: J_as[k][j] at 2 (2 k + j)
    //   (re, im), then the shape coordinates x, y of (antenna a, source s).
    const int XPS = adjs_xps(NA);
    const int wsize = SHAPED ? adjs_welems(NA) : 2 * rows * rows;
    const int xsize = SHAPED ? st * XPS : 4 * st * rows;
    T *Wr = reinterpret_cast<T *>(smem_raw + adjg_setup_bytes(NA));
    T *Wi = Wr + rows * rows;
    T *Xr = Wr + wsize;
    T *Xi = Xr + 2 * st * rows;
    // Point pass operands are blocked by four along the summed index, so a
    // lane's four consecutive k values are one 16/32-byte read and
    // consecutive lanes (rows of W, sources of X) read consecutive blocks:
    //   W[row][col] at ((col / 4) rows + row) 4 + col % 4,
    //   X[n][row]   at ((row / 4) 2 st + n) 4 + row % 4   (point pass)
    auto xi_ = [&](int n, int row) { return ((row >> 2) * (2 * st) + n) * 4 + (row & 3); };
    auto wi_ = [&](int row, int col) { return ((col >> 2) * rows + row) * 4 + (col & 3); };
    J4<T> *Jb = reinterpret_cast<J4<T> *>(Xr); // [st][NA], written over X
    double *sacc = reinterpret_cast<double *>(Xr + xsize);          // [st][10]
    // per-source records of two tiles (the next one is staged during the
    // current one's Jones production)
    double *Lsd0 = sacc + S.sw * 10;                                // [2][st][4]: l, m, n - 1, 1 / n
    T *Lsh0 = reinterpret_cast<T *>(Lsd0 + 2 * st * 4);             // [2][st][8]: shape record
    T *Lst0 = Lsh0 + 2 * st * 8;                                    // [2][st][16]: IQUV, src_bright, sqrt(B)
    T *Luv = Lst0 + 2 * st * 16;                                    // shaped: (u, v) per antenna
    double *Ltab = reinterpret_cast<double *>(Luv + 2 * NA);        // shaped float64: exp table

    tile_setup<T>(R, Luvw, Lpa, Lch, t, c, tid, kThreads);
    stage_antennas<T>(R, Lant, t, c, tid, kThreads);
    if constexpr (SHAPED && sizeof(T) == 8) exp_table_fill(Ltab, tid, kThreads);
    for (int i = tid; i < wsize; i += kThreads) Wr[i] = T(0);
    for (int i = tid; i < S.sw * 10; i += kThreads) sacc[i] = 0.0;
    __syncthreads();
    if (SHAPED)
        for (int a = tid; a < NA; a += kThreads) { Luv[2 * a] = T(Luvw[3 * a]); Luv[2 * a + 1] = T(Luvw[3 * a + 1]); }
    // W from this (t, c)'s visibility adjoints
    // (re, im) of W_pq[i][k]: point pass planes, shaped pass records
    auto wput = [&](int p, int i, int q, int k, T re, T im) {
        if (SHAPED) {
            if (p > q) return;                 // the upper triangle only
            T *w = Wr + adjs_wrow(p, NA) + 8 * (q - p) + 2 * (2 * i + k);
            w[0] = re;
            w[1] = im;
        } else {
            Wr[wi_(2 * p + i, 2 * q + k)] = re;
            Wi[wi_(2 * p + i, 2 * q + k)] = im;
        }
    };
    for (int b = tid; b < nb_ext; b += kThreads) {
        const int p = min(max(R.a1[uint64_t(t) * NBL + b], 0), NA - 1);
        const int q = min(max(R.a2[uint64_t(t) * NBL + b], 0), NA - 1);
        const J4<T> V = jload(A.vbar + ((uint64_t(t) * NBL + b) * NC + c) * 4);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const cx<T> v = V.v[2 * i + k];
                if (p == q) {
                    const cx<T> u = V.v[2 * k + i];   // + (Vb^H)[i][k] = conj(Vb[k][i])
                    wput(p, i, p, k, v.re + u.re, v.im - u.im);
                } else {
                    wput(p, i, q, k, v.re, v.im);
                    wput(q, k, p, i, v.re, -v.im);
                }
            }
    }
    __syncthreads();

    const int np = dext(d.npsrc), ns_all = dext(d.nsrc), ng = dext(d.ngsrc);
    const uint64_t NGS = d.ngsrc.local_size, NSS = d.nssrc.local_size;
    const int sbeg = SHAPED ? np : 0, send = SHAPED ? ns_all : np;
    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint64_t NTT = d.ntime.local_size;
    // Each source's gradient is summed in LDS over a window of sw sources
    // and added to global memory once per window: the float64 atomics'
    // latency is paid once per window, not once per tile.
    // source records of the tile at s0t (ns sources) into buffer buf, one
    // thread per source from thread first on
    auto stage_params = [&](int s0t, int ns, int buf, int first) {
        const int si = tid - first;
        if (si < 0 || si >= ns) return;
        double *Lsd = Lsd0 + buf * st * 4;
        T *Lsh = Lsh0 + buf * st * 8, *Lst = Lst0 + buf * st * 16;
        double l0, m0, nm1;
        const J4<T> B = source_sqrtb<T>(R, Lch, t, s0t + si, l0, m0, nm1);
        Lsd[4 * si] = l0; Lsd[4 * si + 1] = m0; Lsd[4 * si + 2] = nm1; Lsd[4 * si + 3] = 1.0 / (nm1 + 1.0);
        const T *St = R.stokes + (uint64_t(s0t + si) * NTT + t) * 4;
        T *o = Lst + 16 * si;
        o[0] = St[0]; o[1] = St[1]; o[2] = St[2]; o[3] = St[3];
        const SrcBright<T> sbr =
            src_bright<T>(St[0], St[1], St[2], St[3], exp_t<T>(double(R.alpha[uint64_t(s0t + si) * NTT + t]) * Lch[4]));
        o[4] = sbr.rP; o[5] = sbr.s0; o[6] = sbr.it0; o[7] = sbr.is0;
        reinterpret_cast<J4<T> *>(o + 8)[0] = B;
        if (SHAPED) {
            // shape record: gaussian 1, k, k R, el, em, R (k = gauss_scale f);
            // sersic 2, k R', 2 e1 / den, 2 e2 / den, k / den, e1, e2
            // (k = 2 pi f / C, den = 1 - e1^2 - e2^2, R' = R / den)
            const int s = s0t + si;
            T *h = Lsh + 8 * si;
            if (s < np + ng) {
                const int g = s - np;
                const double el = R.gauss[g], em = R.gauss[g + NGS], Rg = R.gauss[g + 2 * NGS];
                const double k = R.gauss_scale * Lch[3];
                h[0] = T(1); h[1] = T(k); h[2] = T(k * Rg); h[3] = T(el); h[4] = T(em); h[5] = T(Rg);
            } else {
                const int g = s - np - ng;
                const double e1 = R.sersic[g], e2 = R.sersic[g + NSS], Rs = R.sersic[g + 2 * NSS];
                const double den = 1.0 - e1 * e1 - e2 * e2, k = R.two_pi_over_c * Lch[3];
                h[0] = T(2); h[1] = T(k * Rs / den); h[2] = T(2.0 * e1 / den); h[3] = T(2.0 * e2 / den);
                h[4] = T(k / den); h[5] = T(e1); h[6] = T(e2);
            }
        }
    };
    if (sbeg < send) stage_params(sbeg, min(st, send - sbeg), 0, 0);
    __syncthreads();
    int wbase = sbeg;
    for (int s0 = sbeg, it = 0; s0 < send; s0 += st, ++it) {
        const int nst = min(st, send - s0);
        double *sac = sacc + (s0 - wbase) * 10;    // this tile's sums
        const double *Lsd = Lsd0 + (it & 1) * st * 4;
        const T *Lsh = Lsh0 + (it & 1) * st * 8, *Lst = Lst0 + (it & 1) * st * 16;
        // the next tile's source records, by the last st threads (idle in
        // the Jones production below when st na < threads)
        if (s0 + st < send) stage_params(s0 + st, min(st, send - s0 - st), (it + 1) & 1, kThreads - st);
        // ---- X of the tile (columns of sources past nst are zero). Each
        // thread owns one (source, antenna) item (st * na <= threads) and
        // keeps its forward state for the backward after the product.
        // Point pass: consecutive threads take consecutive antennas
        // (conflict-free [n][row] writes); shaped: consecutive sources.
        const bool own = tid < nst * NA;
        const int isi = own ? (SHAPED ? tid % nst : tid / NA) : 0;
        const int ia = own ? (SHAPED ? tid / nst : tid % NA) : 0;
        ItemState<T> ist;
        if (own) {
            const double *an = Lant + 4 * ia;
            const T *o = Lst + 16 * isi;
            J4<T> J;
            if (WIDE)
                J = staged_jones<T, 4>(R, Luvw, Lpa, Lant, Lch, c, ia, Lsd[4 * isi], Lsd[4 * isi + 1],
                                       Lsd[4 * isi + 2], reinterpret_cast<const J4<T> *>(o + 8)[0]);
            else
                J = item_forward<T>(R, Luvw + 3 * ia, Lpa[2 * ia], Lpa[2 * ia + 1], an[0], an[1], an[2],
                                    an[3], c, Lch[3], Lch[4], Lsd[4 * isi], Lsd[4 * isi + 1], Lsd[4 * isi + 2],
                                    o[0], o[1], o[2], o[3], SrcBright<T>{o[4], o[5], o[6], o[7]}, ist);
            T *xrec = Xr + isi * XPS + 10 * ia;
            if (SHAPED) {
#pragma unroll
                for (int x = 0; x < 4; ++x) {
                    xrec[2 * x] = J.v[x].re;
                    xrec[2 * x + 1] = J.v[x].im;
                }
            } else {
#pragma unroll
                for (int k = 0; k < 2; ++k)
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        Xr[xi_(2 * isi + j, 2 * ia + k)] = J.v[2 * k + j].re;
                        Xi[xi_(2 * isi + j, 2 * ia + k)] = J.v[2 * k + j].im;
                    }
            }
            if (SHAPED) {
                // shape coordinates, linear in the antenna's (u, v):
                // f_pq = F((x_q - x_p)^2 + (y_q - y_p)^2)
                const T *h = Lsh + 8 * isi;
                const double u = Luvw[3 * ia], v = Luvw[3 * ia + 1];
                double x, y;
                if (h[0] == T(1)) {
                    x = (u * double(h[4]) - v * double(h[3])) * double(h[2]) * adj_gauss_pre<T>();
                    y = (u * double(h[3]) + v * double(h[4])) * double(h[1]) * adj_gauss_pre<T>();
                } else {
                    const double e1 = h[5], e2 = h[6];
                    x = (u * (1.0 + e1) + v * e2) * double(h[1]);
                    y = (u * e2 + v * (1.0 - e1)) * double(h[1]);
                }
                xrec[8] = T(x);
                xrec[9] = T(y);
            }
        }
        for (int i = tid; i < (SHAPED ? 0 : 2 * st * rows); i += kThreads) {
            const int n = (i >> 2) % (2 * st);
            const int r = (i / (8 * st)) * 4 + (i & 3);
            if (n >= 2 * nst || r >= 2 * NA) { Xr[i] = T(0); Xi[i] = T(0); }
        }
        __syncthreads();
        // ---- Jb = sum_q f W_pq J_q
        if (!SHAPED) {
            typedef typename Mfma<T>::acc_t acc_t;
            const int nR = rows / 16, ncb = (2 * st) / 16, nout = nR * ncb;
            const int i16 = lane & 15, g = lane >> 4;
            // three real products per complex one, as in k_adj_point:
            // Re Jb = P1 - P2, Im Jb = P3 - P1 - P2
            acc_t ore[kAdjMaxTiles], oim[kAdjMaxTiles];
            if (wave < kAdjMfmaWaves) {
#pragma unroll
                for (int j = 0; j < kAdjMaxTiles; ++j) {
                    acc_t p1 = acc_t{T(0), T(0), T(0), T(0)}, p2 = p1, p3 = p1;
                    const int o = wave + kAdjMfmaWaves * j;
                    if (o < nout) {
                        const int Rb = o / ncb, Cb = o % ncb;
                        const int wrow = 16 * Rb + Mfma<T>::arow(i16), xn = 16 * Cb + i16;
                        for (int kk = 0; kk < rows; kk += 16) {
                            const T *wr = Wr + wi_(wrow, kk + 4 * g), *wi = Wi + wi_(wrow, kk + 4 * g);
                            const T *xr = Xr + xi_(xn, kk + 4 * g), *xi = Xi + xi_(xn, kk + 4 * g);
                            T a4[4], b4[4], x4[4], y4[4];
#pragma unroll
                            for (int m = 0; m < 4; ++m) {
                                a4[m] = wr[m]; b4[m] = wi[m];
                                x4[m] = xr[m]; y4[m] = xi[m];
                            }
#pragma unroll
                            for (int m = 0; m < 4; ++m) {
                                p1 = Mfma<T>::mma(a4[m], x4[m], p1);
                                p2 = Mfma<T>::mma(b4[m], y4[m], p2);
                                p3 = Mfma<T>::mma(a4[m] + b4[m], x4[m] + y4[m], p3);
                            }
                        }
                    }
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        ore[j][r] = p1[r] - p2[r];
                        oim[j][r] = p3[r] - p1[r] - p2[r];
                    }
                }
            }
            __syncthreads();                       // every X read done
            if (wave < kAdjMfmaWaves) {
#pragma unroll
                for (int j = 0; j < kAdjMaxTiles; ++j) {
                    const int o = wave + kAdjMfmaWaves * j;
                    if (o >= nout) continue;
                    const int Rb = o / ncb, Cb = o % ncb;
                    const int n = 16 * Cb + i16, si = n >> 1, jc = n & 1;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = 16 * Rb + out_row(lane, r);
                        const int p = row >> 1, ip = row & 1;
                        if (p < NA && si < nst) Jb[si * NA + p].v[2 * ip + jc] = cx<T>{ore[j][r], oim[j][r]};
                    }
                }
            }
        } else {
            // one (source, antenna p) per thread: st * NA <= threads. The
            // shape-factor adjoint of a pair is the same in both orders (f
            // and its derivatives are even in the pair's (u, v), and
            // Re <J_p, W_pq J_q> = Re <J_q, W_qp J_p>): it is summed over the
            // pairs q > p only and counted twice. So that the pair loop has
            // no branch that diverges inside a wave (a wave holds several
            // antennas p), q runs in three segments bounded by the wave's
            // smallest and largest p: q below every lane's p (no moments),
            // q among the wave's antennas (moments weighted by q > p), q
            // above every lane's p (moments).
            const bool mine = tid < nst * NA;
            const int si = mine ? tid % nst : 0, p = mine ? tid / nst : 0;
            int plo = mine ? p : NA, phi = mine ? p : -1;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                plo = min(plo, __shfl_xor(plo, o));
                phi = max(phi, __shfl_xor(phi, o));
            }
            const int qa = min(NA, plo), qb = max(qa, min(NA, phi + 1));
            J4<T> acc = jzero<T>();
            T sh0 = T(0), sh1 = T(0), sh2 = T(0);
            if (mine) {
                typedef T v2 __attribute__((ext_vector_type(2)));
                const T *h = Lsh + 8 * si;
                const T *xs = Xr + si * XPS;            // this source's X records
                const J4<T> Jp = jload_pairs(xs + 10 * p);
                const v2 xyp = *reinterpret_cast<const v2 *>(xs + 10 * p + 8);
                const v2 uvp = *reinterpret_cast<const v2 *>(Luv + 2 * p);
                const T xp = xyp.x, yp = xyp.y, up = uvp.x, vp = uvp.y;
                // pair q's X record and (u, v): one running pointer each,
                // fixed offsets within a record (q = 0, 1, ... in order)
                const T *xq = xs;
                const T *uq = Luv;
                auto advance = [&]() {
                    xq += 10;
                    uq += 2;
                };
                auto accumulate = [&](T f, const J4<T> &Y) {
#pragma unroll
                    for (int x = 0; x < 4; ++x) {
                        acc.v[x].re = fma(f, Y.v[x].re, acc.v[x].re);
                        acc.v[x].im = fma(f, Y.v[x].im, acc.v[x].im);
                    }
                };
                // Each pair's shape-factor adjoint is t_q = f Re <J_p, W_pq J_q>
                // (D = 1 + r2 for sersic: t_q = f Re <..> / D) times a form
                // linear in (dx du, dx dv, dy du, dy dv) and, for sersic, r2:
                // those moments are summed over q > p and combined once.
                T m_xu = T(0), m_xv = T(0), m_yu = T(0), m_yv = T(0), m_r = T(0);
                auto moments = [&](T tq, T dx, T dy) {
                    const v2 uv = *reinterpret_cast<const v2 *>(uq);
                    const T du = uv.x - up, dv = uv.y - vp;
                    const T tx = tq * dx, ty = tq * dy;
                    m_xu = fma(tx, du, m_xu);
                    m_xv = fma(tx, dv, m_xv);
                    m_yu = fma(ty, du, m_yu);
                    m_yv = fma(ty, dv, m_yv);
                };
                // one pair (W = W_pq) of kind GAUSS; MOM 0: no moments, 1:
                // moments weighted by q > p, 2: moments
                auto pair = [&](auto gauss_tag, auto mom_tag, int q, const J4<T> &W) {
                    constexpr bool GAUSS = decltype(gauss_tag)::value;
                    constexpr int MOM = decltype(mom_tag)::value;
                    const J4<T> Y = jmul(W, jload_pairs(xq));
                    const v2 xy = *reinterpret_cast<const v2 *>(xq + 8);
                    const T dx = xy.x - xp, dy = xy.y - yp;
                    const T r2 = fma(dx, dx, dy * dy);
                    T f, id = T(1);
                    if constexpr (GAUSS) {
                        if constexpr (sizeof(T) == 8) f = exp_neg(r2, Ltab);
                        else f = gauss_of_scaled(r2);
                    } else {
                        f = sersic_factor(r2, id);
                    }
                    accumulate(f, Y);
                    if constexpr (MOM > 0) {
                        T tq = f * jdot_re(Jp, Y);
                        if constexpr (!GAUSS) tq *= id;
                        if constexpr (MOM == 1) tq = q > p ? tq : T(0);
                        moments(tq, dx, dy);
                        if constexpr (!GAUSS) m_r = fma(tq, r2, m_r);
                    }
                    advance();
                };
                // two pairs in flight in float (registers to spare), one in
                // float64
                auto wblock = [&](const T *w) { return jload(reinterpret_cast<const cx<T> *>(w)); };
                auto walk = [&](auto gauss_tag) {
                    typedef std::integral_constant<int, 0> M0;
                    typedef std::integral_constant<int, 1> M1;
                    typedef std::integral_constant<int, 2> M2;
                    // q below the wave's antennas: W_pq = (block (q, p))^H,
                    // down column p of the triangle
                    const T *wc = Wr + 8 * p;           // block (0, p)
                    int step = 8 * NA - 4;              // to block (q + 1, p)
#pragma unroll (sizeof(T) == 4 ? 2 : 1)
                    for (int q = 0; q < qa; ++q) {
                        pair(gauss_tag, M0(), q, jherm(wblock(wc)));
                        wc += step;
                        step -= 8;
                    }
                    // q among the wave's antennas: either block, per lane
                    for (int q = qa; q < qb; ++q) {
                        const int lo = min(p, q), hi = max(p, q);
                        const J4<T> B = wblock(Wr + adjs_wrow(lo, NA) + 8 * (hi - lo));
                        pair(gauss_tag, M1(), q, q < p ? jherm(B) : B);
                    }
                    // q above the wave's antennas: block (p, q) along row p
                    const T *wr = Wr + adjs_wrow(p, NA) + 8 * (qb - p);
#pragma unroll (sizeof(T) == 4 ? 2 : 1)
                    for (int q = qb; q < NA; ++q) {
                        pair(gauss_tag, M2(), q, wblock(wr));
                        wr += 8;
                    }
                };
                if (h[0] == T(1)) {
                    walk(std::true_type());
                    // gaussian: f = exp(-r2); d f / d(el, em, R) = -2 f k
                    // (dy du - R dx dv, R dx du + dy dv, dx (du em - dv el));
                    // dx, dy here are scaled by g = adj_gauss_pre; x 2 for
                    // the two orders of each pair
                    const T k = h[1], el = h[3], em = h[4], Rg = h[5];
                    const T c4 = T(-4) * k / T(adj_gauss_pre<T>());
                    sh0 = c4 * (m_yu - Rg * m_xv);
                    sh1 = c4 * fma(Rg, m_xu, m_yv);
                    sh2 = c4 * (em * m_xu - el * m_xv);
                } else {
                    walk(std::false_type());
                    // sersic: D = 1 + r2, f = D^-3/2, c = -3 f / D; d f / d(e1,
                    // e2, R) = c (kR' (dx du - dy dv) + a1 r2, kR' (dx dv + dy du)
                    // + a2 r2, k/den (dx dx0 + dy dy0)) with dx0 = du (1 + e1)
                    // + dv e2, dy0 = du e2 + dv (1 - e1); x 2 for the two
                    // orders of each pair
                    const T kR = h[1], a1 = h[2], a2 = h[3], kd = h[4], e1 = h[5], e2 = h[6];
                    sh0 = T(-6) * fma(kR, m_xu - m_yv, a1 * m_r);
                    sh1 = T(-6) * fma(kR, m_xv + m_yu, a2 * m_r);
                    sh2 = T(-6) * kd * ((T(1) + e1) * m_xu + e2 * (m_xv + m_yu) + (T(1) - e1) * m_yv);
                }
            }
            if (!WIDE) __syncthreads();            // every X read done: Jb goes over X
            if (mine) {
                if (!WIDE) Jb[si * NA + p] = acc;
                atomicAdd(sac + si * 10 + 7, double(sh0));
                atomicAdd(sac + si * 10 + 8, double(sh1));
                atomicAdd(sac + si * 10 + 9, double(sh2));
                if (WIDE) {
                    // the item's backward, with the forward state recomputed
                    const double *an = Lant + 4 * p;
                    const T *o = Lst + 16 * si;
                    (void)item_forward<T>(R, Luvw + 3 * p, Lpa[2 * p], Lpa[2 * p + 1], an[0], an[1], an[2],
                                          an[3], c, Lch[3], Lch[4], Lsd[4 * si], Lsd[4 * si + 1],
                                          Lsd[4 * si + 2], o[0], o[1], o[2], o[3],
                                          SrcBright<T>{o[4], o[5], o[6], o[7]}, ist);
                    double g[7];
                    item_backward<T>(ist,
                                     item_chain<T>(R, Luvw + 3 * p, Lpa[2 * p], Lpa[2 * p + 1], an[2], an[3],
                                                   Lch[3], Lch[4], Lsd[4 * si], Lsd[4 * si + 1],
                                                   Lsd[4 * si + 3], ist.okl, ist.okm),
                                     acc, o[7], g);
#pragma unroll
                    for (int x = 0; x < 7; ++x) atomicAdd(sac + si * 10 + x, g[x]);
                }
            }
        }
        __syncthreads();
        // ---- backward of J = E K sqrt(B) per (source, antenna)
        if (own && !WIDE) {
            double g[7];
            const double *an = Lant + 4 * ia;
            item_backward<T>(ist,
                             item_chain<T>(R, Luvw + 3 * ia, Lpa[2 * ia], Lpa[2 * ia + 1], an[2], an[3], Lch[3],
                                           Lch[4], Lsd[4 * isi], Lsd[4 * isi + 1], Lsd[4 * isi + 3], ist.okl,
                                           ist.okm),
                             Jb[isi * NA + ia], Lst[16 * isi + 7], g);
#pragma unroll
            for (int x = 0; x < 7; ++x) atomicAdd(sac + isi * 10 + x, g[x]);
        }
        if (!WIDE) __syncthreads();
        if (s0 + st < send && s0 + 2 * st <= wbase + S.sw) continue;   // window not full
        // ---- this workgroup's share of the window's gradients
        const int nwin = s0 + nst - wbase;
        for (int si = tid; si < nwin; si += kThreads) {
            const int s = wbase + si;
            double *v = sacc + si * 10;
            const double gs = A.gscale;
            if (A.g_lm) {
                atomicAdd(A.g_lm + 2 * uint64_t(s), gs * v[0]);
                atomicAdd(A.g_lm + 2 * uint64_t(s) + 1, gs * v[1]);
            }
            if (A.g_stokes)
                for (int x = 0; x < 4; ++x) atomicAdd(A.g_stokes + (uint64_t(s) * NT + t) * 4 + x, gs * v[2 + x]);
            if (A.g_alpha) atomicAdd(A.g_alpha + uint64_t(s) * NT + t, gs * v[6]);
            if (SHAPED) {
                if (s < np + ng) {
                    if (A.g_gauss)
                        for (int x = 0; x < 3; ++x) atomicAdd(A.g_gauss + (s - np) + x * NGS, gs * v[7 + x]);
                } else if (A.g_sersic) {
                    for (int x = 0; x < 3; ++x) atomicAdd(A.g_sersic + (s - np - ng) + x * NSS, gs * v[7 + x]);
                }
            }
            for (int x = 0; x < 10; ++x) v[x] = 0.0;
        }
        wbase = s0 + st;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------
// Pipelined point pass. Producer waves and matrix-core waves work on
// different source tiles at once, so the product overlaps the vector work
// instead of following it:
//   phase p (between two barriers)
//     producers (waves 0-7): X of tile p (Jones only) into X buffer p % 2;
//       the backward of tile p - 2 (forward state recomputed, Jb from Jb
//       buffer p % 2) into the gradient sums; the gradient window that
//       ended with tile p - 3 to global memory;
//     matrix waves (one per 16-row block of W): Jb of tile p - 1 =
//       W X[(p - 1) % 2] into Jb buffer (p - 1) % 2; the first of them
//       also stages the parameters of tile p + 1 (slot (p + 1) % 4).
// W stays in the matrix waves' registers for the whole (t, c): each holds
// the A operand of its 16-row block, gathered from Vb through the
// baseline map. X columns are n = j st + s (j the Jones column, s the
// source in the tile); X[row][n] sits at (row / 4) XS + 4 n + row % 4,
// the row-block stride XS = 8 st + 4 padded so that producers writing
// different antennas hit different LDS banks. Jb is kept as eight planes
// (Re/Im of the four Jones entries) of [antenna][source].

template <typename T> struct AdjPoint {
    static constexpr int kMaxRb = sizeof(T) == 8 ? 4 : 8;   // 2 na <= 16 kMaxRb
    static constexpr int kMaxCb = 2;                         // st <= 16
    static constexpr int kProducers = 512;
    static constexpr int kThreads = kProducers + 64 * kMaxRb;
    static constexpr int kTaps = sizeof(T) == 4 ? 4 : 2;     // beam taps in flight
    static constexpr int kParams = 20;                       // per source and slot
    // which role stages the next tile's parameters: in float64 the matrix
    // waves (registers to spare), in float the producers (the matrix
    // waves' W leaves no room)
    static constexpr bool kParamsOnMatrix = sizeof(T) == 8;
};

struct AdjPointShape {
    int rows;    // 2 na rounded up to 16
    int nrb;     // rows / 16: matrix waves
    int st;      // sources per tile: 8 or 16, st * na <= 512
    int xs;      // X row-block stride (T)
    int wt;      // tiles per gradient window
};

static __host__ __device__ inline size_t adjp_align(size_t b) { return (b + 31) & ~size_t(31); }

template <typename T>
struct AdjPointLds {
    size_t params, x, jb, sacc, total;
    __host__ __device__ AdjPointLds(int na, const AdjPointShape &S) {
        params = adjg_setup_bytes(na);
        x = params + adjp_align(sizeof(double) * 4 * AdjPoint<T>::kParams * S.st);
        jb = x + adjp_align(sizeof(T) * 2 * 2 * (S.rows / 4) * S.xs);
        sacc = jb + adjp_align(sizeof(T) * 2 * 8 * na * S.st);
        total = sacc + sizeof(double) * 2 * S.wt * S.st * 7;
    }
};

template <typename T, bool BATCHED = false>
__global__ void __launch_bounds__(AdjPoint<T>::kThreads)
k_adj_point(AdjArgs<T> A, AdjPointShape S) {
    typedef AdjPoint<T> C;
    typedef typename Mfma<T>::acc_t acc_t;
    extern __shared__ __align__(32) unsigned char smem_raw[];
    // in a launch over several samples: this workgroup's sample and slot
    const SampleAt at = enter_sample_if<BATCHED>(A);
    const RimeArgs<T> &R = A.R;
    const mbx_dims &d = R.d;
    const int NA = d.na.local_size;
    const uint64_t NT = d.ntime.local_size, NC = d.nchan.local_size, NBL = d.nbl.local_size;
    const int nt_ext = dext(d.ntime), nc_ext = dext(d.nchan), nb_ext = dext(d.nbl);
    const int tid = threadIdx.x;
    int t, c;
    if (!tc_coords(tc_per_xcd(nt_ext, nc_ext), nt_ext, nc_ext, t, c, at.bx)) return;
    const int rows = S.rows, st = S.st, xs = S.xs, nth = C::kProducers + 64 * S.nrb;
    const AdjPointLds<T> L(NA, S);

    double *Luvw = reinterpret_cast<double *>(smem_raw);
    double *Lpa = Luvw + 3 * NA;
    double *Lant = Lpa + 2 * NA;
    double *Lch = Lant + 4 * NA;
    double *Lprm = reinterpret_cast<double *>(smem_raw + L.params);   // [4][kParams][st]
    T *Xbuf = reinterpret_cast<T *>(smem_raw + L.x);                  // [2][re, im][rows/4][xs]
    T *Jbuf = reinterpret_cast<T *>(smem_raw + L.jb);                 // [2][8][na][st]
    double *sacc = reinterpret_cast<double *>(smem_raw + L.sacc);     // [2][wt st][7]
    const int xplane = (rows / 4) * xs, jplane = 8 * NA * st, wsrc = S.wt * st;
    auto prm = [&](int tile, int f) { return Lprm + ((tile & 3) * C::kParams + f) * st; };

    tile_setup<T>(R, Luvw, Lpa, Lch, t, c, tid, nth);
    stage_antennas<T>(R, Lant, t, c, tid, nth);
    // X rows past 2 na read as zero in every tile (producers never write them)
    for (int i = tid; i < 2 * 2 * xplane; i += nth) Xbuf[i] = T(0);
    for (int i = tid; i < 2 * wsrc * 7; i += nth) sacc[i] = 0.0;

    const int np = dext(d.npsrc);
    const int ntiles = (np + st - 1) / st;
    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const bool producer = wave < C::kProducers / 64;
    const int i16 = lane & 15, g = lane >> 4;
    auto tile_n = [&](int it) { return min(st, np - it * st); };

    // parameters of tile `it` by the first matrix wave: l, m, n - 1, I, Q,
    // U, V, sqrt(P) (P = (f / f_ref)^alpha), sqrt(B) (8 reals), 1 / n, then
    // s0, 1 / t0, 1 / s0 (src_bright)
    auto stage_params = [&](int it, int si) {
        if (it >= ntiles || si >= tile_n(it)) return;
        const int s = it * st + si;
        double l0, m0, nm1;
        const J4<T> B = source_sqrtb<T>(R, Lch, t, s, l0, m0, nm1);
        const T *Sk = R.stokes + (uint64_t(s) * NT + t) * 4;
        double v[C::kParams];
        v[0] = l0; v[1] = m0; v[2] = nm1;
        v[3] = Sk[0]; v[4] = Sk[1]; v[5] = Sk[2]; v[6] = Sk[3];
        const SrcBright<T> sbr = src_bright<T>(Sk[0], Sk[1], Sk[2], Sk[3],
                                               exp_t<T>(double(R.alpha[uint64_t(s) * NT + t]) * Lch[4]));
        v[7] = sbr.rP;
        v[17] = sbr.s0; v[18] = sbr.it0; v[19] = sbr.is0;
#pragma unroll
        for (int x = 0; x < 4; ++x) { v[8 + 2 * x] = B.v[x].re; v[9 + 2 * x] = B.v[x].im; }
        v[16] = 1.0 / (nm1 + 1.0);
#pragma unroll
        for (int f = 0; f < C::kParams; ++f) prm(it, f)[si] = v[f];
    };

    auto flush = [&](int done, int wbase) {
        const int ns = min(np, (done + 1) * st) - wbase * st;
        double *sa = sacc + ((done / S.wt) & 1) * wsrc * 7;
        const double gs = A.gscale;
        for (int i = tid; i < ns; i += C::kProducers) {
            const uint64_t s = uint64_t(wbase) * st + i;
            double *v = sa + i * 7;
            if (A.g_lm) {
                atomicAdd(A.g_lm + 2 * s, gs * v[0]);
                atomicAdd(A.g_lm + 2 * s + 1, gs * v[1]);
            }
            if (A.g_stokes)
                for (int x = 0; x < 4; ++x) atomicAdd(A.g_stokes + (s * NT + t) * 4 + x, gs * v[2 + x]);
            if (A.g_alpha) atomicAdd(A.g_alpha + s * NT + t, gs * v[6]);
#pragma unroll
            for (int x = 0; x < 7; ++x) v[x] = 0.0;
        }
    };

    // The two roles run separate loops with the same barriers, so that the
    // matrix waves' W registers are not live in the producers' code.
    if (producer) {
        if (!C::kParamsOnMatrix && wave == 0) stage_params(0, lane);
        __syncthreads();
        int wbase = 0;                 // first tile of the window flushed next
        for (int p = 0; p < ntiles + 2; ++p) {
            // ---- X of tile p: item (si, a), consecutive threads take
            // consecutive sources
            if (p < ntiles) {
                const int nst = tile_n(p);
                if (tid < nst * NA) {
                    const int si = tid % nst, a = tid / nst;
                    J4<T> B;
#pragma unroll
                    for (int x = 0; x < 4; ++x) B.v[x] = {T(prm(p, 8 + 2 * x)[si]), T(prm(p, 9 + 2 * x)[si])};
                    const J4<T> J = staged_jones<T, C::kTaps>(R, Luvw, Lpa, Lant, Lch, c, a, prm(p, 0)[si],
                                                               prm(p, 1)[si], prm(p, 2)[si], B);
                    T *xr = Xbuf + (p & 1) * 2 * xplane, *xi = xr + xplane;
#pragma unroll
                    for (int k = 0; k < 2; ++k)
#pragma unroll
                        for (int j = 0; j < 2; ++j) {
                            const int row = 2 * a + k, o = (row >> 2) * xs + 4 * (j * st + si) + (row & 3);
                            xr[o] = J.v[2 * k + j].re;
                            xi[o] = J.v[2 * k + j].im;
                        }
                }
            }
            // ---- backward of tile p - 2
            if (p >= 2) {
                const int it = p - 2, nst = tile_n(it);
                if (tid < nst * NA) {
                    const int si = tid % nst, a = tid / nst;
                    const double *an = Lant + 4 * a;
                    const T *jb = Jbuf + (it & 1) * jplane + a * st + si;
                    const int js = NA * st;
                    J4<T> Jb;
#pragma unroll
                    for (int x = 0; x < 4; ++x) Jb.v[x] = {jb[(2 * x) * js], jb[(2 * x + 1) * js]};
                    double gr[7];
                    ItemState<T> ist;
                    (void)item_forward<T>(R, Luvw + 3 * a, Lpa[2 * a], Lpa[2 * a + 1], an[0], an[1], an[2],
                                          an[3], c, Lch[3], Lch[4], prm(it, 0)[si], prm(it, 1)[si],
                                          prm(it, 2)[si], T(prm(it, 3)[si]), T(prm(it, 4)[si]),
                                          T(prm(it, 5)[si]), T(prm(it, 6)[si]),
                                          SrcBright<T>{T(prm(it, 7)[si]), T(prm(it, 17)[si]), T(prm(it, 18)[si]),
                                                       T(prm(it, 19)[si])},
                                          ist);
                    item_backward<T>(ist,
                                     item_chain<T>(R, Luvw + 3 * a, Lpa[2 * a], Lpa[2 * a + 1], an[2], an[3],
                                                   Lch[3], Lch[4], prm(it, 0)[si], prm(it, 1)[si],
                                                   prm(it, 16)[si], ist.okl, ist.okm),
                                     Jb, T(prm(it, 19)[si]), gr);
                    double *sa = sacc + ((it / S.wt) & 1) * wsrc * 7 + ((it % S.wt) * st + si) * 7;
#pragma unroll
                    for (int x = 0; x < 7; ++x) atomicAdd(sa + x, gr[x]);
                }
            }
            if (!C::kParamsOnMatrix && wave == 0) stage_params(p + 1, lane);
            // ---- the window that ended with tile p - 3 (summed by phase p - 1)
            const int done = p - 3;
            if (done >= 0 && done % S.wt == S.wt - 1) {
                flush(done, wbase);
                wbase = done + 1;
            }
            __syncthreads();
        }
        // the last window (summed in the last phase)
        if (ntiles > 0 && wbase < ntiles) flush(ntiles - 1, wbase);
    } else {
        // W rows of this matrix wave's block: lane (i16, g) supplies
        // A[row][k] = W[16 rb + arow(i16)][16 kb + 4 g + m] for m = 0..3
        T Wr[C::kMaxRb][4], Wi[C::kMaxRb][4];
        const int rb = wave - C::kProducers / 64;
        {
            const int wrow = 16 * rb + Mfma<T>::arow(i16);
            const int p = wrow >> 1, i = wrow & 1;
#pragma unroll
            for (int kb = 0; kb < C::kMaxRb; ++kb)
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const int col = 16 * kb + 4 * g + m, q = col >> 1, k = col & 1;
                    T wr = T(0), wi = T(0);
                    if (kb < S.nrb && p < NA && q < NA) {
                        const int e = A.bmap[p * NA + q];
                        const int b = e & kBlMask;
                        if (e >= 0 && b < nb_ext) {
                            const cx<T> *v = A.vbar + ((uint64_t(t) * NBL + b) * NC + c) * 4;
                            if (e & MBX_BL_AUTO) {                 // Vb + Vb^H
                                const cx<T> a = v[2 * i + k], h = v[2 * k + i];
                                wr = a.re + h.re; wi = a.im - h.im;
                            } else if (e & MBX_BL_REVERSED) {      // (Vb^H)[i][k]
                                const cx<T> h = v[2 * k + i];
                                wr = h.re; wi = -h.im;
                            } else {
                                const cx<T> a = v[2 * i + k];
                                wr = a.re; wi = a.im;
                            }
                        }
                    }
                    Wr[kb][m] = wr; Wi[kb][m] = wi;
                }
        }
        if (C::kParamsOnMatrix && rb == 0) stage_params(0, lane);
        __syncthreads();
        for (int p = 0; p < ntiles + 2; ++p) {
            // ---- Jb of tile p - 1 on the matrix cores
            if (p >= 1 && p - 1 < ntiles) {
                const int it = p - 1, nst = tile_n(it);
                const T *xr = Xbuf + (it & 1) * 2 * xplane, *xi = xr + xplane;
                T *jb = Jbuf + (it & 1) * jplane;
#pragma unroll
                for (int cb = 0; cb < C::kMaxCb; ++cb) {
                    if (cb >= st / 8) break;
                    // three real products per complex one (Gauss):
                    // P1 = Re W Re X, P2 = Im W Im X, P3 = (Re W + Im W)(Re X + Im X);
                    // Re Jb = P1 - P2, Im Jb = P3 - P1 - P2
                    acc_t p1 = acc_t{T(0), T(0), T(0), T(0)}, p2 = p1, p3 = p1;
                    const int xn = 16 * cb + i16;
#pragma unroll
                    for (int kb = 0; kb < C::kMaxRb; ++kb) {
                        if (kb >= S.nrb) break;
                        const int o = (4 * kb + g) * xs + 4 * xn;
                        T x4[4], y4[4];
#pragma unroll
                        for (int m = 0; m < 4; ++m) { x4[m] = xr[o + m]; y4[m] = xi[o + m]; }
#pragma unroll
                        for (int m = 0; m < 4; ++m) {
                            p1 = Mfma<T>::mma(Wr[kb][m], x4[m], p1);
                            p2 = Mfma<T>::mma(Wi[kb][m], y4[m], p2);
                            p3 = Mfma<T>::mma(Wr[kb][m] + Wi[kb][m], x4[m] + y4[m], p3);
                        }
                    }
                    acc_t ore, oim;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        ore[r] = p1[r] - p2[r];
                        oim[r] = p3[r] - p1[r] - p2[r];
                    }
                    const int j = xn / st, si = xn % st;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = 16 * rb + out_row(lane, r), a = row >> 1, i = row & 1;
                        if (a < NA && si < nst) {
                            const int comp = 2 * (2 * i + j);
                            jb[(comp * NA + a) * st + si] = ore[r];
                            jb[((comp + 1) * NA + a) * st + si] = oim[r];
                        }
                    }
                }
            }
            if (C::kParamsOnMatrix && rb == 0) stage_params(p + 1, lane);
            __syncthreads();
        }
    }
}

template <typename T>
static bool adjp_shape(const mbx_dims &d, AdjPointShape &S, size_t &lds) {
    typedef AdjPoint<T> C;
    const int na = d.na.local_size;
    if (na <= 0) return false;
    S.rows = ((2 * na + 15) / 16) * 16;
    S.nrb = S.rows / 16;
    if (S.nrb > C::kMaxRb) return false;
    S.st = 16 * na <= C::kProducers ? 16 : (8 * na <= C::kProducers ? 8 : 0);
    if (S.st == 0) return false;
    S.xs = 8 * S.st + 4;
    static size_t budget = gemm_lds_budget();
    // the widest gradient window that fits, up to 32 tiles
    S.wt = 1;
    if (AdjPointLds<T>(na, S).total > budget) return false;
    while (S.wt < 32) {
        ++S.wt;
        if (AdjPointLds<T>(na, S).total > budget) { --S.wt; break; }
    }
    lds = AdjPointLds<T>(na, S).total;
    return true;
}

// Tile of k_adj_gemm: one (source, antenna) item per thread, st sources
// (a multiple of 8; at most 16 in the point pass, whose product tiles it
// bounds, 32 in the shaped pass), the most that fit the LDS budget; then
// the widest gradient window up to 32 tiles.
template <typename T>
static bool adjg_shape(const mbx_dims &d, AdjGemmShape &S, size_t &lds, bool shaped,
                       int threads = kAdjThreads) {
    const int na = d.na.local_size;
    if (na <= 0 || na > 64) return false;
    S.rows = ((2 * na + 15) / 16) * 16;
    S.st = std::min(shaped ? 32 : 16, (threads / na) / 8 * 8);
    static size_t budget = gemm_lds_budget();
    auto fits = [&](int st, int sw) { return adjg_lds_bytes<T>(na, S.rows, st, sw, shaped) <= budget; };
    while (S.st > 8 && !fits(S.st, S.st)) S.st -= 8;
    if (S.st < 8 || !fits(S.st, S.st)) return false;
    // the widest flush window that fits, up to 32 tiles
    S.sw = S.st;
    while (S.sw < 32 * S.st && fits(S.st, S.sw + S.st)) S.sw += S.st;
    lds = adjg_lds_bytes<T>(na, S.rows, S.st, S.sw, shaped);
    return shaped || (S.rows / 16) * (S.st / 8) <= kAdjMfmaWaves * kAdjMaxTiles;
}

template <typename T>
bool adjoint_gemm_supported(const mbx_dims &d) {
    AdjGemmShape S;
    size_t lds;
    return adjg_shape<T>(d, S, lds, false) && adjg_shape<T>(d, S, lds, true);
}

// The point pass (if any point sources) and the shaped pass (if any
// gaussian or sersic sources), nwg workgroups each.
template <typename T, bool BATCHED>
static int launch_adjoint_passes(const AdjArgs<T> &Ag, const AdjGemmShape &S, size_t lds, unsigned nwg,
                                 hipStream_t stream) {
    const mbx_dims &d = Ag.R.d;
    if (dext(d.npsrc) > 0) {
        AdjPointShape PS;
        size_t plds;
        if (adjp_shape<T>(d, PS, plds))
            hipLaunchKernelGGL((k_adj_point<T, BATCHED>), dim3(nwg),
                               dim3(AdjPoint<T>::kProducers + 64 * PS.nrb), plds, stream, Ag, PS);
        else
            hipLaunchKernelGGL((k_adj_gemm<T, false, false, BATCHED>), dim3(nwg), dim3(kAdjThreads), lds,
                               stream, Ag, S);
        int r = hip_status(hipGetLastError(), "adjoint point gemm launch");
        if (r) return r;
    }
    if (dext(d.nsrc) > dext(d.npsrc)) {
        // float: sixteen waves with the forward state recomputed (WIDE);
        // float64, and float where the wide tiles do not fit: the state kept
        // in registers (eight waves)
        AdjGemmShape SW;
        size_t wlds;
        bool wide = false;
        if constexpr (sizeof(T) == 4) {
            constexpr int wth = adjg_threads<T, true>();
            wide = adjg_shape<T>(d, SW, wlds, true, wth);
            if (wide)
                hipLaunchKernelGGL((k_adj_gemm<T, true, true, BATCHED>), dim3(nwg), dim3(wth), wlds, stream,
                                   Ag, SW);
        }
        if (!wide) {
            if (!adjg_shape<T>(d, SW, wlds, true))
                return set_error(MBX_E_UNSUPPORTED, "adjoint gemm: shaped layout");
            hipLaunchKernelGGL((k_adj_gemm<T, true, false, BATCHED>), dim3(nwg), dim3(kAdjThreads), wlds,
                               stream, Ag, SW);
        }
        return hip_status(hipGetLastError(), "adjoint shaped launch");
    }
    return MBX_OK;
}

template <typename T>
int launch_adjoint_gemm(const AdjArgs<T> &A, hipStream_t stream, unsigned nrep) {
    const mbx_dims &d = A.R.d;
    AdjGemmShape S;
    size_t lds;
    if (!adjg_shape<T>(d, S, lds, false)) return set_error(MBX_E_UNSUPPORTED, "adjoint gemm: layout");
    // per sample: one workgroup per (time, channel) in the XCD order
    const unsigned nwg = 8 * tc_per_xcd(dext(d.ntime), dext(d.nchan)) * nrep;
    if (A.R.span.wgs && A.R.span.wgs * nrep != nwg)
        return set_error(MBX_E_DIMS, "adjoint gemm: span of %u workgroups per sample", A.R.span.wgs);
    AdjArgs<T> Ag = A;
    Ag.R.ct = 1;
    return A.R.span.wgs ? launch_adjoint_passes<T, true>(Ag, S, lds, nwg, stream)
                        : launch_adjoint_passes<T, false>(Ag, S, lds, nwg, stream);
}

template bool adjoint_gemm_supported<float>(const mbx_dims &);
template bool adjoint_gemm_supported<double>(const mbx_dims &);
template int launch_adjoint_gemm<float>(const AdjArgs<float> &, hipStream_t, unsigned);
template int launch_adjoint_gemm<double>(const AdjArgs<double> &, hipStream_t, unsigned);

}  // namespace mbx
