// Source stage of the adjoint on the matrix cores, one workgroup per
// (time t, channel c). No reference counterpart.
//
// With Vb the visibility adjoint (from k_adj_vis), the Jones adjoint of
// antenna p for source s is
//     Jb_ps = sum_q f_pq,s W_pq J_qs,
// where W is the 2na x 2na Hermitian completion of Vb: W_pq = Vb_b and
// W_qp = Vb_b^H for baseline b = (p, q), W_pp = Vb_b + Vb_b^H for an
// auto-correlation. The workgroup stages W in LDS once (planes Re, Im).
//
// Point pass (f = 1): Jb = W X is a matrix product, X[2q+k][2s+j] =
// J_qs[k][j]; in real arithmetic Re Jb = [Re W | -Im W] [Re X; Im X] and
// Im Jb = [Im W | Re W] [Re X; Im X], on v_mfma_*_16x16x4 with 16x16
// output tiles (antenna rows x (source, column)).
// Shaped pass (gaussian, sersic): one thread per (source, antenna p)
// walks q on the vector units, with W and J from LDS, and also forms the
// shape-factor adjoint d chi^2/d f_pq = Re <J_p, W_pq J_q> (summed over
// both orders of a pair, this is 2 Re <J_p, Vb J_q> per baseline).
// Both passes then run the per-(source, antenna) backward of J = E K
// sqrt(B) (antenna_backward) and reduce each source's gradient in LDS
// before one set of float64 atomics per source and workgroup.
//
// Per source tile: source parameters | X (all waves, one (source,
// antenna) item each, whose forward state -- beam Jones and its l, m
// derivatives, phase, sqrt(B0) -- stays in registers) | Jb (MFMA waves,
// or all waves) | backward from the kept state | flush per window, with a
// barrier between the phases. Jb is held in registers across the barrier
// that ends the product and then written over X's LDS.

#include "mbx_backward.hpp"
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

template <typename T>
static inline size_t adjg_lds_bytes(int na, int rows, int st, int sw) {
    return adjg_setup_bytes(na) + 2 * size_t(rows) * rows * sizeof(T)   // W planes
           + 2 * size_t(2 * st) * rows * sizeof(T)                      // X planes (then Jb)
           + size_t(sw) * 10 * sizeof(double)                           // gradient sums
           + size_t(st) * 4 * sizeof(double)                            // l, m, n - 1
           + size_t(st) * 20 * sizeof(T);                               // shape, I Q U V P, sqrt(B)
}

template <typename T, bool SHAPED>
__global__ void __launch_bounds__(kAdjThreads)
k_adj_gemm(AdjArgs<T> A, AdjGemmShape S) {
    extern __shared__ __align__(32) unsigned char smem_raw[];
    const RimeArgs<T> &R = A.R;
    const mbx_dims &d = R.d;
    const int NA = d.na.local_size;
    const uint64_t NT = d.ntime.local_size, NC = d.nchan.local_size, NBL = d.nbl.local_size;
    const int nt_ext = dext(d.ntime), nc_ext = dext(d.nchan), nb_ext = dext(d.nbl);
    const int tid = threadIdx.x;
    // XCD order of kChanGroup (mbx_tile.hpp)
    int t, c;
    if (!tc_coords(tc_per_xcd(nt_ext, nc_ext), nt_ext, nc_ext, t, c)) return;
    const int rows = S.rows, st = S.st;

    double *Luvw = reinterpret_cast<double *>(smem_raw);
    double *Lpa = Luvw + 3 * NA;
    double *Lant = Lpa + 2 * NA;
    double *Lch = Lant + 4 * NA;
    T *Wr = reinterpret_cast<T *>(smem_raw + adjg_setup_bytes(NA));
    T *Wi = Wr + rows * rows;
    // X (re, im planes): point pass [n][row] (MFMA operand reads along the
    // rows), shaped pass [row][n] (a wave reads one row for many sources);
    // n = 2 s + j, row = 2 a + k
    T *Xr = Wi + rows * rows;
    T *Xi = Xr + 2 * st * rows;
    // Point pass operands are blocked by four along the summed index, so a
    // lane's four consecutive k values are one 16/32-byte read and
    // consecutive lanes (rows of W, sources of X) read consecutive blocks:
    //   W[row][col] at ((col / 4) rows + row) 4 + col % 4,
    //   X[n][row]   at ((row / 4) 2 st + n) 4 + row % 4   (point pass)
    auto xi_ = [&](int n, int row) {
        return SHAPED ? row * (2 * st) + n : ((row >> 2) * (2 * st) + n) * 4 + (row & 3);
    };
    auto wi_ = [&](int row, int col) { return ((col >> 2) * rows + row) * 4 + (col & 3); };
    J4<T> *Jb = reinterpret_cast<J4<T> *>(Xr); // [st][NA], written over X
    double *sacc = reinterpret_cast<double *>(Xi + 2 * st * rows);   // [st][10]
    double *Lsd = sacc + S.sw * 10;                                 // [st][4]: l, m, n - 1
    T *Lsh = reinterpret_cast<T *>(Lsd + st * 4);                   // [st][4]: shape
    T *Lst = Lsh + st * 4;                                          // [st][16]: IQUV P . sqrt(B)

    tile_setup<T>(R, Luvw, Lpa, Lch, t, c, tid, kAdjThreads);
    stage_antennas<T>(R, Lant, t, c, tid, kAdjThreads);
    for (int i = tid; i < 2 * rows * rows; i += kAdjThreads) Wr[i] = T(0);
    for (int i = tid; i < S.sw * 10; i += kAdjThreads) sacc[i] = 0.0;
    __syncthreads();
    // W from this (t, c)'s visibility adjoints
    for (int b = tid; b < nb_ext; b += kAdjThreads) {
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
                    Wr[wi_(2 * p + i, 2 * p + k)] = v.re + u.re;
                    Wi[wi_(2 * p + i, 2 * p + k)] = v.im - u.im;
                } else {
                    Wr[wi_(2 * p + i, 2 * q + k)] = v.re;
                    Wi[wi_(2 * p + i, 2 * q + k)] = v.im;
                    Wr[wi_(2 * q + k, 2 * p + i)] = v.re;
                    Wi[wi_(2 * q + k, 2 * p + i)] = -v.im;
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
    int wbase = sbeg;
    for (int s0 = sbeg; s0 < send; s0 += st) {
        const int nst = min(st, send - s0);
        double *sac = sacc + (s0 - wbase) * 10;    // this tile's sums
        // ---- the tile's source parameters, once
        for (int si = tid; si < nst; si += kAdjThreads) {
            double l0, m0, nm1;
            const J4<T> B = source_sqrtb<T>(R, Lch, t, s0 + si, l0, m0, nm1);
            Lsd[4 * si] = l0; Lsd[4 * si + 1] = m0; Lsd[4 * si + 2] = nm1;
            const T *St = R.stokes + (uint64_t(s0 + si) * NTT + t) * 4;
            T *o = Lst + 16 * si;
            o[0] = St[0]; o[1] = St[1]; o[2] = St[2]; o[3] = St[3];
            o[4] = exp_t<T>(double(R.alpha[uint64_t(s0 + si) * NTT + t]) * Lch[4]);
            reinterpret_cast<J4<T> *>(o + 8)[0] = B;
        }
        __syncthreads();
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
            const J4<T> J = item_forward<T>(R, Luvw + 3 * ia, Lpa[2 * ia], Lpa[2 * ia + 1], an[0], an[1],
                                            an[2], an[3], c, Lch[3], Lch[4], Lsd[4 * isi],
                                            Lsd[4 * isi + 1], o[0], o[1], o[2], o[3], o[4], ist);
#pragma unroll
            for (int k = 0; k < 2; ++k)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    Xr[xi_(2 * isi + j, 2 * ia + k)] = J.v[2 * k + j].re;
                    Xi[xi_(2 * isi + j, 2 * ia + k)] = J.v[2 * k + j].im;
                }
        }
        for (int i = tid; i < 2 * st * rows; i += kAdjThreads) {
            const int n = SHAPED ? i % (2 * st) : (i >> 2) % (2 * st);
            const int r = SHAPED ? i / (2 * st) : (i / (8 * st)) * 4 + (i & 3);
            if (n >= 2 * nst || r >= 2 * NA) { Xr[i] = T(0); Xi[i] = T(0); }
        }
        if (SHAPED)
            for (int si = tid; si < nst; si += kAdjThreads) {
                const int s = s0 + si;
                T *o = Lsh + 4 * si;
                if (s < np + ng) {
                    const int g = s - np;
                    o[0] = R.gauss[g]; o[1] = R.gauss[g + NGS]; o[2] = R.gauss[g + 2 * NGS]; o[3] = T(1);
                } else {
                    const int g = s - np - ng;
                    const T e1 = R.sersic[g], e2 = R.sersic[g + NSS];
                    o[0] = e1; o[1] = e2; o[2] = R.sersic[g + 2 * NSS] / (T(1) - e1 * e1 - e2 * e2);
                    o[3] = T(2);
                }
            }
        __syncthreads();
        // ---- Jb = sum_q f W_pq J_q
        if (!SHAPED) {
            typedef typename Mfma<T>::acc_t acc_t;
            const int nR = rows / 16, ncb = (2 * st) / 16, nout = nR * ncb;
            const int i16 = lane & 15, g = lane >> 4;
            acc_t ore[kAdjMaxTiles], oim[kAdjMaxTiles];
            if (wave < kAdjMfmaWaves) {
#pragma unroll
                for (int j = 0; j < kAdjMaxTiles; ++j) {
                    ore[j] = acc_t{T(0), T(0), T(0), T(0)};
                    oim[j] = acc_t{T(0), T(0), T(0), T(0)};
                    const int o = wave + kAdjMfmaWaves * j;
                    if (o >= nout) continue;
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
                            ore[j] = Mfma<T>::mma(a4[m], x4[m], ore[j]);
                            ore[j] = Mfma<T>::mma(-b4[m], y4[m], ore[j]);
                            oim[j] = Mfma<T>::mma(b4[m], x4[m], oim[j]);
                            oim[j] = Mfma<T>::mma(a4[m], y4[m], oim[j]);
                        }
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
            // one (source, antenna p) per thread: st * NA <= threads
            const bool mine = tid < nst * NA;
            const int si = mine ? tid % nst : 0, p = mine ? tid / nst : 0;
            J4<T> acc = jzero<T>();
            T sh0 = T(0), sh1 = T(0), sh2 = T(0);
            if (mine) {
                const T *sp = Lsh + 4 * si;
                const int ty = int(sp[3]);
                const T kf = T((ty == 1 ? R.gauss_scale : R.two_pi_over_c) * Lch[3]);
                J4<T> Jp;
#pragma unroll
                for (int k = 0; k < 2; ++k)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        Jp.v[2 * k + j] = {Xr[xi_(2 * si + j, 2 * p + k)], Xi[xi_(2 * si + j, 2 * p + k)]};
                const double up = Luvw[3 * p], vp = Luvw[3 * p + 1];
#pragma unroll 1
                for (int q = 0; q < NA; ++q) {
                    J4<T> Wq, Jq;
#pragma unroll
                    for (int i = 0; i < 2; ++i)
#pragma unroll
                        for (int k = 0; k < 2; ++k) {
                            Wq.v[2 * i + k] = {Wr[wi_(2 * p + i, 2 * q + k)], Wi[wi_(2 * p + i, 2 * q + k)]};
                            Jq.v[2 * i + k] = {Xr[xi_(2 * si + k, 2 * q + i)], Xi[xi_(2 * si + k, 2 * q + i)]};
                        }
                    const J4<T> Y = jmul(Wq, Jq);
                    T df[3];
                    const T f = shape_factor_grad<T>(ty, T(Luvw[3 * q] - up), T(Luvw[3 * q + 1] - vp),
                                                     sp[0], sp[1], sp[2], kf, df);
#pragma unroll
                    for (int x = 0; x < 4; ++x) {
                        acc.v[x].re = fma(f, Y.v[x].re, acc.v[x].re);
                        acc.v[x].im = fma(f, Y.v[x].im, acc.v[x].im);
                    }
                    const T fb = jdot_re(Jp, Y);
                    sh0 = fma(fb, df[0], sh0); sh1 = fma(fb, df[1], sh1); sh2 = fma(fb, df[2], sh2);
                }
            }
            __syncthreads();                       // every X read done
            if (mine) {
                Jb[si * NA + p] = acc;
                atomicAdd(sac + si * 10 + 7, double(sh0));
                atomicAdd(sac + si * 10 + 8, double(sh1));
                atomicAdd(sac + si * 10 + 9, double(sh2));
            }
        }
        __syncthreads();
        // ---- backward of J = E K sqrt(B) per (source, antenna)
        if (own) {
            double g[7];
            item_backward<T>(ist, Jb[isi * NA + ia], g);
#pragma unroll
            for (int x = 0; x < 7; ++x) atomicAdd(sac + isi * 10 + x, g[x]);
        }
        __syncthreads();
        if (s0 + st < send && s0 + 2 * st <= wbase + S.sw) continue;   // window not full
        // ---- this workgroup's share of the window's gradients
        const int nwin = s0 + nst - wbase;
        for (int si = tid; si < nwin; si += kAdjThreads) {
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

template <typename T>
static bool adjg_shape(const mbx_dims &d, AdjGemmShape &S, size_t &lds) {
    const int na = d.na.local_size;
    if (na <= 0 || na > 64) return false;
    S.rows = ((2 * na + 15) / 16) * 16;
    S.st = std::min(16, (kAdjThreads / na) / 8 * 8);
    if (S.st < 8) return false;
    static size_t budget = gemm_lds_budget();
    auto fits = [&](int st, int sw) { return adjg_lds_bytes<T>(na, S.rows, st, sw) <= budget; };
    if (!fits(S.st, S.st) && S.st > 8) S.st = 8;
    if (!fits(S.st, S.st)) return false;
    // the widest flush window that fits, up to 32 tiles
    S.sw = S.st;
    while (S.sw < 32 * S.st && fits(S.st, S.sw + S.st)) S.sw += S.st;
    lds = adjg_lds_bytes<T>(na, S.rows, S.st, S.sw);
    return (S.rows / 16) * (S.st / 8) <= kAdjMfmaWaves * kAdjMaxTiles;
}

template <typename T>
bool adjoint_gemm_supported(const mbx_dims &d) {
    AdjGemmShape S;
    size_t lds;
    return adjg_shape<T>(d, S, lds);
}

template <typename T>
int launch_adjoint_gemm(const AdjArgs<T> &A, hipStream_t stream) {
    const mbx_dims &d = A.R.d;
    AdjGemmShape S;
    size_t lds;
    if (!adjg_shape<T>(d, S, lds)) return set_error(MBX_E_UNSUPPORTED, "adjoint gemm: layout");
    const unsigned nwg = 8 * tc_per_xcd(dext(d.ntime), dext(d.nchan));
    AdjArgs<T> Ag = A;
    Ag.R.ct = 1;
    if (dext(d.npsrc) > 0) {
        hipLaunchKernelGGL((k_adj_gemm<T, false>), dim3(nwg), dim3(kAdjThreads), lds, stream, Ag, S);
        int r = hip_status(hipGetLastError(), "adjoint point gemm launch");
        if (r) return r;
    }
    if (dext(d.nsrc) > dext(d.npsrc)) {
        hipLaunchKernelGGL((k_adj_gemm<T, true>), dim3(nwg), dim3(kAdjThreads), lds, stream, Ag, S);
        return hip_status(hipGetLastError(), "adjoint shaped launch");
    }
    return MBX_OK;
}

template bool adjoint_gemm_supported<float>(const mbx_dims &);
template bool adjoint_gemm_supported<double>(const mbx_dims &);
template int launch_adjoint_gemm<float>(const AdjArgs<float> &, hipStream_t);
template int launch_adjoint_gemm<double>(const AdjArgs<double> &, hipStream_t);

}  // namespace mbx
