// Adjoint of the fused RIME forward: d chi^2 / d(lm, stokes, alpha,
// gaussian and sersic shapes, gains), recomputing the Jones terms instead
// of storing per-source visibilities. No reference counterpart.
//
// Notation: for a real chi^2 and a complex quantity z, zb = d chi^2 / d z*
// (Wirtinger), so d chi^2 = 2 Re sum conj(zb) dz and the gradient with
// respect to Re z + i Im z is 2 zb. For 2x2 Jones matrices <A, B> =
// sum conj(A_ij) B_ij.
//   Mb = -w (O - M)                       (chi^2 = sum w |O - M|^2, flagged -> 0)
//   Vb = G_p^H Mb G_q,   Gb_p += Mb G_q V^H,   Gb_q += Mb^H G_p V
//   Jb_p += f Vb J_q,    Jb_q += f Vb^H J_p,   d chi^2/d f = 2 Re <J_p, Vb J_q>
//   J = E X, X = K sqrt(B):  Eb = Jb X^H,  Xb = E^H Jb,
//   d chi^2/d phi = -2 Im(K conj(sum Xb conj(sqrtB))),  sqrtBb = conj(K) Xb
// and the chain rule on to the parameters (phase, beam weights, sqrt(B)).
//
// Launches: k_adj_vis (per time, channel tile; parallel over baselines)
// writes Vb to the workspace and reduces the gain adjoint in LDS;
// k_adj_src (per time, channel tile) recomputes the Jones tile of each
// source tile, accumulates Jb per (antenna, source) in registers by
// walking the antenna's row of the baseline map, then projects Jb onto the
// source parameters per (source, antenna, channel) item.

#include <algorithm>

#include "mbx_tile.hpp"

namespace mbx {

template <typename T>
__device__ __forceinline__ J4<T> jherm(const J4<T> &A) {
    J4<T> r;
    r.v[0] = {A.v[0].re, -A.v[0].im};
    r.v[1] = {A.v[2].re, -A.v[2].im};
    r.v[2] = {A.v[1].re, -A.v[1].im};
    r.v[3] = {A.v[3].re, -A.v[3].im};
    return r;
}

template <typename T>
__device__ __forceinline__ J4<T> jadd(const J4<T> &A, const J4<T> &B) {
    J4<T> r;
#pragma unroll
    for (int i = 0; i < 4; ++i) r.v[i] = {A.v[i].re + B.v[i].re, A.v[i].im + B.v[i].im};
    return r;
}

// Re <A, B>
template <typename T>
__device__ __forceinline__ T jdot_re(const J4<T> &A, const J4<T> &B) {
    T s = T(0);
#pragma unroll
    for (int i = 0; i < 4; ++i) s += A.v[i].re * B.v[i].re + A.v[i].im * B.v[i].im;
    return s;
}

template <typename T> struct AdjArgs {
    RimeArgs<T> R;
    const cx<T> *ekb;        // forward's pre-gain sums
    cx<T> *vbar;             // workspace: Vb per visibility
    const int32_t *bmap;     // (na x na) baseline map
    double *g_lm, *g_stokes, *g_alpha, *g_gauss, *g_sersic;
    double *g_G;             // (ntime, na, nchan, 4) complex128 as re, im pairs
    double gscale;           // upstream d L / d chi^2
    int ct2, nctile2;        // channel tiling of k_adj_vis
};

constexpr int kBlMask = (1 << 29) - 1;

// ---------------------------------------------------------------------
// Visibility stage: Mb, Vb and the gain adjoint.

template <typename T>
__global__ void __launch_bounds__(256)
k_adj_vis(AdjArgs<T> A) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double *Gacc = reinterpret_cast<double *>(smem_raw);   // [NA][ct2][8]
    const RimeArgs<T> &R = A.R;
    const mbx_dims &d = R.d;
    const int NA = d.na.local_size;
    const uint64_t NC = d.nchan.local_size, NBL = d.nbl.local_size;
    const int ct2 = A.ct2;
    const int tid = threadIdx.x;
    const int ctile = blockIdx.x % A.nctile2;
    const int t = blockIdx.x / A.nctile2;
    const int c0 = ctile * ct2;
    const int nb_ext = dext(d.nbl), nc_ext = dext(d.nchan);
    const bool active = t < int(dext(d.ntime));
    for (int i = tid; i < NA * ct2 * 8; i += 256) Gacc[i] = 0.0;
    __syncthreads();
    const bool weights = R.options & MBX_APPLY_WEIGHTS;
    if (active) {
        const int n = nb_ext * ct2;
        for (int i = tid; i < n; i += 256) {
            const int k = i % ct2, b = i / ct2;
            const int c = c0 + k;
            if (c >= nc_ext) continue;
            const int p = min(max(R.a1[uint64_t(t) * NBL + b], 0), NA - 1);
            const int q = min(max(R.a2[uint64_t(t) * NBL + b], 0), NA - 1);
            const uint64_t vi = (uint64_t(t) * NBL + b) * NC + c;
            const J4<T> V = jload(A.ekb + vi * 4);
            const J4<T> Gp = jload(R.G + ((uint64_t(t) * NA + p) * NC + c) * 4);
            const J4<T> Gq = jload(R.G + ((uint64_t(t) * NA + q) * NC + c) * 4);
            const J4<T> M = jmul_h(jmul(Gp, V), Gq);
            const J4<T> O = jload(R.obs + vi * 4);
            const uchar4 f4 = *reinterpret_cast<const uchar4 *>(R.flag + vi * 4);
            const uint8_t fa[4] = {f4.x, f4.y, f4.z, f4.w};
            J4<T> Mb;
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                T we = weights ? R.wv[vi * 4 + x] : T(1);
                if (fa[x] > 0) we = T(0);
                Mb.v[x] = {-we * (O.v[x].re - M.v[x].re), -we * (O.v[x].im - M.v[x].im)};
            }
            jstore(A.vbar + vi * 4, jmul(jmul(jherm(Gp), Mb), Gq));
            if (A.g_G) {
                const J4<T> gp = jmul_h(jmul(Mb, Gq), V);         // Mb G_q V^H
                const J4<T> gq = jmul(jmul(jherm(Mb), Gp), V);    // Mb^H G_p V
                double *ap = Gacc + (p * ct2 + k) * 8, *aq = Gacc + (q * ct2 + k) * 8;
#pragma unroll
                for (int x = 0; x < 4; ++x) {
                    atomicAdd(ap + 2 * x, double(gp.v[x].re));
                    atomicAdd(ap + 2 * x + 1, double(gp.v[x].im));
                    atomicAdd(aq + 2 * x, double(gq.v[x].re));
                    atomicAdd(aq + 2 * x + 1, double(gq.v[x].im));
                }
            }
        }
    }
    __syncthreads();
    if (active && A.g_G) {
        // torch convention: d/dRe + i d/dIm = 2 Gb
        const double s = 2.0 * A.gscale;
        for (int i = tid; i < NA * ct2 * 8; i += 256) {
            const int x = i % 8, k = (i / 8) % ct2, a = i / (8 * ct2);
            const int c = c0 + k;
            if (c < nc_ext) A.g_G[((uint64_t(t) * NA + a) * NC + c) * 8 + x] = s * Gacc[i];
        }
    }
}

// ---------------------------------------------------------------------
// Per-item backward of J = E K sqrt(B) onto (l, m, I, Q, U, V, alpha).

template <typename T>
__device__ __forceinline__ void antenna_backward(const RimeArgs<T> &A, const double *Luvw,
                                                 const double *Lpa, const double *Lch, int t,
                                                 int s, int a, int c, int k, const J4<T> &Jb,
                                                 double *g /* [7] */) {
    const mbx_dims &d = A.d;
    const uint64_t NT = d.ntime.local_size, NA = d.na.local_size, NC = d.nchan.local_size;
    const double l0 = A.lm[2 * uint64_t(s)], m0 = A.lm[2 * uint64_t(s) + 1];
    const double n = sqrt(1.0 - l0 * l0 - m0 * m0);
    const double nm1 = n - 1.0;
    // ---- beam value and its derivatives along the grid fractions
    const T *pet = A.pe + ((uint64_t(t) * NA + a) * NC + c) * 2;
    const T *asc = A.ascale + (uint64_t(a) * NC + c) * 2;
    const double ps = Lpa[2 * a], pc = Lpa[2 * a + 1];
    const double al = double(asc[0]), am = double(asc[1]);
    const double lr = (l0 * pc - m0 * ps + double(pet[0])) * al;
    const double mr = (l0 * ps + m0 * pc + double(pet[1])) * am;
    const double LW1 = double(d.beam_lw.local_size - 1), MH1 = double(d.beam_mh.local_size - 1);
    const double vl_raw = A.lscale * (lr - A.ll), vm_raw = A.mscale * (mr - A.lmin);
    const double vl = fmin(LW1, fmax(0.0, vl_raw)), vm = fmin(MH1, fmax(0.0, vm_raw));
    const bool okl = vl_raw >= 0.0 && vl_raw <= LW1, okm = vm_raw >= 0.0 && vm_raw <= MH1;
    const double gl0 = floor(vl), gm0 = floor(vm);
    const T ld = T(vl - gl0), md = T(vm - gm0), chd = T(Lch[CH * k + 2]);
    const uint32_t MH = d.beam_mh.local_size, NUD = d.beam_nud.local_size;
    const uint32_t il0 = uint32_t(gl0), il1 = uint32_t(fmin(gl0 + 1.0, LW1));
    const uint32_t im0 = uint32_t(gm0), im1 = uint32_t(fmin(gm0 + 1.0, MH1));
    const uint32_t ic0 = uint32_t(Lch[CH * k]), ic1 = uint32_t(Lch[CH * k + 1]);
    cx<T> S[4], dSl[4], dSm[4];
    T Am[4], dAl[4], dAm[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) {
        S[x] = dSl[x] = dSm[x] = {T(0), T(0)};
        Am[x] = dAl[x] = dAm[x] = T(0);
    }
#pragma unroll 1
    for (int tap = 0; tap < 8; ++tap) {
        const bool hl = tap & 1, hm = (tap >> 1) & 1, hc = tap >> 2;
        const uint32_t il = hl ? il1 : il0, im = hm ? im1 : im0, ic = hc ? ic1 : ic0;
        const T wl = hl ? ld : T(1) - ld, wm = hm ? md : T(1) - md, wc = hc ? chd : T(1) - chd;
        const T w = wl * wm * wc;
        const T wdl = (hl ? T(1) : T(-1)) * wm * wc;
        const T wdm = wl * (hm ? T(1) : T(-1)) * wc;
        const BeamTap<T> *e = A.beam + ((uint64_t(il * MH + im) * NUD) + ic) * 4;
#pragma unroll
        for (int x = 0; x < 4; ++x) {
            const BeamTap<T> v = e[x];
            S[x].re = fma(w, v.re, S[x].re);     S[x].im = fma(w, v.im, S[x].im);
            dSl[x].re = fma(wdl, v.re, dSl[x].re); dSl[x].im = fma(wdl, v.im, dSl[x].im);
            dSm[x].re = fma(wdm, v.re, dSm[x].re); dSm[x].im = fma(wdm, v.im, dSm[x].im);
            Am[x] = fma(w, v.amp, Am[x]);
            dAl[x] = fma(wdl, v.amp, dAl[x]);
            dAm[x] = fma(wdm, v.amp, dAm[x]);
        }
    }
    J4<T> E, dEl, dEm;
#pragma unroll
    for (int x = 0; x < 4; ++x) {
        const T n2 = S[x].re * S[x].re + S[x].im * S[x].im;
        if (n2 > T(0) && isfinite(n2)) {
            const T inv = T(1) / sqrt(n2);
            const T g = Am[x] * inv;
            E.v[x] = {S[x].re * g, S[x].im * g};
            // dE = dS g + S dA / |S| - S A Re(conj(S) dS) / |S|^3
            const T rl = (S[x].re * dSl[x].re + S[x].im * dSl[x].im) * inv * inv;
            const T rm = (S[x].re * dSm[x].re + S[x].im * dSm[x].im) * inv * inv;
            const T cl = dAl[x] * inv - g * rl, cm = dAm[x] * inv - g * rm;
            dEl.v[x] = {dSl[x].re * g + S[x].re * cl, dSl[x].im * g + S[x].im * cl};
            dEm.v[x] = {dSm[x].re * g + S[x].re * cm, dSm[x].im * g + S[x].im * cm};
        } else {
            E.v[x] = {S[x].re * Am[x], S[x].im * Am[x]};
            dEl.v[x] = {dSl[x].re * Am[x] + S[x].re * dAl[x], dSl[x].im * Am[x] + S[x].im * dAl[x]};
            dEm.v[x] = {dSm[x].re * Am[x] + S[x].re * dAm[x], dSm[x].im * Am[x] + S[x].im * dAm[x]};
        }
    }
    // ---- phase and sqrt(B)
    const double *xa = Luvw + 3 * a;
    const double f = Lch[CH * k + 3];
    const double phi = -A.two_pi_over_c * f * (xa[2] * nm1 + xa[1] * m0 + xa[0] * l0);
    T sn, cs;
    PhaseCS<T>::eval(phi, &sn, &cs);
    const cx<T> K = {cs, sn};
    const T *St = A.stokes + (uint64_t(s) * NT + t) * 4;
    const T I = St[0], Q = St[1], U = St[2], V = St[3];
    const double lnr = Lch[CH * k + 4];
    const T P = exp_t<T>(double(A.alpha[uint64_t(s) * NT + t]) * lnr);
    const T rP = sqrt(P);
    const T s0 = sqrt(I * I - Q * Q - U * U - V * V);
    const T t0 = sqrt(T(2) * I + T(2) * s0);
    J4<T> B0s;                            // sqrt(B0), B0 = unscaled brightness
    B0s.v[0] = {I + Q + s0, T(0)}; B0s.v[1] = {U, V}; B0s.v[2] = {U, -V}; B0s.v[3] = {I - Q + s0, T(0)};
    const T it0 = t0 != T(0) ? T(1) / t0 : T(0);
    if (t0 != T(0)) B0s = jscale(B0s, it0);
    const J4<T> sB = jscale(B0s, rP);      // sqrt(B) = sqrt(P) sqrt(B0)
    J4<T> X;
#pragma unroll
    for (int x = 0; x < 4; ++x) X.v[x] = cmul(K, sB.v[x]);
    // ---- adjoints
    const J4<T> Eb = jmul_h(Jb, X);                 // Jb X^H
    const J4<T> Xb = jmul(jherm(E), Jb);            // E^H Jb
    const T g_ld = T(2) * jdot_re(Eb, dEl);
    const T g_md = T(2) * jdot_re(Eb, dEm);
    // Kb = sum Xb conj(sqrtB); d chi^2/d phi = -2 Im(K conj(Kb))
    cx<T> Kb = {T(0), T(0)};
#pragma unroll
    for (int x = 0; x < 4; ++x) {
        Kb.re += Xb.v[x].re * sB.v[x].re + Xb.v[x].im * sB.v[x].im;
        Kb.im += Xb.v[x].im * sB.v[x].re - Xb.v[x].re * sB.v[x].im;
    }
    const T g_phi = T(-2) * (K.im * Kb.re - K.re * Kb.im);
    J4<T> Zb;                                       // sqrtB adjoint = conj(K) Xb
#pragma unroll
    for (int x = 0; x < 4; ++x) Zb.v[x] = cmul(cx<T>{K.re, -K.im}, Xb.v[x]);
    // alpha: d sqrtB / d alpha = lnr / 2 sqrtB
    g[6] = double(lnr) * double(jdot_re(Zb, sB));
    // stokes through sqrt(B0): d sqrtB0/dx = (dB0/dx + ds/dx 1)/t - sqrtB0 (dx_I + ds/dx)/t^2
    if (t0 != T(0) && s0 != T(0)) {
        const T ZI = Zb.v[0].re + Zb.v[3].re;
        const T ZB = jdot_re(Zb, B0s);
        const T rdB[4] = {Zb.v[0].re + Zb.v[3].re, Zb.v[0].re - Zb.v[3].re,
                          Zb.v[1].re + Zb.v[2].re, Zb.v[1].im - Zb.v[2].im};
        const T is0 = T(1) / s0;
        const T dsx[4] = {I * is0, -Q * is0, -U * is0, -V * is0};
#pragma unroll
        for (int x = 0; x < 4; ++x) {
            const T dt = (x == 0 ? T(1) : T(0)) + dsx[x];
            g[2 + x] = 2.0 * double(rP) * double((rdB[x] + dsx[x] * ZI) * it0 - ZB * dt * it0 * it0);
        }
    } else {
        g[2] = g[3] = g[4] = g[5] = 0.0;
    }
    // l, m: phase and beam-weight paths
    const double dphi_dl = -A.two_pi_over_c * f * (xa[0] - xa[2] * l0 / n);
    const double dphi_dm = -A.two_pi_over_c * f * (xa[1] - xa[2] * m0 / n);
    const double dvl_dl = okl ? A.lscale * al * pc : 0.0, dvl_dm = okl ? -A.lscale * al * ps : 0.0;
    const double dvm_dl = okm ? A.mscale * am * ps : 0.0, dvm_dm = okm ? A.mscale * am * pc : 0.0;
    g[0] = double(g_phi) * dphi_dl + double(g_ld) * dvl_dl + double(g_md) * dvm_dl;
    g[1] = double(g_phi) * dphi_dm + double(g_ld) * dvl_dm + double(g_md) * dvm_dm;
}

// Derivatives of the shape factor f of one pair with respect to the three
// shape parameters of its source (gaussian el, em, R or sersic e1, e2, R).
template <typename T>
__device__ __forceinline__ T shape_factor_grad(int ty, T u, T v, T p0, T p1, T p2, T kf,
                                               T *df /* [3] */) {
    if (ty == 1) {
        // x = (u em - v el) k R, y = (u el + v em) k, f = exp(-(x^2 + y^2))
        const T a = u * p1 - v * p0, bq = u * p0 + v * p1;
        const T x = a * kf * p2, y = bq * kf;
        const T f = exp(-(x * x + y * y));
        const T m2f = T(-2) * f;
        df[0] = m2f * (x * (-v * kf * p2) + y * (u * kf));
        df[1] = m2f * (x * (u * kf * p2) + y * (v * kf));
        df[2] = m2f * (x * a * kf);
        return f;
    }
    // sersic: p2 holds R' = R / (1 - e1^2 - e2^2); D = 1 + x^2 + y^2, f = D^-3/2
    const T den = T(1) - p0 * p0 - p1 * p1;
    const T x0 = u * (T(1) + p0) + v * p1, y0 = u * p1 + v * (T(1) - p0);
    const T x = x0 * kf * p2, y = y0 * kf * p2;
    const T D = T(1) + x * x + y * y;
    const T f = T(1) / (D * sqrt(D));
    const T c = T(-3) * f / D;                        // df/dx = c x, df/dy = c y
    const T dR1 = p2 * T(2) * p0 / den, dR2 = p2 * T(2) * p1 / den;   // dR'/de1, dR'/de2
    df[0] = c * (x * (u * kf * p2 + x0 * kf * dR1) + y * (-v * kf * p2 + y0 * kf * dR1));
    df[1] = c * (x * (v * kf * p2 + x0 * kf * dR2) + y * (u * kf * p2 + y0 * kf * dR2));
    df[2] = c * (x * x0 * kf + y * y0 * kf) / den;
    return f;
}

// ---------------------------------------------------------------------
// Source stage.

template <typename T, int STB>
__global__ void __launch_bounds__(512)
k_adj_src(AdjArgs<T> A) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const RimeArgs<T> &R = A.R;
    const mbx_dims &d = R.d;
    const int NA = d.na.local_size;
    const uint64_t NT = d.ntime.local_size, NC = d.nchan.local_size, NBL = d.nbl.local_size;
    const int ct = R.ct;
    const int tid = threadIdx.x, nth = blockDim.x;
    const int ctile = blockIdx.x % R.nctile;
    const int t = blockIdx.x / R.nctile;
    Tile<T> L(smem_raw, NA, ct, STB);
    double *sacc = L.part + (nth + 63) / 64;          // [STB][10]
    const bool active = t < int(dext(d.ntime));
    const int c0 = ctile * ct;
    const int nc_ext = dext(d.nchan);
    if (active) tile_setup(R, L, t, c0, tid, nth);
    for (int i = tid; i < STB * 10; i += nth) sacc[i] = 0.0;
    const int p = tid / ct, k = tid % ct;
    const int c = c0 + k;
    const bool mine = active && p < NA && c < nc_ext;
    const int np = dext(d.npsrc), ng = dext(d.ngsrc), ns_all = dext(d.nsrc);
    const uint64_t NG = d.ngsrc.local_size, NS = d.nssrc.local_size;
    __syncthreads();
    const T up = mine ? T(L.uvw[3 * p]) : T(0), vp = mine ? T(L.uvw[3 * p + 1]) : T(0);
    const T fc = mine ? T(L.ch[CH * k + 3]) : T(0);
    const T kf_g = T(R.gauss_scale) * fc, kf_s = T(R.two_pi_over_c) * fc;

    for (int s0 = 0; s0 < ns_all; s0 += STB) {
        const int nst = min(STB, ns_all - s0);
        if (active) tile_phase0(R, L, t, c0, s0, nst, tid, nth);
        __syncthreads();
        if (active) tile_phase1(R, L, t, c0, nst, tid, nth);
        __syncthreads();
        J4<T> Jb[STB];
        T sh[STB][3];
#pragma unroll
        for (int si = 0; si < STB; ++si) {
            Jb[si] = jzero<T>();
            sh[si][0] = sh[si][1] = sh[si][2] = T(0);
        }
        if (mine) {
            const int32_t *row = A.bmap + uint64_t(p) * NA;
            for (int q = 0; q < NA; ++q) {
                const int code = row[q];
                if (code < 0) continue;
                const int b = code & kBlMask;
                const bool rev = code & MBX_BL_REVERSED, aut = code & MBX_BL_AUTO;
                const J4<T> Wb = jload(A.vbar + ((uint64_t(t) * NBL + b) * NC + c) * 4);
                const J4<T> W = rev ? jherm(Wb) : (aut ? jadd(Wb, jherm(Wb)) : Wb);
                const T u = T(L.uvw[3 * q]) - up, v = T(L.uvw[3 * q + 1]) - vp;
#pragma unroll
                for (int si = 0; si < STB; ++si) {
                    if (si >= nst) break;
                    const J4<T> Y = jmul(W, L.J[(si * NA + q) * ct + k]);
                    const int ty = int(L.shape[4 * si + 3]);
                    if (ty == 0) {
#pragma unroll
                        for (int x = 0; x < 4; ++x) {
                            Jb[si].v[x].re += Y.v[x].re;
                            Jb[si].v[x].im += Y.v[x].im;
                        }
                    } else {
                        T df[3];
                        const T f = shape_factor_grad<T>(ty, u, v, L.shape[4 * si], L.shape[4 * si + 1],
                                                         L.shape[4 * si + 2], ty == 1 ? kf_g : kf_s, df);
#pragma unroll
                        for (int x = 0; x < 4; ++x) {
                            Jb[si].v[x].re = fma(f, Y.v[x].re, Jb[si].v[x].re);
                            Jb[si].v[x].im = fma(f, Y.v[x].im, Jb[si].v[x].im);
                        }
                        if (!rev) {
                            // d chi^2/d f = 2 Re <J_p, Vb J_q> (auto pairs: W = Vb + Vb^H)
                            const T fb = (aut ? T(1) : T(2)) *
                                         jdot_re(L.J[(si * NA + p) * ct + k], Y);
                            sh[si][0] = fma(fb, df[0], sh[si][0]);
                            sh[si][1] = fma(fb, df[1], sh[si][1]);
                            sh[si][2] = fma(fb, df[2], sh[si][2]);
                        }
                    }
                }
            }
        }
        __syncthreads();                               // every Jones read done
        if (mine) {
#pragma unroll
            for (int si = 0; si < STB; ++si) {
                if (si >= nst) break;
                L.J[(si * NA + p) * ct + k] = Jb[si];
                if (int(L.shape[4 * si + 3]) != 0) {
                    atomicAdd(sacc + si * 10 + 7, double(sh[si][0]));
                    atomicAdd(sacc + si * 10 + 8, double(sh[si][1]));
                    atomicAdd(sacc + si * 10 + 9, double(sh[si][2]));
                }
            }
        }
        __syncthreads();
        if (active) {
            const int n3 = nst * NA * ct;
            for (int i = tid; i < n3; i += nth) {
                const int kk = i % ct, a = (i / ct) % NA, si = i / (ct * NA);
                const int cc = c0 + kk;
                if (cc >= nc_ext) continue;
                double g[7];
                antenna_backward<T>(R, L.uvw, L.pa, L.ch, t, s0 + si, a, cc, kk, L.J[i], g);
#pragma unroll
                for (int x = 0; x < 7; ++x) atomicAdd(sacc + si * 10 + x, g[x]);
            }
        }
        __syncthreads();
        if (active) {
            for (int si = tid; si < nst; si += nth) {
                const int s = s0 + si;
                const double *v = sacc + si * 10;
                const double gs = A.gscale;
                if (A.g_lm) {
                    atomicAdd(A.g_lm + 2 * uint64_t(s), gs * v[0]);
                    atomicAdd(A.g_lm + 2 * uint64_t(s) + 1, gs * v[1]);
                }
                if (A.g_stokes)
                    for (int x = 0; x < 4; ++x)
                        atomicAdd(A.g_stokes + (uint64_t(s) * NT + t) * 4 + x, gs * v[2 + x]);
                if (A.g_alpha) atomicAdd(A.g_alpha + uint64_t(s) * NT + t, gs * v[6]);
                if (s >= np && s < np + ng && A.g_gauss) {
                    const int g = s - np;
                    for (int x = 0; x < 3; ++x) atomicAdd(A.g_gauss + g + x * NG, gs * v[7 + x]);
                } else if (s >= np + ng && A.g_sersic) {
                    const int g = s - np - ng;
                    for (int x = 0; x < 3; ++x) atomicAdd(A.g_sersic + g + x * NS, gs * v[7 + x]);
                }
            }
        }
        __syncthreads();
        for (int i = tid; i < STB * 10; i += nth) sacc[i] = 0.0;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_fill_zero(double *p, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

static void fill_zero(double *p, uint64_t n, hipStream_t st) {
    if (!p || n == 0) return;
    unsigned g = (unsigned)std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_zero, dim3(g), dim3(256), 0, st, p, n);
}

int validate_inputs(const mbx_dims *d, const mbx_inputs *in, const mbx_beam_extents *b,
                    const mbx_consts *k, int options, const char *who);

template <typename T>
static size_t adjoint_ws_bytes(const mbx_dims &d) {
    size_t beam = size_t(d.beam_lw.local_size) * d.beam_mh.local_size * d.beam_nud.local_size * 4 *
                  sizeof(BeamTap<T>);
    beam = (beam + 255) & ~size_t(255);
    size_t vbar = size_t(d.ntime.local_size) * d.nbl.local_size * d.nchan.local_size * 4 * sizeof(cx<T>);
    return beam + vbar;
}

template <typename T> constexpr int adj_stb() { return sizeof(T) == 4 ? 8 : 4; }

template <typename T>
static int adjoint(const mbx_dims *d, const mbx_inputs *in, const void *ekb, const int32_t *bmap,
                   const mbx_grads *gr, const mbx_beam_extents *b, const mbx_consts *k, int options,
                   double gscale, void *ws, size_t ws_bytes, hipStream_t stream) {
    int r = validate_inputs(d, in, b, k, options, "adjoint");
    if (r) return r;
    if (!ekb || !bmap || !gr) return set_error(MBX_E_NULL, "adjoint: NULL ekb_vis, baseline map or grads");
    if (d->nsrc.lower_extent != 0 || d->nsrc.upper_extent != d->nsrc.global_size)
        return set_error(MBX_E_EXTENT, "adjoint: the source extents must cover every source");
    size_t need = adjoint_ws_bytes<T>(*d);
    if (!ws || ws_bytes < need)
        return set_error(MBX_E_WORKSPACE, "adjoint: workspace needs %zu bytes (got %zu)", need, ws_bytes);
    const int na = d->na.local_size;
    const uint64_t NT = d->ntime.local_size, NC = d->nchan.local_size;
    const uint64_t ns = d->nsrc.local_size;
    fill_zero(gr->lm, ns * 2, stream);
    fill_zero(gr->stokes, ns * NT * 4, stream);
    fill_zero(gr->alpha, ns * NT, stream);
    fill_zero(gr->gauss_shape, uint64_t(d->ngsrc.local_size) * 3, stream);
    fill_zero(gr->sersic_shape, uint64_t(d->nssrc.local_size) * 3, stream);
    fill_zero(static_cast<double *>(gr->G_term), NT * na * NC * 8, stream);
    if (dext(d->ntime) == 0 || dext(d->nbl) == 0 || dext(d->nchan) == 0)
        return hip_status(hipGetLastError(), "adjoint (empty extents)");

    const uint64_t ncube = uint64_t(d->beam_lw.local_size) * d->beam_mh.local_size *
                           d->beam_nud.local_size * 4;
    BeamTap<T> *beam = static_cast<BeamTap<T> *>(ws);
    size_t beam_bytes = (ncube * sizeof(BeamTap<T>) + 255) & ~size_t(255);
    cx<T> *vbar = reinterpret_cast<cx<T> *>(static_cast<char *>(ws) + beam_bytes);
    launch_beam_prepare<T>(static_cast<const cx<T> *>(in->E_beam), beam, ncube, stream);

    AdjArgs<T> A;
    RimeArgs<T> &R = A.R;
    fill_args<T>(R, d, in, b, k, options);
    R.beam = beam;
    A.ekb = static_cast<const cx<T> *>(ekb);
    A.vbar = vbar;
    A.bmap = bmap;
    A.g_lm = gr->lm; A.g_stokes = gr->stokes; A.g_alpha = gr->alpha;
    A.g_gauss = gr->gauss_shape; A.g_sersic = gr->sersic_shape;
    A.g_G = static_cast<double *>(gr->G_term);
    // chi^2 = sum w |R|^2 / sigma^2 without weights: fold 1/sigma^2 into the scale
    A.gscale = gscale * ((options & MBX_APPLY_WEIGHTS) ? 1.0 : 1.0 / k->sigma_sqrd);

    // visibility stage
    A.ct2 = std::max(1, std::min<int>(8, d->nchan.local_size));
    while (A.ct2 > 1 && size_t(na) * A.ct2 * 8 * sizeof(double) > 48 * 1024) --A.ct2;
    A.nctile2 = (d->nchan.local_size + A.ct2 - 1) / A.ct2;
    size_t lds1 = size_t(na) * A.ct2 * 8 * sizeof(double);
    hipLaunchKernelGGL(k_adj_vis<T>, dim3(unsigned(NT * A.nctile2)), dim3(256), lds1, stream, A);
    r = hip_status(hipGetLastError(), "adjoint vis launch");
    if (r) return r;

    // source stage: one thread per (antenna, channel) of the tile
    constexpr int STB = adj_stb<T>();
    int ct = std::max(1, std::min<int>(512 / na, d->nchan.local_size));
    auto lds_of = [&](int c) {
        int th = ((na * c + 63) / 64) * 64;
        return tile_lds_bytes<T>(na, c, STB, th / 64) + sizeof(double) * STB * 10;
    };
    while (ct > 1 && lds_of(ct) > 60 * 1024) --ct;
    if (na > 512) return set_error(MBX_E_UNSUPPORTED, "adjoint: more than 512 antennas");
    R.ct = ct;
    R.st = STB;
    R.nctile = (d->nchan.local_size + ct - 1) / ct;
    R.nbchunk = 1;
    const int threads = ((na * ct + 63) / 64) * 64;
    hipLaunchKernelGGL((k_adj_src<T, STB>), dim3(unsigned(NT * R.nctile)), dim3(threads),
                       lds_of(ct), stream, A);
    return hip_status(hipGetLastError(), "adjoint source launch");
}

}  // namespace mbx

using namespace mbx;

extern "C" {

size_t mbx_adjoint_workspace_bytes(const mbx_dims *d, int is_double) {
    if (!d) return 0;
    return is_double ? adjoint_ws_bytes<double>(*d) : adjoint_ws_bytes<float>(*d);
}

int mbx_adjoint_f32(const mbx_dims *d, const mbx_inputs *in, const void *ekb_vis,
                    const int32_t *bmap, const mbx_grads *grads, const mbx_beam_extents *beam,
                    const mbx_consts *k, int options, double grad_scale, void *ws,
                    size_t ws_bytes, void *s) {
    return adjoint<float>(d, in, ekb_vis, bmap, grads, beam, k, options, grad_scale, ws, ws_bytes,
                          (hipStream_t)s);
}
int mbx_adjoint_f64(const mbx_dims *d, const mbx_inputs *in, const void *ekb_vis,
                    const int32_t *bmap, const mbx_grads *grads, const mbx_beam_extents *beam,
                    const mbx_consts *k, int options, double grad_scale, void *ws,
                    size_t ws_bytes, void *s) {
    return adjoint<double>(d, in, ekb_vis, bmap, grads, beam, k, options, grad_scale, ws, ws_bytes,
                           (hipStream_t)s);
}

}  // extern "C"
