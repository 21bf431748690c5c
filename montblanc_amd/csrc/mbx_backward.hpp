// Backward (adjoint) helpers shared by the single-model adjoint and the
// batched kernels: Jones-matrix adjoint algebra, the per-(source, antenna,
// channel) backward of J = E K sqrt(B) onto (l, m, I, Q, U, V, alpha), and
// the shape-factor derivatives. Conventions: zb = d chi^2 / d z* for a
// complex z (Wirtinger), <A, B> = sum conj(A_ij) B_ij.
#pragma once

#include "mbx_tile.hpp"

namespace mbx {

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

// enter_sample (mbx_tile.hpp) for the adjoint: also the slot's pre-gain
// sums and Vb, and the sample's rows / columns of the gradient arrays
// (concatenated like the source arrays).
template <typename T>
__device__ __forceinline__ SampleAt enter_sample(AdjArgs<T> &A) {
    const SampleAt at = enter_sample(A.R);
    if (A.R.span.wgs) {
        const uint64_t off = uint64_t(at.j) * A.R.span.vstride, NT = A.R.d.ntime.local_size;
        A.ekb += off;
        A.vbar += off;
        if (A.g_lm) A.g_lm += 2 * uint64_t(at.src0);
        if (A.g_stokes) A.g_stokes += uint64_t(at.src0) * NT * 4;
        if (A.g_alpha) A.g_alpha += uint64_t(at.src0) * NT;
        if (A.g_gauss) A.g_gauss += at.goff;
        if (A.g_sersic) A.g_sersic += at.soff;
    }
    return at;
}


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

// Re <A, B> (one fused multiply-add chain)
template <typename T>
__device__ __forceinline__ T jdot_re(const J4<T> &A, const J4<T> &B) {
    T s = A.v[0].re * B.v[0].re;
    s = fma(A.v[0].im, B.v[0].im, s);
#pragma unroll
    for (int i = 1; i < 4; ++i) {
        s = fma(A.v[i].re, B.v[i].re, s);
        s = fma(A.v[i].im, B.v[i].im, s);
    }
    return s;
}

// ---------------------------------------------------------------------
// Per-item backward of J = E K sqrt(B) onto (l, m, I, Q, U, V, alpha).

// What the backward of one (source, antenna, channel) item keeps from its
// forward: the beam Jones and its derivatives along l and m, the phase
// term, the brightness terms and whether the beam position was inside the
// grid. What is cheap to recompute from the tile's LDS records (the
// chain-rule factors, sqrt(B0)) is left out, so that the state held across
// a source tile's pair loop stays small.
template <typename T> struct ItemState {
    J4<T> E, dEl, dEm;
    cx<T> K;
    T rP, s0, it0, I, Q, U, V;
    bool okl, okm;
};

// Brightness terms of one source at one (time, channel), shared by its
// items: sqrt(P) (P = (f/f_ref)^alpha), s0 = sqrt(det B0), 1 / t0 with
// t0 = sqrt(tr B0 + 2 s0) (0 where t0 = 0), 1 / s0 (0 where s0 = 0).
template <typename T> struct SrcBright { T rP, s0, it0, is0; };

template <typename T>
__device__ __forceinline__ SrcBright<T> src_bright(T I, T Q, T U, T V, T P) {
    SrcBright<T> b;
    b.rP = sqrt(P);
    b.s0 = sqrt(I * I - Q * Q - U * U - V * V);
    const T t0 = sqrt(T(2) * I + T(2) * b.s0);
    b.it0 = t0 != T(0) ? T(1) / t0 : T(0);
    b.is0 = b.s0 != T(0) ? T(1) / b.s0 : T(0);
    return b;
}

// Chain-rule factors of one item from (l, m) to its phase and its beam
// grid position, and ln(f / f_ref) (inputs as item_forward).
struct ItemChain {
    double lnr, dphi_dl, dphi_dm, dvl_dl, dvl_dm, dvm_dl, dvm_dm;
};

// rn = 1 / n, n = sqrt(1 - l0^2 - m0^2): a per-source value the caller
// keeps, so that no item evaluates a square root or a division for it.
template <typename T>
__device__ __forceinline__ ItemChain item_chain(const RimeArgs<T> &A, const double *xa, double ps,
                                                double pc, double al, double am, double f, double lnr,
                                                double l0, double m0, double rn, bool okl, bool okm) {
    ItemChain ch;
    ch.lnr = lnr;
    ch.dphi_dl = -A.two_pi_over_c * f * (xa[0] - xa[2] * (l0 * rn));
    ch.dphi_dm = -A.two_pi_over_c * f * (xa[1] - xa[2] * (m0 * rn));
    ch.dvl_dl = okl ? A.lscale * al * pc : 0.0; ch.dvl_dm = okl ? -A.lscale * al * ps : 0.0;
    ch.dvm_dl = okm ? A.mscale * am * ps : 0.0; ch.dvm_dm = okm ? A.mscale * am * pc : 0.0;
    return ch;
}

// Forward of one item with the derivative information its backward needs.
// Inputs: the source's l, m; its Stokes I, Q, U, V and brightness terms
// (src_bright) at this (time, channel); the antenna's uvw (xa),
// parallactic sin/cos (ps, pc), pointing error (pel, pem) and scaling
// (al, am); the channel c (beam plane), its frequency f and
// lnr = log(f/f_ref); nm1 = n - 1 of the source (n = sqrt(1 - l0^2 - m0^2),
// computed once per source by the caller). Returns J = E K sqrt(B).
template <typename T>
__device__ __forceinline__ J4<T> item_forward(const RimeArgs<T> &A, const double *xa, double ps,
                                              double pc, double pel, double pem, double al,
                                              double am, int c, double f, double lnr, double l0,
                                              double m0, double nm1, T I, T Q, T U, T V,
                                              const SrcBright<T> &sb, ItemState<T> &st) {
    const mbx_dims &d = A.d;
    // ---- beam value and its derivatives along the grid fractions
    const double lr = (l0 * pc - m0 * ps + pel) * al;
    const double mr = (l0 * ps + m0 * pc + pem) * am;
    const double LW1 = double(d.beam_lw.local_size - 1), MH1 = double(d.beam_mh.local_size - 1);
    const double vl_raw = A.lscale * (lr - A.ll), vm_raw = A.mscale * (mr - A.lmin);
    const double vl = fmin(LW1, fmax(0.0, vl_raw)), vm = fmin(MH1, fmax(0.0, vm_raw));
    const bool okl = vl_raw >= 0.0 && vl_raw <= LW1, okm = vm_raw >= 0.0 && vm_raw <= MH1;
    const double gl0 = floor(vl), gm0 = floor(vm);
    const T ld = T(vl - gl0), md = T(vm - gm0);
    const uint32_t MH = d.beam_mh.local_size;
    // tap byte offsets in the channel's plane, as in beam_jones_at
    const uint32_t cell = uint32_t(sizeof(BeamCell<T>));
    const uint32_t o00 = uint32_t(__umul24(uint32_t(__umul24(uint32_t(gl0), MH)) + uint32_t(gm0), cell));
    const uint32_t dl = gl0 < LW1 ? uint32_t(__umul24(MH, cell)) : 0u, dm = gm0 < MH1 ? cell : 0u;
    const char *plane = reinterpret_cast<const char *>(A.beam + uint64_t(c) * d.beam_lw.local_size * MH);
    // the four taps: cells (g0, g0), (g0 + 1, g0), (g0, g0 + 1), (g0 + 1, g0 + 1)
    // in (l, m); each of the 12 reals (re, im, |.| of 4 correlations) and its
    // derivatives along the grid fractions from the differences
    // a = v10 - v00, c = v01 - v00, e = v11 - v01 - a:
    //   v = v00 + ld a + md (c + ld e),  dv/dld = a + md e,  dv/dmd = c + ld e
    const BeamCell<T> e00 = *reinterpret_cast<const BeamCell<T> *>(plane + o00);
    const BeamCell<T> e10 = *reinterpret_cast<const BeamCell<T> *>(plane + o00 + dl);
    const BeamCell<T> e01 = *reinterpret_cast<const BeamCell<T> *>(plane + o00 + dm);
    const BeamCell<T> e11 = *reinterpret_cast<const BeamCell<T> *>(plane + o00 + dl + dm);
    // ---- phase, evaluated while the four cells are in flight
    const double phi = -A.two_pi_over_c * f * (xa[2] * nm1 + xa[1] * m0 + xa[0] * l0);
    T sn, cs;
    PhaseCS<T>::eval(phi, &sn, &cs);
    const cx<T> K = {cs, sn};
    T Bv[12], dBl[12], dBm[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) {
        const T a = e10.v[i] - e00.v[i], cc = e01.v[i] - e00.v[i];
        const T ee = (e11.v[i] - e01.v[i]) - a;
        const T cm = fma(ld, ee, cc);
        Bv[i] = fma(ld, a, fma(md, cm, e00.v[i]));
        dBl[i] = fma(md, ee, a);
        dBm[i] = cm;
    }
    cx<T> S[4], dSl[4], dSm[4];
    T Am[4], dAl[4], dAm[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) {
        S[x] = {Bv[3 * x], Bv[3 * x + 1]};
        dSl[x] = {dBl[3 * x], dBl[3 * x + 1]};
        dSm[x] = {dBm[3 * x], dBm[3 * x + 1]};
        Am[x] = Bv[3 * x + 2];
        dAl[x] = dBl[3 * x + 2];
        dAm[x] = dBm[3 * x + 2];
    }
    J4<T> E, dEl, dEm;
#pragma unroll
    for (int x = 0; x < 4; ++x) {
        const T n2 = S[x].re * S[x].re + S[x].im * S[x].im;
        if (n2 > T(0) && isfinite(n2)) {
            const T inv = rsqrt_acc(n2);
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
    // ---- sqrt(B)
    const T rP = sb.rP, s0 = sb.s0, it0 = sb.it0;
    J4<T> B0s;                            // sqrt(B0), B0 = unscaled brightness
    B0s.v[0] = {I + Q + s0, T(0)}; B0s.v[1] = {U, V}; B0s.v[2] = {U, -V}; B0s.v[3] = {I - Q + s0, T(0)};
    if (it0 != T(0)) B0s = jscale(B0s, it0);
    const J4<T> sB = jscale(B0s, rP);      // sqrt(B) = sqrt(P) sqrt(B0)
    J4<T> X;
#pragma unroll
    for (int x = 0; x < 4; ++x) X.v[x] = cmul(K, sB.v[x]);
    st.E = E; st.dEl = dEl; st.dEm = dEm; st.K = K;
    st.rP = rP; st.s0 = s0; st.it0 = it0; st.I = I; st.Q = Q; st.U = U; st.V = V;
    st.okl = okl; st.okm = okm;
    return jmul(E, X);
}

// Backward of one item: d chi^2 / d(l, m, I, Q, U, V, alpha) into g from
// the Jones adjoint Jb, the item's forward state, its chain-rule factors
// (item_chain) and the source's 1 / s0 (src_bright).
template <typename T>
__device__ __forceinline__ void item_backward(const ItemState<T> &st, const ItemChain &ch,
                                              const J4<T> &Jb, T is0, double *g /* [7] */) {
    const J4<T> &E = st.E, &dEl = st.dEl, &dEm = st.dEm;
    const cx<T> K = st.K;
    const T rP = st.rP, s0 = st.s0, it0 = st.it0, I = st.I, Q = st.Q, U = st.U, V = st.V;
    const T t0 = it0 != T(0) ? T(1) : T(0);   // nonzero marker: it0 = 1/t0 or 0
    const double lnr = ch.lnr;
    J4<T> B0s;                                // sqrt(B0), as item_forward
    B0s.v[0] = {I + Q + s0, T(0)}; B0s.v[1] = {U, V}; B0s.v[2] = {U, -V}; B0s.v[3] = {I - Q + s0, T(0)};
    if (t0 != T(0)) B0s = jscale(B0s, it0);
    const J4<T> sB = jscale(B0s, rP);
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
    g[0] = double(g_phi) * ch.dphi_dl + double(g_ld) * ch.dvl_dl + double(g_md) * ch.dvm_dl;
    g[1] = double(g_phi) * ch.dphi_dm + double(g_ld) * ch.dvl_dm + double(g_md) * ch.dvm_dm;
}

// Forward and backward of one item in one call (inputs as item_forward).
template <typename T>
__device__ __forceinline__ void antenna_backward_at(const RimeArgs<T> &A, const double *xa,
                                                    double ps, double pc, double pel, double pem,
                                                    double al, double am, int c, double f,
                                                    double lnr, double l0, double m0, T I, T Q,
                                                    T U, T V, T P, const J4<T> &Jb,
                                                    double *g /* [7] */) {
    ItemState<T> st;
    const double n = sqrt(1.0 - l0 * l0 - m0 * m0);
    const SrcBright<T> sb = src_bright<T>(I, Q, U, V, P);
    (void)item_forward<T>(A, xa, ps, pc, pel, pem, al, am, c, f, lnr, l0, m0, n - 1.0, I, Q, U, V, sb, st);
    item_backward<T>(st, item_chain<T>(A, xa, ps, pc, al, am, f, lnr, l0, m0, 1.0 / n, st.okl, st.okm), Jb,
                     sb.is0, g);
}

// The same item reading its inputs from global memory and the tile's
// per-antenna (Luvw, Lpa) and channel (Lch, record k) LDS arrays.
template <typename T>
__device__ __forceinline__ void antenna_backward(const RimeArgs<T> &A, const double *Luvw,
                                                 const double *Lpa, const double *Lch, int t,
                                                 int s, int a, int c, int k, const J4<T> &Jb,
                                                 double *g /* [7] */) {
    const mbx_dims &d = A.d;
    const uint64_t NT = d.ntime.local_size, NA = d.na.local_size, NC = d.nchan.local_size;
    const T *pet = A.pe + ((uint64_t(t) * NA + a) * NC + c) * 2;
    const T *asc = A.ascale + (uint64_t(a) * NC + c) * 2;
    const T *St = A.stokes + (uint64_t(s) * NT + t) * 4;
    const double lnr = Lch[CH * k + 4];
    const T P = exp_t<T>(double(A.alpha[uint64_t(s) * NT + t]) * lnr);
    antenna_backward_at<T>(A, Luvw + 3 * a, Lpa[2 * a], Lpa[2 * a + 1], double(pet[0]), double(pet[1]),
                           double(asc[0]), double(asc[1]), c, Lch[CH * k + 3], lnr,
                           A.lm[2 * uint64_t(s)], A.lm[2 * uint64_t(s) + 1], St[0], St[1], St[2], St[3],
                           P, Jb, g);
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

}  // namespace mbx
