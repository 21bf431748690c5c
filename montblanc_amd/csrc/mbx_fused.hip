// Fused forward RIME for gfx950.
//
// Replaces the four-node v4 pipeline (B_sqrt -> E_beam -> EKB_sqrt ->
// sum_coherencies -> chi^2 sum) with one kernel that never writes the
// per-source scratch arrays (jones is 42 GB in float64 at 64 antennas,
// 100 times, 64 channels and 1600 sources).
//
// Work decomposition
//   workgroup = (time t, channel tile of `ct` channels, baseline chunk)
//   per source tile of `st` sources:
//     phase 0  source parameters and sqrt(B) per (source, channel) -> LDS
//     phase 1  per-antenna Jones J = E.K.sqrt(B) per (source, antenna,
//              channel) -> LDS   (beam taps from a prepared cube with the
//              tap amplitudes precomputed, phase argument in float64)
//     phase 2  every thread owns NI (baseline, channel) items and keeps
//              their 2x2 coherency sums in registers:
//              V_pq += f_s(pq) J_ps J_qs^H  (f = 1, gaussian or sersic)
//   epilogue   M = G_p V G_q^H, flags, model/residual output, chi^2 terms,
//              float64 workgroup partial of chi^2.
// A final one-block kernel sums the partials in a fixed order.
//
// Source-batch semantics of the reference (RimeSumCoherencies.py:183-190,
// 307-316) are kept: with nsrc.lower_extent > 0 the running sum is read
// from ekb_vis; with nsrc.upper_extent < nsrc.global_size it is written
// to ekb_vis and the epilogue is skipped.

#include "mbx_common.hpp"

namespace mbx {

template <typename T> struct BeamTap { T re, im, amp, pad; };

void launch_chi2_final(const double *partial, unsigned n, double scale, double *X2, hipStream_t st);

template <typename T>
__global__ void __launch_bounds__(256)
k_beam_prepare(const cx<T> *__restrict__ E, BeamTap<T> *__restrict__ out, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        cx<T> e = E[i];
        out[i] = {e.re, e.im, sqrt(e.re * e.re + e.im * e.im), T(0)};
    }
}

template <typename T> struct FwdArgs {
    mbx_dims d;
    const T *uvw, *lm, *stokes, *alpha, *freq, *ref_freq, *pa, *pe, *ascale, *gauss, *sersic, *wv;
    const BeamTap<T> *beam;
    const int32_t *a1, *a2;
    const uint8_t *flag;
    const cx<T> *obs, *G;
    cx<T> *model;
    T *chi2;
    cx<T> *ekb;
    double *partial;
    double ll, lmin, lscale, mscale, lfreq, fscale;
    double gauss_scale, two_pi_over_c;
    int options;
    int ct, st, nbchunk, bchunk, nctile;
};

struct FwdPlan {
    int ct, st, ni, nbchunk, bchunk, nctile;
    uint64_t nwg;
    size_t lds_bytes;
};

constexpr int kThreads = 256;

template <typename T>
static size_t fwd_lds_bytes(int na, int ct, int st) {
    return sizeof(double) * (3 * na + 2 * na)      // uvw, parallactic sin/cos
         + sizeof(double) * 4 * ct                  // gc0, gc1, chd, freq
         + sizeof(double) * 3 * st + sizeof(T) * 4 * st  // l, m, n-1; shape params + type
         + sizeof(J4<T>) * st * ct                  // sqrt(B)
         + sizeof(J4<T>) * st * na * ct             // per-antenna Jones tile
         + sizeof(double) * (kThreads / 64)         // chi^2 reduction
         + 16;                                      // alignment of the Jones tiles
}

template <typename T>
static FwdPlan plan_forward(const mbx_dims &d) {
    FwdPlan p;
    const int na = d.na.local_size;
    const int nb = d.nbl.local_size;
    p.ct = 1;
    int need = (nb * p.ct + kThreads - 1) / kThreads;
    p.ni = need <= 1 ? 1 : need <= 2 ? 2 : need <= 4 ? 4 : 8;
    int per_chunk = p.ni * kThreads / p.ct;
    p.nbchunk = (nb + per_chunk - 1) / per_chunk;
    p.bchunk = (nb + p.nbchunk - 1) / p.nbchunk;
    p.nctile = (d.nchan.local_size + p.ct - 1) / p.ct;
    const size_t budget = 60 * 1024;
    int st = 32;
    while (st > 1 && fwd_lds_bytes<T>(na, p.ct, st) > budget) st >>= 1;
    p.st = st;
    p.lds_bytes = fwd_lds_bytes<T>(na, p.ct, st);
    p.nwg = uint64_t(d.ntime.local_size) * p.nctile * p.nbchunk;
    return p;
}

// Per-antenna Jones J = E . (K * sqrt(B)) of one (source, time, antenna,
// channel); E from the prepared cube (RimeEBeam.py:233-343 semantics).
template <typename T>
__device__ __forceinline__ J4<T> antenna_jones(const FwdArgs<T> &A, int t, int a, int c,
                                               double l0, double m0, double nm1,
                                               const J4<T> &sqB, double ps, double pc,
                                               const double *uvw_a, double gc0, double gc1,
                                               T chd, double f) {
    const mbx_dims &d = A.d;
    const uint64_t NA = d.na.local_size, NC = d.nchan.local_size;
    const T *pet = A.pe + ((uint64_t(t) * NA + a) * NC + c) * 2;
    const T *asc = A.ascale + (uint64_t(a) * NC + c) * 2;
    double l = (l0 * pc - m0 * ps + double(pet[0])) * double(asc[0]);
    double m = (l0 * ps + m0 * pc + double(pet[1])) * double(asc[1]);
    const double LW1 = double(d.beam_lw.local_size - 1), MH1 = double(d.beam_mh.local_size - 1);
    double vl = fmin(LW1, fmax(0.0, A.lscale * (l - A.ll)));
    double gl0 = floor(vl), gl1 = fmin(gl0 + 1.0, LW1);
    T ld = T(vl - gl0);
    double vm = fmin(MH1, fmax(0.0, A.mscale * (m - A.lmin)));
    double gm0 = floor(vm), gm1 = fmin(gm0 + 1.0, MH1);
    T md = T(vm - gm0);
    const uint64_t MH = d.beam_mh.local_size, NUD = d.beam_nud.local_size;
    const uint64_t il[2] = {uint64_t(gl0), uint64_t(gl1)};
    const uint64_t im[2] = {uint64_t(gm0), uint64_t(gm1)};
    const uint64_t ic[2] = {uint64_t(gc0), uint64_t(gc1)};
    const T wl[2] = {T(1) - ld, ld}, wm[2] = {T(1) - md, md}, wc[2] = {T(1) - chd, chd};
    cx<T> acc[4];
    T amp[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) { acc[p] = {T(0), T(0)}; amp[p] = T(0); }
#pragma unroll
    for (int kl = 0; kl < 2; ++kl)
#pragma unroll
        for (int km = 0; km < 2; ++km) {
            T wlm = wl[kl] * wm[km];
            const BeamTap<T> *row = A.beam + ((il[kl] * MH + im[km]) * NUD) * 4;
#pragma unroll
            for (int kc = 0; kc < 2; ++kc) {
                T w = wlm * wc[kc];
                const BeamTap<T> *tap = row + ic[kc] * 4;
#pragma unroll
                for (int p = 0; p < 4; ++p) {
                    BeamTap<T> e = tap[p];
                    acc[p].re = fma(w, e.re, acc[p].re);
                    acc[p].im = fma(w, e.im, acc[p].im);
                    amp[p] = fma(w, e.amp, amp[p]);
                }
            }
        }
    J4<T> E;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        T n2 = acc[p].re * acc[p].re + acc[p].im * acc[p].im;
        T g = (n2 > T(0) && isfinite(n2)) ? amp[p] / sqrt(n2) : amp[p];
        E.v[p] = {acc[p].re * g, acc[p].im * g};
    }
    double sn, cs;
    phase_sincos(uvw_a[0], uvw_a[1], uvw_a[2], l0, m0, nm1, f, A.two_pi_over_c, &sn, &cs);
    cx<T> K = {T(cs), T(sn)};
    J4<T> KB;
#pragma unroll
    for (int p = 0; p < 4; ++p) KB.v[p] = cmul(K, sqB.v[p]);
    return jmul(E, KB);
}

template <typename T>
__device__ __forceinline__ T gauss_f(T u, T v, T el, T em, T R, T k) {
    T x = (u * em - v * el) * k * R;
    T y = (u * el + v * em) * k;
    return exp(-(x * x + y * y));
}

template <typename T>
__device__ __forceinline__ T sersic_f(T u, T v, T e1, T e2, T Rp, T k) {
    // Rp = R / (1 - e1^2 - e2^2) precomputed per source
    T x = (u * (T(1) + e1) + v * e2) * k * Rp;
    T y = (u * e2 + v * (T(1) - e1)) * k * Rp;
    T den = T(1) + x * x + y * y;
    return T(1) / (den * sqrt(den));
}

template <typename T, int NI>
__global__ void __launch_bounds__(kThreads)
k_forward(FwdArgs<T> A) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const mbx_dims &d = A.d;
    const int NA = d.na.local_size;
    const uint64_t NT = d.ntime.local_size, NC = d.nchan.local_size, NBL = d.nbl.local_size;
    const int ct = A.ct, st = A.st;
    const int tid = threadIdx.x;

    // Workgroup -> (baseline chunk, channel tile, time).
    uint64_t wg = blockIdx.x;
    const int bc = wg % A.nbchunk;
    wg /= A.nbchunk;
    const int ctile = wg % A.nctile;
    const int t = wg / A.nctile;

    const int nt_ext = dext(d.ntime), nc_ext = dext(d.nchan), nb_ext = dext(d.nbl);
    double *sh_part;

    // LDS carve-up (all offsets multiples of 8 bytes; J4 tiles 16-aligned).
    double *s_uvw = reinterpret_cast<double *>(smem_raw);
    double *s_pa = s_uvw + 3 * NA;
    double *s_ch = s_pa + 2 * NA;            // [ct][4]: gc0 gc1 chd freq
    double *s_src = s_ch + 4 * ct;           // [st][3]: l m n-1
    T *s_shape = reinterpret_cast<T *>(s_src + 3 * st);  // [st][4]: p0 p1 p2 (type in slot 3)
    uintptr_t off = reinterpret_cast<uintptr_t>(s_shape + 4 * st);
    off = (off + 15) & ~uintptr_t(15);
    J4<T> *s_sqB = reinterpret_cast<J4<T> *>(off);     // [st][ct]
    J4<T> *s_J = s_sqB + st * ct;                       // [st][NA][ct]
    sh_part = reinterpret_cast<double *>(s_J + st * NA * ct);

    const bool active_wg = t < nt_ext;
    const int c0 = ctile * ct;
    const int b0 = bc * A.bchunk;
    const int b1 = min(nb_ext, b0 + A.bchunk);
    const int nitems = max(0, b1 - b0) * ct;

    if (active_wg) {
        for (int a = tid; a < NA; a += kThreads) {
            const T *x = A.uvw + (uint64_t(t) * NA + a) * 3;
            s_uvw[3 * a] = x[0]; s_uvw[3 * a + 1] = x[1]; s_uvw[3 * a + 2] = x[2];
            double sn, cs;
            sincos(double(A.pa[uint64_t(t) * NA + a]), &sn, &cs);
            s_pa[2 * a] = sn; s_pa[2 * a + 1] = cs;
        }
        const double NUD1 = double(d.beam_nud.local_size - 1);
        for (int k = tid; k < ct; k += kThreads) {
            int c = min(c0 + k, nc_ext - 1);
            double f = A.freq[c];
            double vc = fmin(NUD1, fmax(0.0, A.fscale * (f - A.lfreq)));
            double g0 = floor(vc);
            s_ch[4 * k] = g0;
            s_ch[4 * k + 1] = fmin(g0 + 1.0, NUD1);
            s_ch[4 * k + 2] = vc - g0;
            s_ch[4 * k + 3] = f;
        }
    }

    // Per-item state.
    J4<T> acc[NI];
    int ip[NI], iq[NI], icl[NI];
    T iu[NI], iv[NI];
#pragma unroll
    for (int k = 0; k < NI; ++k) {
        int it = tid + k * kThreads;
        bool ok = active_wg && it < nitems;
        int b = b0 + (ok ? it / ct : 0);
        icl[k] = ok ? it % ct : 0;
        ip[k] = 0; iq[k] = 0;
        if (active_wg && b < nb_ext) {
            // antenna indices are data: clamp so a bad index cannot read
            // outside the per-antenna arrays
            ip[k] = min(max(A.a1[uint64_t(t) * NBL + b], 0), NA - 1);
            iq[k] = min(max(A.a2[uint64_t(t) * NBL + b], 0), NA - 1);
        }
        acc[k] = jzero<T>();
        if (ok && d.nsrc.lower_extent > 0) {
            int c = c0 + icl[k];
            if (c < nc_ext) acc[k] = jload(A.ekb + ((uint64_t(t) * NBL + b) * NC + c) * 4);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NI; ++k) {
        iu[k] = T(s_uvw[3 * iq[k]] - s_uvw[3 * ip[k]]);
        iv[k] = T(s_uvw[3 * iq[k] + 1] - s_uvw[3 * ip[k] + 1]);
    }

    const int np = dext(d.npsrc), ng = dext(d.ngsrc), ns_all = dext(d.nsrc);
    const uint64_t NG = d.ngsrc.local_size, NS = d.nssrc.local_size;

    for (int s0 = 0; s0 < ns_all; s0 += st) {
        const int nst = min(st, ns_all - s0);
        if (active_wg) {
            // phase 0: per-source parameters and sqrt(B) per (source, channel)
            for (int i = tid; i < nst; i += kThreads) {
                int s = s0 + i;
                double l = A.lm[2 * uint64_t(s)], m = A.lm[2 * uint64_t(s) + 1];
                s_src[3 * i] = l; s_src[3 * i + 1] = m;
                s_src[3 * i + 2] = sqrt(1.0 - l * l - m * m) - 1.0;
                T p0 = T(0), p1 = T(0), p2 = T(0), ty = T(0);
                if (s >= np && s < np + ng) {
                    int g = s - np;
                    p0 = A.gauss[g]; p1 = A.gauss[g + NG]; p2 = A.gauss[g + 2 * NG]; ty = T(1);
                } else if (s >= np + ng) {
                    int g = s - np - ng;
                    p0 = A.sersic[g]; p1 = A.sersic[g + NS];
                    p2 = A.sersic[g + 2 * NS] / (T(1) - p0 * p0 - p1 * p1); ty = T(2);
                }
                s_shape[4 * i] = p0; s_shape[4 * i + 1] = p1; s_shape[4 * i + 2] = p2;
                s_shape[4 * i + 3] = ty;
            }
            for (int i = tid; i < nst * ct; i += kThreads) {
                int si = i / ct, k = i % ct;
                int s = s0 + si;
                int c = min(c0 + k, nc_ext - 1);
                const T *S = A.stokes + (uint64_t(s) * NT + t) * 4;
                s_sqB[i] = brightness_sqrt<T>(S[0], S[1], S[2], S[3], A.alpha[uint64_t(s) * NT + t],
                                              A.freq[c], A.ref_freq[c]);
            }
        }
        __syncthreads();
        if (active_wg) {
            // phase 1: per-antenna Jones tile
            const int n1 = nst * NA * ct;
            for (int i = tid; i < n1; i += kThreads) {
                int k = i % ct;
                int a = (i / ct) % NA;
                int si = i / (ct * NA);
                int c = min(c0 + k, nc_ext - 1);
                s_J[i] = antenna_jones<T>(A, t, a, c, s_src[3 * si], s_src[3 * si + 1],
                                          s_src[3 * si + 2], s_sqB[si * ct + k], s_pa[2 * a],
                                          s_pa[2 * a + 1], s_uvw + 3 * a, s_ch[4 * k],
                                          s_ch[4 * k + 1], T(s_ch[4 * k + 2]), s_ch[4 * k + 3]);
            }
        }
        __syncthreads();
        if (active_wg) {
            // phase 2: coherency accumulation
            for (int si = 0; si < nst; ++si) {
                const J4<T> *Js = s_J + si * NA * ct;
                const T ty = s_shape[4 * si + 3];
                if (ty == T(0)) {
#pragma unroll
                    for (int k = 0; k < NI; ++k) {
                        if (tid + k * kThreads < nitems)
                            jacc_h(acc[k], Js[ip[k] * ct + icl[k]], Js[iq[k] * ct + icl[k]]);
                    }
                } else {
                    const T p0 = s_shape[4 * si], p1 = s_shape[4 * si + 1], p2 = s_shape[4 * si + 2];
                    const bool gauss = ty == T(1);
                    const T kscale = T(gauss ? A.gauss_scale : A.two_pi_over_c);
#pragma unroll
                    for (int k = 0; k < NI; ++k) {
                        if (tid + k * kThreads < nitems) {
                            T kf = kscale * T(s_ch[4 * icl[k] + 3]);
                            T f = gauss ? gauss_f(iu[k], iv[k], p0, p1, p2, kf)
                                        : sersic_f(iu[k], iv[k], p0, p1, p2, kf);
                            jacc_h(acc[k], jscale(Js[ip[k] * ct + icl[k]], f), Js[iq[k] * ct + icl[k]]);
                        }
                    }
                }
            }
        }
        __syncthreads();
    }

    // Epilogue.
    const bool final_batch = d.nsrc.upper_extent >= d.nsrc.global_size;
    double term_sum = 0.0;
    if (active_wg) {
#pragma unroll
        for (int k = 0; k < NI; ++k) {
            int it = tid + k * kThreads;
            if (it >= nitems) continue;
            int b = b0 + it / ct;
            int c = c0 + icl[k];
            if (c >= nc_ext) continue;
            uint64_t vi = (uint64_t(t) * NBL + b) * NC + c;
            if (A.ekb) jstore(A.ekb + vi * 4, acc[k]);
            if (!final_batch) continue;
            const uint64_t gp = ((uint64_t(t) * NA + ip[k]) * NC + c) * 4;
            const uint64_t gq = ((uint64_t(t) * NA + iq[k]) * NC + c) * 4;
            J4<T> M = jmul_h(jmul(jload(A.G + gp), acc[k]), jload(A.G + gq));
            J4<T> O = jload(A.obs + vi * 4);
            uchar4 f4 = *reinterpret_cast<const uchar4 *>(A.flag + vi * 4);
            const uint8_t fa[4] = {f4.x, f4.y, f4.z, f4.w};
            J4<T> R;
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                if (fa[p] > 0) { M.v[p] = {T(0), T(0)}; O.v[p] = {T(0), T(0)}; }
                R.v[p] = {O.v[p].re - M.v[p].re, O.v[p].im - M.v[p].im};
            }
            jstore(A.model + vi * 4, (A.options & MBX_RESIDUAL_OUTPUT) ? R : M);
            T term = T(0);
            if (A.options & MBX_APPLY_WEIGHTS) {
                const T *w = A.wv + vi * 4;
#pragma unroll
                for (int p = 0; p < 4; ++p)
                    term += (R.v[p].re * R.v[p].re + R.v[p].im * R.v[p].im) * w[p];
            } else {
#pragma unroll
                for (int p = 0; p < 4; ++p) term += R.v[p].re * R.v[p].re + R.v[p].im * R.v[p].im;
            }
            A.chi2[vi] = term;
            term_sum += double(term);
        }
    }
    // Workgroup reduction of the chi^2 partial (wave64 shuffles, then LDS).
    for (int off2 = 32; off2 > 0; off2 >>= 1) term_sum += __shfl_down(term_sum, off2, 64);
    if ((tid & 63) == 0) sh_part[tid >> 6] = term_sum;
    __syncthreads();
    if (tid == 0) {
        double s = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) s += sh_part[w];
        A.partial[blockIdx.x] = s;
    }
}

template <typename T>
static size_t forward_ws_bytes(const mbx_dims &d) {
    FwdPlan p = plan_forward<T>(d);
    size_t beam = size_t(d.beam_lw.local_size) * d.beam_mh.local_size * d.beam_nud.local_size * 4 *
                  sizeof(BeamTap<T>);
    beam = (beam + 255) & ~size_t(255);
    return beam + p.nwg * sizeof(double);
}

template <typename T>
static int forward(const mbx_dims *d, const mbx_inputs *in, const mbx_outputs *out,
                   const mbx_beam_extents *b, const mbx_consts *k, int options, void *ws,
                   size_t ws_bytes, hipStream_t stream) {
    int r = check_dims(d, true);
    if (r) return r;
    if (!in || !out || !b || !k) return set_error(MBX_E_NULL, "forward: NULL argument struct");
    if (!in->uvw || !in->frequency || !in->ref_frequency || !in->parallactic_angles ||
        !in->point_errors || !in->antenna_scaling || !in->E_beam || !in->antenna1 ||
        !in->antenna2 || !in->flag || !in->observed_vis || !in->G_term)
        return set_error(MBX_E_NULL, "forward: NULL input array");
    if (dext(d->nsrc) > 0 && (!in->lm || !in->stokes || !in->alpha))
        return set_error(MBX_E_NULL, "forward: NULL source array");
    if (dext(d->ngsrc) > 0 && !in->gauss_shape) return set_error(MBX_E_NULL, "forward: gauss_shape is NULL");
    if (dext(d->nssrc) > 0 && !in->sersic_shape) return set_error(MBX_E_NULL, "forward: sersic_shape is NULL");
    if ((options & MBX_APPLY_WEIGHTS) && !in->weight_vector)
        return set_error(MBX_E_NULL, "forward: weight_vector is NULL");
    if (options & ~(MBX_APPLY_WEIGHTS | MBX_RESIDUAL_OUTPUT))
        return set_error(MBX_E_UNSUPPORTED, "forward: unknown option bits 0x%x", options);
    const bool final_batch = d->nsrc.upper_extent >= d->nsrc.global_size;
    if (final_batch && (!out->model_vis || !out->chi_sqrd_result || !out->X2))
        return set_error(MBX_E_NULL, "forward: NULL output array");
    if ((!final_batch || d->nsrc.lower_extent > 0) && !out->ekb_vis)
        return set_error(MBX_E_NULL, "forward: a partial source batch needs ekb_vis");
    if (!(b->ul != b->ll && b->um != b->lm && b->ufreq != b->lfreq))
        return set_error(MBX_E_DIMS, "forward: degenerate beam extents");
    size_t need = forward_ws_bytes<T>(*d);
    if (!ws || ws_bytes < need)
        return set_error(MBX_E_WORKSPACE, "forward: workspace needs %zu bytes (got %zu)", need, ws_bytes);

    FwdPlan p = plan_forward<T>(*d);
    const bool empty = dext(d->ntime) == 0 || dext(d->nbl) == 0 || dext(d->nchan) == 0;
    const uint64_t ncube = uint64_t(d->beam_lw.local_size) * d->beam_mh.local_size *
                           d->beam_nud.local_size * 4;
    BeamTap<T> *beam = static_cast<BeamTap<T> *>(ws);
    size_t beam_bytes = (ncube * sizeof(BeamTap<T>) + 255) & ~size_t(255);
    double *partial = reinterpret_cast<double *>(static_cast<char *>(ws) + beam_bytes);

    unsigned g = (unsigned)std::min<uint64_t>((ncube + 255) / 256, 65536);
    hipLaunchKernelGGL(k_beam_prepare<T>, dim3(g), dim3(256), 0, stream,
                       static_cast<const cx<T> *>(in->E_beam), beam, ncube);

    FwdArgs<T> A;
    A.d = *d;
    A.uvw = (const T *)in->uvw; A.lm = (const T *)in->lm; A.stokes = (const T *)in->stokes;
    A.alpha = (const T *)in->alpha; A.freq = (const T *)in->frequency;
    A.ref_freq = (const T *)in->ref_frequency; A.pa = (const T *)in->parallactic_angles;
    A.pe = (const T *)in->point_errors; A.ascale = (const T *)in->antenna_scaling;
    A.gauss = (const T *)in->gauss_shape; A.sersic = (const T *)in->sersic_shape;
    A.wv = (const T *)in->weight_vector; A.beam = beam;
    A.a1 = in->antenna1; A.a2 = in->antenna2; A.flag = in->flag;
    A.obs = (const cx<T> *)in->observed_vis; A.G = (const cx<T> *)in->G_term;
    A.model = (cx<T> *)out->model_vis; A.chi2 = (T *)out->chi_sqrd_result;
    A.ekb = (cx<T> *)out->ekb_vis; A.partial = partial;
    A.ll = b->ll; A.lmin = b->lm;
    A.lscale = double(d->beam_lw.local_size - 1) / (b->ul - b->ll);
    A.mscale = double(d->beam_mh.local_size - 1) / (b->um - b->lm);
    A.lfreq = b->lfreq;
    A.fscale = double(d->beam_nud.local_size - 1) / (b->ufreq - b->lfreq);
    A.gauss_scale = k->gauss_scale; A.two_pi_over_c = k->two_pi_over_c;
    A.options = options;
    A.ct = p.ct; A.st = p.st; A.nbchunk = p.nbchunk; A.bchunk = p.bchunk; A.nctile = p.nctile;

    dim3 grid((unsigned)p.nwg), block(kThreads);
    if (empty) {
        if (final_batch) launch_chi2_final(partial, 0, 1.0, out->X2, stream);
        return hip_status(hipGetLastError(), "forward (empty extents)");
    }
    switch (p.ni) {
    case 1: hipLaunchKernelGGL((k_forward<T, 1>), grid, block, p.lds_bytes, stream, A); break;
    case 2: hipLaunchKernelGGL((k_forward<T, 2>), grid, block, p.lds_bytes, stream, A); break;
    case 4: hipLaunchKernelGGL((k_forward<T, 4>), grid, block, p.lds_bytes, stream, A); break;
    default: hipLaunchKernelGGL((k_forward<T, 8>), grid, block, p.lds_bytes, stream, A); break;
    }
    r = hip_status(hipGetLastError(), "forward launch");
    if (r) return r;
    if (final_batch) {
        double scale = (options & MBX_APPLY_WEIGHTS) ? 1.0 : 1.0 / k->sigma_sqrd;
        launch_chi2_final(partial, (unsigned)p.nwg, scale, out->X2, stream);
        r = hip_status(hipGetLastError(), "forward chi2 launch");
    }
    return r;
}

}  // namespace mbx

using namespace mbx;

extern "C" {

size_t mbx_forward_workspace_bytes(const mbx_dims *d, int is_double) {
    if (!d) return 0;
    return is_double ? forward_ws_bytes<double>(*d) : forward_ws_bytes<float>(*d);
}

int mbx_forward_f32(const mbx_dims *d, const mbx_inputs *in, const mbx_outputs *out,
                    const mbx_beam_extents *beam, const mbx_consts *k, int options, void *ws,
                    size_t ws_bytes, void *s) {
    return forward<float>(d, in, out, beam, k, options, ws, ws_bytes, (hipStream_t)s);
}
int mbx_forward_f64(const mbx_dims *d, const mbx_inputs *in, const mbx_outputs *out,
                    const mbx_beam_extents *beam, const mbx_consts *k, int options, void *ws,
                    size_t ws_bytes, void *s) {
    return forward<double>(d, in, out, beam, k, options, ws, ws_bytes, (hipStream_t)s);
}

}  // extern "C"
