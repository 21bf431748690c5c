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

#include "mbx_backward.hpp"

namespace mbx {
template <typename T> bool adjoint_gemm_supported(const mbx_dims &d);
template <typename T> int launch_adjoint_gemm(const AdjArgs<T> &A, hipStream_t stream, unsigned nrep = 1);
}  // namespace mbx

namespace mbx {

// ---------------------------------------------------------------------
// Visibility stage: Mb, Vb and the gain adjoint.

template <typename T, bool BATCHED = false>
__global__ void __launch_bounds__(256)
k_adj_vis(AdjArgs<T> A) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double *Gacc = reinterpret_cast<double *>(smem_raw);   // [NA][ct2][8]
    // in a launch over several samples: this workgroup's sample and slot
    const SampleAt at = enter_sample_if<BATCHED>(A);
    const RimeArgs<T> &R = A.R;
    const mbx_dims &d = R.d;
    const int NA = d.na.local_size;
    const uint64_t NC = d.nchan.local_size, NBL = d.nbl.local_size;
    const int ct2 = A.ct2;
    const int tid = threadIdx.x;
    const int ctile = at.bx % A.nctile2;
    const int t = at.bx / A.nctile2;
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
    size_t beam = beam_planes_bytes<T>(d);
    size_t vbar = size_t(d.ntime.local_size) * d.nbl.local_size * d.nchan.local_size * 4 * sizeof(cx<T>);
    return beam + vbar;
}

template <typename T> constexpr int adj_stb() { return sizeof(T) == 4 ? 8 : 4; }

// Channel tiling of k_adj_vis (gain adjoint sums of a tile in LDS); its
// LDS bytes.
template <typename T>
static size_t adj_vis_tiling(const mbx_dims &d, AdjArgs<T> &A) {
    const int na = d.na.local_size;
    A.ct2 = std::max(1, std::min<int>(8, d.nchan.local_size));
    while (A.ct2 > 1 && size_t(na) * A.ct2 * 8 * sizeof(double) > 48 * 1024) --A.ct2;
    A.nctile2 = (d.nchan.local_size + A.ct2 - 1) / A.ct2;
    return size_t(na) * A.ct2 * 8 * sizeof(double);
}

// Adjoint of nrep samples (A.R.span set with its slot0; A.ekb and A.vbar
// the first slot's buffers; gradients zeroed by the caller): the
// visibility stage and the matrix-core source passes, each pass one
// launch over all nrep samples.
template <typename T>
int launch_adjoint_samples(AdjArgs<T> A, unsigned nrep, hipStream_t stream) {
    const mbx_dims &d = A.R.d;
    if (!adjoint_gemm_supported<T>(d))
        return set_error(MBX_E_UNSUPPORTED, "batched adjoint: %u antennas", d.na.local_size);
    if (A.g_G) return set_error(MBX_E_UNSUPPORTED, "batched adjoint: gain gradients are not produced");
    const uint64_t NT = d.ntime.local_size;
    const size_t lds1 = adj_vis_tiling(d, A);
    A.R.span.wgs = unsigned(NT * A.nctile2);
    hipLaunchKernelGGL((k_adj_vis<T, true>), dim3(unsigned(NT * A.nctile2 * nrep)), dim3(256), lds1, stream, A);
    const int r = hip_status(hipGetLastError(), "batched adjoint vis launch");
    if (r) return r;
    A.R.span.wgs = 8 * tc_per_xcd(dext(d.ntime), dext(d.nchan));
    return launch_adjoint_gemm<T>(A, stream, nrep);
}
template int launch_adjoint_samples<float>(AdjArgs<float>, unsigned, hipStream_t);
template int launch_adjoint_samples<double>(AdjArgs<double>, unsigned, hipStream_t);

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

    BeamCell<T> *beam = static_cast<BeamCell<T> *>(ws);
    cx<T> *vbar = reinterpret_cast<cx<T> *>(static_cast<char *>(ws) + beam_planes_bytes<T>(*d));
    launch_beam_planes<T>(*d, in->E_beam, in->frequency, b, beam, stream);

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
    const size_t lds1 = adj_vis_tiling(*d, A);
    hipLaunchKernelGGL(k_adj_vis<T>, dim3(unsigned(NT * A.nctile2)), dim3(256), lds1, stream, A);
    r = hip_status(hipGetLastError(), "adjoint vis launch");
    if (r) return r;

    // source stage: on the matrix cores with the visibility adjoint staged
    // in LDS where it fits (na <= 64), else one thread per (antenna,
    // channel) of a tile
    if (adjoint_gemm_supported<T>(*d)) return launch_adjoint_gemm<T>(A, stream);
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
