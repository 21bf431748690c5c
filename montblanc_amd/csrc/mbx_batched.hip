// Batched RIME: many sky-model samples (each with its own, unevenly sized
// source list) evaluated against one shared set of visibility-space inputs
// (uvw, beam, gains, observed visibilities, flags, weights). No reference
// counterpart. Two strategies (mbx_batch_tuning.strategy):
//
// 1: the matrix-core kernels of the single-model path (BATCHED instances
// of k_forward_mx, k_adj_vis, k_adj_point, k_adj_gemm) each launched once
// over a group of samples, every sample a grid of (time, channel)
// workgroups (SampleSpan, mbx_tile.hpp); per-slot chi^2 from each slot's
// partials (k_chi2_slots). Host loop: batched_mx below.
//
// 0: ONE persistent launch on the vector units, k_batched below. Work
// items are (sample, time, channel tile); the host orders samples by
// decreasing source count so the longest items start first, and the
// workgroups pull items from an atomic counter until the queue is empty.
// Per item a workgroup
//   1. accumulates the coherency sums of the sample's sources for every
//      baseline (per-antenna Jones tiles in LDS, sums in registers),
//   2. forms M = G_p V G_q^H and the chi^2 terms against the shared
//      observations; the float64 workgroup sum goes to X2[sample] by an
//      atomic add,
//   3. when gradients are requested, writes Vb = G_p^H Mb G_q to the
//      workgroup's scratch slot and runs the adjoint source loop of
//      mbx_adjoint.hip for the sample (recomputing its Jones tiles),
//      adding into the sample's slices of the gradient arrays.
// Model visibilities of the samples are never written to HBM.

#include <algorithm>

#include "mbx_backward.hpp"
#include "mbx_gemm.hpp"

namespace mbx {

template <typename T> bool adjoint_gemm_supported(const mbx_dims &d);
template <typename T> int launch_adjoint_samples(AdjArgs<T> A, unsigned nrep, hipStream_t stream);

struct BatchDev {
    const int32_t *src_offset, *npsrc, *ngsrc, *nssrc, *gauss_offset, *sersic_offset, *order;
    int nsamples;
};

template <typename T> struct BatchArgs {
    RimeArgs<T> R;          // shared inputs; source pointers are the concatenated arrays
    BatchDev b;
    double *X2;             // [nsamples]
    cx<T> *slots;           // per-workgroup Vb scratch: [nwg][nbl][ct]
    int *counter;           // work-queue head
    const int32_t *bmap;
    double *g_lm, *g_stokes, *g_alpha, *g_gauss, *g_sersic;
    double inv_sigma2;      // 1/sigma^2 (1 with weights)
    int nitems, grads;
    int st_fwd;             // sources per forward tile
};

constexpr int kBatThreads = 512;

// Per-sample view of the shared argument block: source pointers offset to
// the sample's slices and source-count extents set to the sample's.
template <typename T>
__device__ __forceinline__ void sample_view(const BatchArgs<T> &B, int k, RimeArgs<T> &R,
                                            int &src0) {
    const uint64_t NT = B.R.d.ntime.local_size;
    src0 = B.b.src_offset[k];
    const int np = B.b.npsrc[k], ng = B.b.ngsrc[k], nss = B.b.nssrc[k];
    R.lm = B.R.lm + 2 * uint64_t(src0);
    R.stokes = B.R.stokes + uint64_t(src0) * NT * 4;
    R.alpha = B.R.alpha + uint64_t(src0) * NT;
    R.gauss = B.R.gauss ? B.R.gauss + B.b.gauss_offset[k] : nullptr;
    R.sersic = B.R.sersic ? B.R.sersic + B.b.sersic_offset[k] : nullptr;
    R.d.npsrc.lower_extent = 0; R.d.npsrc.upper_extent = np;
    R.d.ngsrc.lower_extent = 0; R.d.ngsrc.upper_extent = ng;
    R.d.nssrc.lower_extent = 0; R.d.nssrc.upper_extent = nss;
    R.d.nsrc.lower_extent = 0; R.d.nsrc.upper_extent = np + ng + nss;
    R.d.nsrc.global_size = np + ng + nss;
}

template <typename T, int NI, int RS>
__global__ void __launch_bounds__(kBatThreads)
k_batched(BatchArgs<T> B) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int s_item;
    const mbx_dims &d = B.R.d;
    const int NA = d.na.local_size;
    const uint64_t NC = d.nchan.local_size, NBL = d.nbl.local_size, NT = d.ntime.local_size;
    const int ct = B.R.ct;
    const int tid = threadIdx.x;
    const int nb_ext = dext(d.nbl), nc_ext = dext(d.nchan), nt_ext = dext(d.ntime);
    const int per_sample = nt_ext * B.R.nctile;
    cx<T> *slot = B.slots + uint64_t(blockIdx.x) * NBL * ct * 4;
    const bool weights = B.R.options & MBX_APPLY_WEIGHTS;

    for (;;) {
        if (tid == 0) s_item = atomicAdd(B.counter, 1);
        __syncthreads();
        const int item = s_item;
        __syncthreads();
        if (item >= B.nitems) break;
        const int k = B.b.order[item / per_sample];
        const int rem = item % per_sample;
        const int t = rem / B.R.nctile;
        const int c0 = (rem % B.R.nctile) * ct;
        RimeArgs<T> R = B.R;
        int src0;
        sample_view(B, k, R, src0);
        const int ns_all = dext(R.d.nsrc);

        // ---- forward: NI (baseline, channel) items per thread
        const int st = B.st_fwd;
        Tile<T> L(smem_raw, NA, ct, st);
        R.st = st;
        tile_setup(R, L, t, c0, tid, kBatThreads);
        const int nitems = nb_ext * ct;
        J4<T> acc[NI];
        int ip[NI], iq[NI];
        T iu[NI], iv[NI];
#pragma unroll
        for (int x = 0; x < NI; ++x) {
            const int it = tid + x * kBatThreads;
            const int b = it < nitems ? it / ct : 0;
            ip[x] = min(max(R.a1[uint64_t(t) * NBL + b], 0), NA - 1);
            iq[x] = min(max(R.a2[uint64_t(t) * NBL + b], 0), NA - 1);
            acc[x] = jzero<T>();
        }
        __syncthreads();
#pragma unroll
        for (int x = 0; x < NI; ++x) {
            iu[x] = T(L.uvw[3 * iq[x]] - L.uvw[3 * ip[x]]);
            iv[x] = T(L.uvw[3 * iq[x] + 1] - L.uvw[3 * ip[x] + 1]);
        }
        for (int s0 = 0; s0 < ns_all; s0 += st) {
            const int nst = min(st, ns_all - s0);
            tile_phase0(R, L, t, c0, s0, nst, tid, kBatThreads);
            __syncthreads();
            tile_phase1(R, L, t, c0, nst, tid, kBatThreads);
            __syncthreads();
            for (int si = 0; si < nst; ++si) {
                const J4<T> *Js = L.J + si * NA * ct;
                const T ty = L.shape[4 * si + 3];
                const T p0 = L.shape[4 * si], p1 = L.shape[4 * si + 1], p2 = L.shape[4 * si + 2];
#pragma unroll
                for (int x = 0; x < NI; ++x) {
                    const int it = tid + x * kBatThreads;
                    if (it >= nitems) continue;
                    const int kk = it % ct;
                    const J4<T> &Jp = Js[ip[x] * ct + kk];
                    const J4<T> &Jq = Js[iq[x] * ct + kk];
                    if (ty == T(0)) {
                        jacc_h(acc[x], Jp, Jq);
                    } else {
                        const T fc = T(L.ch[CH * kk + 3]);
                        const T f = ty == T(1) ? gauss_f(iu[x], iv[x], p0, p1, p2, T(R.gauss_scale) * fc)
                                               : sersic_f(iu[x], iv[x], p0, p1, p2, T(R.two_pi_over_c) * fc);
                        jacc_h(acc[x], jscale(Jp, f), Jq);
                    }
                }
            }
            __syncthreads();
        }
        // ---- chi^2 and the visibility adjoint Vb
        double term = 0.0;
#pragma unroll
        for (int x = 0; x < NI; ++x) {
            const int it = tid + x * kBatThreads;
            if (it >= nitems) continue;
            const int b = it / ct, kk = it % ct, c = c0 + kk;
            if (c >= nc_ext) continue;
            const uint64_t vi = (uint64_t(t) * NBL + b) * NC + c;
            const J4<T> Gp = jload(R.G + ((uint64_t(t) * NA + ip[x]) * NC + c) * 4);
            const J4<T> Gq = jload(R.G + ((uint64_t(t) * NA + iq[x]) * NC + c) * 4);
            const J4<T> M = jmul_h(jmul(Gp, acc[x]), Gq);
            const J4<T> O = jload(R.obs + vi * 4);
            const uchar4 f4 = *reinterpret_cast<const uchar4 *>(R.flag + vi * 4);
            const uint8_t fa[4] = {f4.x, f4.y, f4.z, f4.w};
            J4<T> Mb;
            T tt = T(0);
#pragma unroll
            for (int z = 0; z < 4; ++z) {
                T w = weights ? R.wv[vi * 4 + z] : T(1);
                if (fa[z] > 0) w = T(0);
                const T rr = O.v[z].re - M.v[z].re, ri = O.v[z].im - M.v[z].im;
                tt += w * (rr * rr + ri * ri);
                Mb.v[z] = {-w * rr, -w * ri};
            }
            term += double(tt);
            if (B.grads) jstore(slot + (uint64_t(b) * ct + kk) * 4, jmul(jmul(jherm(Gp), Mb), Gq));
        }
        wg_partial(term, L.part, &L.part[8], tid, kBatThreads);
        if (tid == 0) atomicAdd(B.X2 + k, L.part[8] * B.inv_sigma2);
        __syncthreads();
        if (!B.grads) continue;

        // ---- adjoint source loop: thread (antenna p, channel kk, group g)
        // accumulates Jb for sources g, g + ngrp, ... of each tile (RS per thread)
        const int rows = NA * ct;
        const int ngrp = max(1, kBatThreads / rows);
        const int SB = ngrp * RS;                       // sources per adjoint tile
        Tile<T> LA(smem_raw, NA, ct, SB);
        double *sacc = LA.part + kBatThreads / 64;      // [SB][10]
        R.st = SB;
        for (int i = tid; i < SB * 10; i += kBatThreads) sacc[i] = 0.0;
        const int grp = tid / rows, p = (tid % rows) / ct, kk = tid % ct;
        const int c = c0 + kk;
        const bool mine = grp < ngrp && c < nc_ext;
        const uint64_t NG = d.ngsrc.local_size, NS = d.nssrc.local_size;
        __syncthreads();
        const T up = T(LA.uvw[3 * p]), vp = T(LA.uvw[3 * p + 1]);
        const T fc = T(LA.ch[CH * kk + 3]);
        const T kf_g = T(R.gauss_scale) * fc, kf_s = T(R.two_pi_over_c) * fc;
        const int np = dext(R.d.npsrc), ng = dext(R.d.ngsrc);
        for (int s0 = 0; s0 < ns_all; s0 += SB) {
            const int nst = min(SB, ns_all - s0);
            tile_phase0(R, LA, t, c0, s0, nst, tid, kBatThreads);
            __syncthreads();
            tile_phase1(R, LA, t, c0, nst, tid, kBatThreads);
            __syncthreads();
            J4<T> Jb[RS];
            T sh[RS][3];
#pragma unroll
            for (int r = 0; r < RS; ++r) {
                Jb[r] = jzero<T>();
                sh[r][0] = sh[r][1] = sh[r][2] = T(0);
            }
            if (mine) {
                const int32_t *row = B.bmap + uint64_t(p) * NA;
                for (int q = 0; q < NA; ++q) {
                    const int code = row[q];
                    if (code < 0) continue;
                    const int b = code & ((1 << 29) - 1);
                    const bool rev = code & MBX_BL_REVERSED, aut = code & MBX_BL_AUTO;
                    const J4<T> Wb = jload(slot + (uint64_t(b) * ct + kk) * 4);
                    J4<T> W = rev ? jherm(Wb) : Wb;
                    if (aut) {
                        const J4<T> H = jherm(Wb);
#pragma unroll
                        for (int z = 0; z < 4; ++z) W.v[z] = {Wb.v[z].re + H.v[z].re, Wb.v[z].im + H.v[z].im};
                    }
                    const T u = T(LA.uvw[3 * q]) - up, v = T(LA.uvw[3 * q + 1]) - vp;
#pragma unroll
                    for (int r = 0; r < RS; ++r) {
                        const int si = grp + r * ngrp;
                        if (si >= nst) break;
                        const J4<T> Y = jmul(W, LA.J[(si * NA + q) * ct + kk]);
                        const int ty = int(LA.shape[4 * si + 3]);
                        if (ty == 0) {
#pragma unroll
                            for (int z = 0; z < 4; ++z) {
                                Jb[r].v[z].re += Y.v[z].re;
                                Jb[r].v[z].im += Y.v[z].im;
                            }
                        } else {
                            T df[3];
                            const T f = shape_factor_grad<T>(ty, u, v, LA.shape[4 * si], LA.shape[4 * si + 1],
                                                        LA.shape[4 * si + 2], ty == 1 ? kf_g : kf_s, df);
#pragma unroll
                            for (int z = 0; z < 4; ++z) {
                                Jb[r].v[z].re = fma(f, Y.v[z].re, Jb[r].v[z].re);
                                Jb[r].v[z].im = fma(f, Y.v[z].im, Jb[r].v[z].im);
                            }
                            if (!rev) {
                                const T fb = (aut ? T(1) : T(2)) *
                                             jdot_re(LA.J[(si * NA + p) * ct + kk], Y);
                                sh[r][0] = fma(fb, df[0], sh[r][0]);
                                sh[r][1] = fma(fb, df[1], sh[r][1]);
                                sh[r][2] = fma(fb, df[2], sh[r][2]);
                            }
                        }
                    }
                }
            }
            __syncthreads();
            if (mine) {
#pragma unroll
                for (int r = 0; r < RS; ++r) {
                    const int si = grp + r * ngrp;
                    if (si >= nst) break;
                    LA.J[(si * NA + p) * ct + kk] = Jb[r];
                    if (int(LA.shape[4 * si + 3]) != 0) {
                        atomicAdd(sacc + si * 10 + 7, double(sh[r][0]));
                        atomicAdd(sacc + si * 10 + 8, double(sh[r][1]));
                        atomicAdd(sacc + si * 10 + 9, double(sh[r][2]));
                    }
                }
            }
            __syncthreads();
            const int n3 = nst * NA * ct;
            for (int i = tid; i < n3; i += kBatThreads) {
                const int k3 = i % ct, a = (i / ct) % NA, si = i / (ct * NA);
                if (c0 + k3 >= nc_ext) continue;
                double g[7];
                antenna_backward<T>(R, LA.uvw, LA.pa, LA.ch, t, s0 + si, a, c0 + k3, k3, LA.J[i], g);
#pragma unroll
                for (int z = 0; z < 7; ++z) atomicAdd(sacc + si * 10 + z, g[z]);
            }
            __syncthreads();
            for (int si = tid; si < nst; si += kBatThreads) {
                const int s = src0 + s0 + si;         // index in the concatenated arrays
                const double *vv = sacc + si * 10;
                const double gs = B.inv_sigma2;
                if (B.g_lm) {
                    atomicAdd(B.g_lm + 2 * uint64_t(s), gs * vv[0]);
                    atomicAdd(B.g_lm + 2 * uint64_t(s) + 1, gs * vv[1]);
                }
                if (B.g_stokes)
                    for (int z = 0; z < 4; ++z)
                        atomicAdd(B.g_stokes + (uint64_t(s) * NT + t) * 4 + z, gs * vv[2 + z]);
                if (B.g_alpha) atomicAdd(B.g_alpha + uint64_t(s) * NT + t, gs * vv[6]);
                const int sl = s0 + si;
                if (sl >= np && sl < np + ng && B.g_gauss) {
                    const int gi = B.b.gauss_offset[k] + sl - np;
                    for (int z = 0; z < 3; ++z) atomicAdd(B.g_gauss + gi + z * NG, gs * vv[7 + z]);
                } else if (sl >= np + ng && B.g_sersic) {
                    const int gi = B.b.sersic_offset[k] + sl - np - ng;
                    for (int z = 0; z < 3; ++z) atomicAdd(B.g_sersic + gi + z * NS, gs * vv[7 + z]);
                }
            }
            __syncthreads();
            for (int i = tid; i < SB * 10; i += kBatThreads) sacc[i] = 0.0;
            __syncthreads();
        }
    }
}

__global__ void __launch_bounds__(256) k_fill_batch(double *p, uint64_t n, int *counter) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
    if (counter && blockIdx.x == 0 && threadIdx.x == 0) *counter = 0;
}

static void zero_doubles(double *p, uint64_t n, hipStream_t st) {
    if (!p || n == 0) return;
    unsigned g = (unsigned)std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_batch, dim3(g), dim3(256), 0, st, p, n, (int *)nullptr);
}

int validate_inputs(const mbx_dims *d, const mbx_inputs *in, const mbx_beam_extents *b,
                    const mbx_consts *k, int options, const char *who);

struct BatPlan {
    int ct, ni, st_fwd, rs, nctile, nwg;
    size_t lds;
};

template <typename T> constexpr int bat_rs() { return sizeof(T) == 4 ? 2 : 1; }

template <typename T>
static int plan_batched(const mbx_dims &d, const mbx_batch_tuning *tune, BatPlan &p) {
    const int na = d.na.local_size, nb = d.nbl.local_size;
    p.ct = (tune && tune->channels_per_item > 0) ? tune->channels_per_item : 1;
    p.ct = std::max(1, std::min<int>(p.ct, d.nchan.local_size));
    const int need = (nb * p.ct + kBatThreads - 1) / kBatThreads;
    if (need > 8) return set_error(MBX_E_UNSUPPORTED, "batched: %d baselines x %d channels per "
                                   "item exceed 8 per thread; use fewer channels per item", nb, p.ct);
    p.ni = need <= 1 ? 1 : need <= 2 ? 2 : need <= 4 ? 4 : 8;
    if (na * p.ct > kBatThreads)
        return set_error(MBX_E_UNSUPPORTED, "batched: %d antennas x %d channels exceed %d threads",
                         na, p.ct, kBatThreads);
    p.rs = bat_rs<T>();
    const int ngrp = std::max(1, kBatThreads / (na * p.ct));
    const int sb = ngrp * p.rs;
    int st = (tune && tune->sources_per_tile > 0) ? tune->sources_per_tile : 32;
    size_t budget = 60 * 1024;
    while (st > 1 && tile_lds_bytes<T>(na, p.ct, st, kBatThreads / 64 + 2) > budget) st >>= 1;
    p.st_fwd = st;
    size_t lf = tile_lds_bytes<T>(na, p.ct, st, kBatThreads / 64 + 2);
    size_t la = tile_lds_bytes<T>(na, p.ct, sb, kBatThreads / 64) + sizeof(double) * sb * 10;
    p.lds = std::max(lf, la);
    if (p.lds > 64 * 1024) return set_error(MBX_E_UNSUPPORTED, "batched: LDS tile too large (%zu)", p.lds);
    p.nctile = (d.nchan.local_size + p.ct - 1) / p.ct;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 256;
    p.nwg = (tune && tune->workgroups > 0) ? tune->workgroups : 2 * cus;
    return MBX_OK;
}

// ---- strategy 1: the matrix-core passes over groups of samples ---------

static uint64_t mx_vstride(const mbx_dims &d) {
    return uint64_t(d.ntime.local_size) * d.nbl.local_size * d.nchan.local_size * 4;
}

static size_t mx_partial_bytes(const mbx_dims &d, const mbx_batch_tuning *tune) {
    return (uint64_t(tune->samples_per_group) * gemm_workgroups(d) * sizeof(double) + 255) & ~size_t(255);
}

// Beam planes, the chi^2 partials of a group's samples, then per slot of a
// group the running / pre-gain sums, then per slot Vb.
template <typename T>
static size_t batched_mx_ws_bytes(const mbx_dims &d, const mbx_batch_tuning *tune) {
    if (sizeof(T) == 8) {
        set_error(MBX_E_UNSUPPORTED, "batched: the matrix-core strategy is float only");
        return 0;
    }
    if (tune->samples_per_group <= 0) {
        set_error(MBX_E_DIMS, "batched: samples_per_group must be positive (got %d)", tune->samples_per_group);
        return 0;
    }
    return beam_planes_bytes<T>(d) + mx_partial_bytes(d, tune) +
           2 * uint64_t(tune->samples_per_group) * mx_vstride(d) * sizeof(cx<T>);
}

// X2[order[slot0 + j]] = scale x (the sum of slot j's nwg partials, in a
// fixed order): one block per slot.
__global__ void __launch_bounds__(256)
k_chi2_slots(const double *__restrict__ partial, unsigned nwg, const int32_t *__restrict__ order,
             double scale, double *__restrict__ X2) {
    __shared__ double red[256];
    const double *p = partial + uint64_t(blockIdx.x) * nwg;
    double acc = 0.0;
    for (unsigned i = threadIdx.x; i < nwg; i += 256) acc += p[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (unsigned w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) X2[order[blockIdx.x]] = red[0] * scale;
}

// Gradients (if any) are zeroed by the caller.
template <typename T>
static int batched_mx(const mbx_dims *d, const mbx_inputs *in, const mbx_batch *bt, double *X2,
                      const mbx_grads *gr, bool grads, const int32_t *bmap, const mbx_beam_extents *b,
                      const mbx_consts *k, const mbx_batch_tuning *tune, int options, void *ws,
                      size_t ws_bytes, hipStream_t stream) {
    if constexpr (sizeof(T) == 8) {
        (void)d; (void)in; (void)bt; (void)X2; (void)gr; (void)grads; (void)bmap; (void)b; (void)k;
        (void)tune; (void)options; (void)ws; (void)ws_bytes; (void)stream;
        return set_error(MBX_E_UNSUPPORTED, "batched: the matrix-core strategy is float only");
    } else {
        if (!bmap) return set_error(MBX_E_NULL, "batched: the matrix-core strategy needs the baseline map");
        if (!batched_mx_supported(*d))
            return set_error(MBX_E_UNSUPPORTED, "batched: the matrix-core passes do not fit %u antennas",
                             d->na.local_size);
        if (grads && !adjoint_gemm_supported<T>(*d))
            return set_error(MBX_E_UNSUPPORTED, "batched: the matrix-core adjoint does not fit %u antennas",
                             d->na.local_size);
        if (tune->samples_per_group <= 0)
            return set_error(MBX_E_DIMS, "batched: samples_per_group must be positive (got %d)",
                             tune->samples_per_group);
        const size_t need = batched_mx_ws_bytes<T>(*d, tune);
        if (!ws || ws_bytes < need)
            return set_error(MBX_E_WORKSPACE, "batched: workspace needs %zu bytes (got %zu)", need, ws_bytes);
        if (dext(d->ntime) == 0 || dext(d->nbl) == 0 || dext(d->nchan) == 0)
            return hip_status(hipGetLastError(), "batched (empty extents)");
        const uint64_t nwg = gemm_workgroups(*d), vstride = mx_vstride(*d);
        const unsigned G = unsigned(tune->samples_per_group);
        char *base = static_cast<char *>(ws);
        BeamCell<T> *beam = reinterpret_cast<BeamCell<T> *>(base);
        double *partial = reinterpret_cast<double *>(base + beam_planes_bytes<T>(*d));
        cx<T> *buf = reinterpret_cast<cx<T> *>(base + beam_planes_bytes<T>(*d) + mx_partial_bytes(*d, tune));
        cx<T> *vbar = buf + uint64_t(G) * vstride;
        launch_beam_planes<T>(*d, in->E_beam, in->frequency, b, beam, stream);

        RimeArgs<T> A;
        fill_args<T>(A, d, in, b, k, options);
        A.beam = beam;
        A.partial = partial;
        A.ekb = grads ? buf : nullptr;      // pre-gain sums kept for the adjoint
        A.span = SampleSpan{bt->src_offset, bt->npsrc, bt->ngsrc, bt->nssrc, bt->gauss_offset,
                            bt->sersic_offset, bt->order, unsigned(nwg), 0, vstride};
        const double scale = (options & MBX_APPLY_WEIGHTS) ? 1.0 : 1.0 / k->sigma_sqrd;
        AdjArgs<T> J;
        if (grads) {
            J.R = A;
            J.R.ekb = nullptr;
            J.R.partial = nullptr;
            J.ekb = buf;
            J.vbar = vbar;
            J.bmap = bmap;
            J.g_lm = gr->lm; J.g_stokes = gr->stokes; J.g_alpha = gr->alpha;
            J.g_gauss = gr->gauss_shape; J.g_sersic = gr->sersic_shape;
            J.g_G = nullptr;
            J.gscale = scale;
        }
        const bool any_point = dext(d->npsrc) > 0;
        for (int slot0 = 0; slot0 < bt->nsamples; slot0 += int(G)) {
            const unsigned nrep = unsigned(std::min<int>(int(G), bt->nsamples - slot0));
            A.span.slot0 = slot0;
            int r = launch_batched_mx(A, bmap, buf, any_point, nrep, stream);
            if (r) return r;
            hipLaunchKernelGGL(k_chi2_slots, dim3(nrep), dim3(256), 0, stream, partial, unsigned(nwg),
                               bt->order + slot0, scale, X2);
            if ((r = hip_status(hipGetLastError(), "batched chi2 launch"))) return r;
            if (grads) {
                J.R.span = A.span;
                if ((r = launch_adjoint_samples<T>(J, nrep, stream))) return r;
            }
        }
        return MBX_OK;
    }
}

template <typename T>
static size_t batched_ws_bytes(const mbx_dims &d, const mbx_batch_tuning *tune) {
    if (tune && tune->strategy == 1) return batched_mx_ws_bytes<T>(d, tune);
    BatPlan p;
    if (plan_batched<T>(d, tune, p)) return 0;
    size_t beam = beam_planes_bytes<T>(d);
    size_t slots = size_t(p.nwg) * d.nbl.local_size * p.ct * 4 * sizeof(cx<T>);
    return beam + slots + 256;
}

template <typename T>
static int batched(const mbx_dims *d, const mbx_inputs *in, const mbx_batch *bt, double *X2,
                   const mbx_grads *gr, const int32_t *bmap, const mbx_beam_extents *b,
                   const mbx_consts *k, const mbx_batch_tuning *tune, int options, void *ws,
                   size_t ws_bytes, hipStream_t stream) {
    int r = validate_inputs(d, in, b, k, options, "batched");
    if (r) return r;
    if (!bt || !X2) return set_error(MBX_E_NULL, "batched: NULL batch or X2");
    if (bt->nsamples <= 0) return set_error(MBX_E_DIMS, "batched: no samples");
    if (!bt->src_offset || !bt->npsrc || !bt->ngsrc || !bt->nssrc || !bt->gauss_offset ||
        !bt->sersic_offset || !bt->order)
        return set_error(MBX_E_NULL, "batched: NULL batch array");
    const bool grads = gr && (gr->lm || gr->stokes || gr->alpha || gr->gauss_shape || gr->sersic_shape);
    if (grads && !bmap) return set_error(MBX_E_NULL, "batched: gradients need the baseline map");
    if (gr && gr->G_term) return set_error(MBX_E_UNSUPPORTED, "batched: gain gradients are not produced");
    if (tune && tune->strategy != 0 && tune->strategy != 1)
        return set_error(MBX_E_UNSUPPORTED, "batched: unknown strategy %d", tune->strategy);
    const bool mx = tune && tune->strategy == 1;
    BatPlan p;
    if (!mx && (r = plan_batched<T>(*d, tune, p))) return r;
    // shapes the matrix strategy rejects: its own checks report them
    // before the workspace is sized (its size function returns 0 for them)
    if (mx && (sizeof(T) == 8 || !bmap || tune->samples_per_group <= 0))
        return batched_mx<T>(d, in, bt, X2, gr, grads, bmap, b, k, tune, options, ws, ws_bytes, stream);
    size_t need = batched_ws_bytes<T>(*d, tune);
    if (!ws || ws_bytes < need)
        return set_error(MBX_E_WORKSPACE, "batched: workspace needs %zu bytes (got %zu)", need, ws_bytes);
    const uint64_t NT = d->ntime.local_size;
    const uint64_t ns = d->nsrc.local_size;
    // zero X2 and the gradients (fill kernels: safe under stream capture);
    // the persistent launch's work counter sits in its workspace's last
    // 256 bytes
    hipLaunchKernelGGL(k_fill_batch, dim3(1), dim3(256), 0, stream, X2, uint64_t(bt->nsamples),
                       mx ? (int *)nullptr : reinterpret_cast<int *>(static_cast<char *>(ws) + need - 256));
    if (grads) {
        zero_doubles(gr->lm, ns * 2, stream);
        zero_doubles(gr->stokes, ns * NT * 4, stream);
        zero_doubles(gr->alpha, ns * NT, stream);
        zero_doubles(gr->gauss_shape, uint64_t(d->ngsrc.local_size) * 3, stream);
        zero_doubles(gr->sersic_shape, uint64_t(d->nssrc.local_size) * 3, stream);
    }
    if (mx) return batched_mx<T>(d, in, bt, X2, gr, grads, bmap, b, k, tune, options, ws, ws_bytes, stream);
    if (dext(d->ntime) == 0 || dext(d->nbl) == 0 || dext(d->nchan) == 0)
        return hip_status(hipGetLastError(), "batched (empty extents)");
    BeamCell<T> *beam = static_cast<BeamCell<T> *>(ws);
    const size_t beam_bytes = beam_planes_bytes<T>(*d);
    launch_beam_planes<T>(*d, in->E_beam, in->frequency, b, beam, stream);

    BatchArgs<T> B;
    fill_args<T>(B.R, d, in, b, k, options);
    B.R.beam = beam;
    B.R.ct = p.ct;
    B.R.nctile = p.nctile;
    B.b.src_offset = bt->src_offset; B.b.npsrc = bt->npsrc; B.b.ngsrc = bt->ngsrc;
    B.b.nssrc = bt->nssrc; B.b.gauss_offset = bt->gauss_offset; B.b.sersic_offset = bt->sersic_offset;
    B.b.order = bt->order; B.b.nsamples = bt->nsamples;
    B.X2 = X2;
    B.slots = reinterpret_cast<cx<T> *>(static_cast<char *>(ws) + beam_bytes);
    B.counter = reinterpret_cast<int *>(static_cast<char *>(ws) + need - 256);
    B.bmap = bmap;
    B.g_lm = grads ? gr->lm : nullptr; B.g_stokes = grads ? gr->stokes : nullptr;
    B.g_alpha = grads ? gr->alpha : nullptr; B.g_gauss = grads ? gr->gauss_shape : nullptr;
    B.g_sersic = grads ? gr->sersic_shape : nullptr;
    B.inv_sigma2 = (options & MBX_APPLY_WEIGHTS) ? 1.0 : 1.0 / k->sigma_sqrd;
    B.nitems = int(int64_t(bt->nsamples) * dext(d->ntime) * p.nctile);
    B.grads = grads;
    B.st_fwd = p.st_fwd;
    dim3 grid(p.nwg), block(kBatThreads);
    constexpr int RS = bat_rs<T>();
    switch (p.ni) {
    case 1: hipLaunchKernelGGL((k_batched<T, 1, RS>), grid, block, p.lds, stream, B); break;
    case 2: hipLaunchKernelGGL((k_batched<T, 2, RS>), grid, block, p.lds, stream, B); break;
    case 4: hipLaunchKernelGGL((k_batched<T, 4, RS>), grid, block, p.lds, stream, B); break;
    default: hipLaunchKernelGGL((k_batched<T, 8, RS>), grid, block, p.lds, stream, B); break;
    }
    return hip_status(hipGetLastError(), "batched launch");
}

}  // namespace mbx

using namespace mbx;

extern "C" {

size_t mbx_batched_workspace_bytes(const mbx_dims *d, int is_double, const mbx_batch_tuning *tune) {
    if (!d) return 0;
    return is_double ? batched_ws_bytes<double>(*d, tune) : batched_ws_bytes<float>(*d, tune);
}

int mbx_batched_f32(const mbx_dims *d, const mbx_inputs *in, const mbx_batch *batch, double *X2,
                    const mbx_grads *grads, const int32_t *baseline_map,
                    const mbx_beam_extents *beam, const mbx_consts *k,
                    const mbx_batch_tuning *tune, int options, void *ws, size_t ws_bytes, void *s) {
    return batched<float>(d, in, batch, X2, grads, baseline_map, beam, k, tune, options, ws,
                          ws_bytes, (hipStream_t)s);
}
int mbx_batched_f64(const mbx_dims *d, const mbx_inputs *in, const mbx_batch *batch, double *X2,
                    const mbx_grads *grads, const int32_t *baseline_map,
                    const mbx_beam_extents *beam, const mbx_consts *k,
                    const mbx_batch_tuning *tune, int options, void *ws, size_t ws_bytes, void *s) {
    return batched<double>(d, in, batch, X2, grads, baseline_map, beam, k, tune, options, ws,
                           ws_bytes, (hipStream_t)s);
}

}  // extern "C"
