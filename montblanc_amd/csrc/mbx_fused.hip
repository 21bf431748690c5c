// Fused forward RIME for gfx950.
//
// Replaces the four-node v4 pipeline (B_sqrt -> E_beam -> EKB_sqrt ->
// sum_coherencies -> chi^2 sum, v4/RimeSolver.py:36-40) with one kernel
// that never writes the per-source scratch arrays (jones alone is 42 GB
// in float64 at 64 antennas, 100 times, 64 channels and 1600 sources).
//
// A workgroup owns (time t, a tile of `ct` channels, a chunk of
// baselines). Per tile of `st` sources it stages sqrt(B) and the
// per-antenna Jones in LDS (mbx_tile.hpp) and then accumulates the
// coherency sums V_pq += f_s(pq) J_ps J_qs^H in registers:
//   k_forward_blk  pair-block kernel: each thread owns a PB x QB block of
//                  antenna pairs (from a host-built plan), so each Jones
//                  loaded from LDS feeds PB (or QB) products;
//   k_forward_gen  generic kernel for antenna layouts that change with
//                  time: each thread owns NI baselines.
// The epilogue applies gains, flags and chi^2 and leaves one float64
// partial per workgroup; a one-block kernel sums the partials in order.
//
// Source-batch semantics of the reference (RimeSumCoherencies.py:183-190,
// 307-316): with nsrc.lower_extent > 0 the running sum is read from
// ekb_vis; with nsrc.upper_extent < nsrc.global_size it is written there
// and the epilogue is skipped.

#include <algorithm>
#include <mutex>

#include "mbx_gemm.hpp"

namespace mbx {

void launch_chi2_final(const double *partial, unsigned n, double scale, double *X2, hipStream_t st);

constexpr int kGenThreads = 256;

// Workgroup index -> (time, baseline chunk, channel tile) in the XCD
// order of kChanGroup (mbx_tile.hpp): channel tile fastest over a group,
// then time, then baseline chunk, then channel group. False for the
// padding workgroups past the grid.
template <typename T>
__device__ __forceinline__ bool wg_coords(const RimeArgs<T> &A, int &bc, int &ctile, int &t) {
    uint64_t w = xcd_slot(A.per_xcd);
    const uint64_t NT = A.d.ntime.local_size;
    const int ci = int(w % kChanGroup);
    w /= kChanGroup;
    t = int(w % NT);
    w /= NT;
    bc = int(w % A.nbchunk);
    ctile = int(w / A.nbchunk) * kChanGroup + ci;
    return ctile < A.nctile;
}

// ---------------------------------------------------------------------
// Generic kernel: NI baselines per thread, any antenna layout.

template <typename T, int NI>
__global__ void __launch_bounds__(kGenThreads)
k_forward_gen(RimeArgs<T> A) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const mbx_dims &d = A.d;
    const int NA = d.na.local_size;
    const uint64_t NC = d.nchan.local_size, NBL = d.nbl.local_size;
    const int ct = A.ct, st = A.st;
    const int tid = threadIdx.x;
    int bc, ctile, t;
    if (!wg_coords(A, bc, ctile, t)) {       // padding workgroup
        if (threadIdx.x == 0) A.partial[blockIdx.x] = 0.0;
        return;
    }
    Tile<T> L(smem_raw, NA, ct, st);
    const int nt_ext = dext(d.ntime), nc_ext = dext(d.nchan), nb_ext = dext(d.nbl);
    const bool active = t < nt_ext;
    const int c0 = ctile * ct;
    const int b0 = bc * A.bchunk;
    const int b1 = min(nb_ext, b0 + A.bchunk);
    const int nitems = max(0, b1 - b0) * ct;
    if (active) tile_setup(A, L, t, c0, tid, kGenThreads);

    J4<T> acc[NI];
    int ip[NI], iq[NI], icl[NI];
    T iu[NI], iv[NI];
#pragma unroll
    for (int k = 0; k < NI; ++k) {
        int it = tid + k * kGenThreads;
        bool ok = active && it < nitems;
        int b = b0 + (ok ? it / ct : 0);
        icl[k] = ok ? it % ct : 0;
        ip[k] = 0; iq[k] = 0;
        if (active && b < nb_ext) {
            // antenna indices are data: clamp so a bad index cannot read
            // outside the per-antenna arrays
            ip[k] = min(max(A.a1[uint64_t(t) * NBL + b], 0), NA - 1);
            iq[k] = min(max(A.a2[uint64_t(t) * NBL + b], 0), NA - 1);
        }
        acc[k] = jzero<T>();
        if (ok && A.vin && c0 + icl[k] < nc_ext)
            acc[k] = jload(A.vin + ((uint64_t(t) * NBL + b) * NC + c0 + icl[k]) * 4);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NI; ++k) {
        iu[k] = T(L.uvw[3 * iq[k]] - L.uvw[3 * ip[k]]);
        iv[k] = T(L.uvw[3 * iq[k] + 1] - L.uvw[3 * ip[k] + 1]);
    }
    const int ns_all = dext(d.nsrc);
    for (int s0 = A.s_begin; s0 < ns_all; s0 += st) {
        const int nst = min(st, ns_all - s0);
        if (active) tile_phase0(A, L, t, c0, s0, nst, tid, kGenThreads);
        __syncthreads();
        if (active) tile_phase1(A, L, t, c0, nst, tid, kGenThreads);
        __syncthreads();
        if (active) {
            for (int si = 0; si < nst; ++si) {
                const J4<T> *Js = L.J + si * NA * ct;
                const T ty = L.shape[4 * si + 3];
                if (ty == T(0)) {
#pragma unroll
                    for (int k = 0; k < NI; ++k)
                        if (tid + k * kGenThreads < nitems)
                            jacc_h(acc[k], Js[ip[k] * ct + icl[k]], Js[iq[k] * ct + icl[k]]);
                } else {
                    const T p0 = L.shape[4 * si], p1 = L.shape[4 * si + 1], p2 = L.shape[4 * si + 2];
                    const bool gauss = ty == T(1);
                    const T kscale = T(gauss ? A.gauss_scale : A.two_pi_over_c);
#pragma unroll
                    for (int k = 0; k < NI; ++k)
                        if (tid + k * kGenThreads < nitems) {
                            T kf = kscale * T(L.ch[CH * icl[k] + 3]);
                            T f = gauss ? gauss_f(iu[k], iv[k], p0, p1, p2, kf)
                                        : sersic_f(iu[k], iv[k], p0, p1, p2, kf);
                            jacc_h(acc[k], jscale(Js[ip[k] * ct + icl[k]], f), Js[iq[k] * ct + icl[k]]);
                        }
                }
            }
        }
        __syncthreads();
    }
    const bool final_batch = d.nsrc.upper_extent >= d.nsrc.global_size;
    double term_sum = 0.0;
    if (active) {
#pragma unroll
        for (int k = 0; k < NI; ++k) {
            int it = tid + k * kGenThreads;
            if (it >= nitems || c0 + icl[k] >= nc_ext) continue;
            term_sum += vis_out(A, t, b0 + it / ct, c0 + icl[k], ip[k], iq[k], acc[k], final_batch);
        }
    }
    wg_partial(term_sum, L.part, A.partial + blockIdx.x, tid, kGenThreads);
}

// ---------------------------------------------------------------------
// Pair-block kernel. Plan entry of a block: p0, q0, then PB*QB baseline
// indices (slot i*QB + j is pair (p0 + i, q0 + j); -1 where no baseline).

// Workgroup size limit of the pair-block kernel: 768 threads leave 168
// VGPRs per lane (float), 512 threads leave 256 (double, whose 2x2 Jones
// accumulators take twice the registers).
template <typename T> constexpr int blk_max_threads() { return sizeof(T) == 4 ? 768 : 512; }

template <typename T, int PB, int QB>
__global__ void __launch_bounds__(blk_max_threads<T>())
k_forward_blk(RimeArgs<T> A) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const mbx_dims &d = A.d;
    const int NA = d.na.local_size;
    const uint64_t NC = d.nchan.local_size, NBL = d.nbl.local_size;
    const int ct = A.ct, st = A.st;
    const int tid = threadIdx.x, nth = blockDim.x;
    int bc, ctile, t;
    if (!wg_coords(A, bc, ctile, t)) {       // padding workgroup
        if (threadIdx.x == 0) A.partial[blockIdx.x] = 0.0;
        return;
    }
    Tile<T> L(smem_raw, NA, ct, st);
    const int nt_ext = dext(d.ntime), nc_ext = dext(d.nchan);
    const bool active = t < nt_ext;
    const int c0 = ctile * ct;
    const int blk_lo = bc * A.bchunk;
    const int nbk = max(0, min(A.bchunk, A.nblocks - blk_lo));
    const bool mine = active && tid < nbk * ct;
    const int blk = blk_lo + (mine ? tid / ct : 0);
    const int k = mine ? tid % ct : 0;
    const bool chan_ok = c0 + k < nc_ext;
    if (active) tile_setup(A, L, t, c0, tid, nth);

    constexpr int W = 2 + PB * QB;
    const int32_t *plan = A.blocks + uint64_t(min(blk, A.nblocks - 1)) * W;
    int pi[PB], qj[QB];
    const int p0 = plan[0], q0 = plan[1];
#pragma unroll
    for (int i = 0; i < PB; ++i) pi[i] = min(max(p0 + i, 0), NA - 1);
#pragma unroll
    for (int j = 0; j < QB; ++j) qj[j] = min(max(q0 + j, 0), NA - 1);

    J4<T> acc[PB][QB];
#pragma unroll
    for (int i = 0; i < PB; ++i)
#pragma unroll
        for (int j = 0; j < QB; ++j) {
            acc[i][j] = jzero<T>();
            const int b = mine ? plan[2 + i * QB + j] : -1;
            if (b >= 0 && chan_ok && A.vin)
                acc[i][j] = jload(A.vin + ((uint64_t(t) * NBL + b) * NC + c0 + k) * 4);
        }
    __syncthreads();
    T up[PB], vp[PB], uq[QB], vq[QB];
#pragma unroll
    for (int i = 0; i < PB; ++i) { up[i] = T(L.uvw[3 * pi[i]]); vp[i] = T(L.uvw[3 * pi[i] + 1]); }
#pragma unroll
    for (int j = 0; j < QB; ++j) { uq[j] = T(L.uvw[3 * qj[j]]); vq[j] = T(L.uvw[3 * qj[j] + 1]); }

    const int ns_all = dext(d.nsrc);
    for (int s0 = A.s_begin; s0 < ns_all; s0 += st) {
        const int nst = min(st, ns_all - s0);
        if (active) tile_phase0(A, L, t, c0, s0, nst, tid, nth);
        __syncthreads();
        if (active) tile_phase1(A, L, t, c0, nst, tid, nth);
        __syncthreads();
        if (mine) {
            const T kf_g = T(A.gauss_scale) * T(L.ch[CH * k + 3]);
            const T kf_s = T(A.two_pi_over_c) * T(L.ch[CH * k + 3]);
            for (int si = 0; si < nst; ++si) {
                const J4<T> *Js = L.J + si * NA * ct + k;
                const T ty = L.shape[4 * si + 3];
                J4<T> Jp[PB];
#pragma unroll
                for (int i = 0; i < PB; ++i) Jp[i] = Js[pi[i] * ct];
                if (ty == T(0)) {
#pragma unroll
                    for (int j = 0; j < QB; ++j) {
                        const J4<T> Jq = Js[qj[j] * ct];
#pragma unroll
                        for (int i = 0; i < PB; ++i) jacc_h(acc[i][j], Jp[i], Jq);
                    }
                } else {
                    const T s0p = L.shape[4 * si], s1p = L.shape[4 * si + 1], s2p = L.shape[4 * si + 2];
                    const bool gauss = ty == T(1);
#pragma unroll
                    for (int j = 0; j < QB; ++j) {
                        const J4<T> Jq = Js[qj[j] * ct];
#pragma unroll
                        for (int i = 0; i < PB; ++i) {
                            const T u = uq[j] - up[i], v = vq[j] - vp[i];
                            const T f = gauss ? gauss_f(u, v, s0p, s1p, s2p, kf_g)
                                              : sersic_f(u, v, s0p, s1p, s2p, kf_s);
                            jacc_h(acc[i][j], jscale(Jp[i], f), Jq);
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
    const bool final_batch = d.nsrc.upper_extent >= d.nsrc.global_size;
    double term_sum = 0.0;
    if (mine && chan_ok) {
#pragma unroll
        for (int i = 0; i < PB; ++i)
#pragma unroll
            for (int j = 0; j < QB; ++j) {
                const int b = plan[2 + i * QB + j];
                if (b >= 0) term_sum += vis_out(A, t, b, c0 + k, pi[i], qj[j], acc[i][j], final_batch);
            }
    }
    wg_partial(term_sum, L.part, A.partial + blockIdx.x, tid, nth);
}

// ---------------------------------------------------------------------
// Double-buffered pair-block kernel. Waves [0, cth/64) are consumers (one
// pair block and channel each, accumulators in registers); the remaining
// waves are producers that build the Jones tile of the next source tile
// in the other LDS buffer while the consumers work on the current one. One
// barrier per source tile; neither role holds the other's registers.

template <typename T> constexpr int pc_max_threads() { return sizeof(T) == 4 ? 1024 : 768; }

template <typename T, int PB, int QB>
__device__ __forceinline__ void pc_consume(const RimeArgs<T> &A, const PcTile<T> &L, int buf,
                                           int nst, int k, const int *pi, const int *qj,
                                           const T *up, const T *vp, const T *uq, const T *vq,
                                           J4<T> (&acc)[PB][QB]) {
    const int NA = A.d.na.local_size;
    const int ct = A.ct;
    const J4<T> *Jb = L.J + buf * L.jstride + k;
    const T *shape = L.shape + buf * 4 * A.st;
    const T kf_g = T(A.gauss_scale) * T(L.ch[CH * k + 3]);
    const T kf_s = T(A.two_pi_over_c) * T(L.ch[CH * k + 3]);
    for (int si = 0; si < nst; ++si) {
        const J4<T> *Js = Jb + si * NA * ct;
        const T ty = shape[4 * si + 3];
        J4<T> Jp[PB];
#pragma unroll
        for (int i = 0; i < PB; ++i) Jp[i] = Js[pi[i] * ct];
        if (ty == T(0)) {
#pragma unroll
            for (int j = 0; j < QB; ++j) {
                const J4<T> Jq = Js[qj[j] * ct];
#pragma unroll
                for (int i = 0; i < PB; ++i) jacc_h(acc[i][j], Jp[i], Jq);
            }
        } else {
            const T s0p = shape[4 * si], s1p = shape[4 * si + 1], s2p = shape[4 * si + 2];
            // the source's shape coordinates, linear in the pair's (u, v):
            // x = u c1 + v c2, y = u c3 + v c4 (gaussian: k = gauss_scale f,
            // float pre-scaled by sqrt(log2 e) for the hardware exp2;
            // sersic: k = 2 pi f / C times R')
            T c1, c2, c3, c4;
            if (ty == T(1)) {
                const T kg = kf_g * kGaussExp2Scale<T>(), kR = kg * s2p;
                c1 = s1p * kR; c2 = -s0p * kR; c3 = s0p * kg; c4 = s1p * kg;
            } else {
                const T kR = kf_s * s2p;
                c1 = (T(1) + s0p) * kR; c2 = s1p * kR; c3 = c2; c4 = (T(1) - s0p) * kR;
            }
            T f[PB][QB];
            auto r2_of = [&](int i, int j) {
                const T u = uq[j] - up[i], v = vq[j] - vp[i];
                const T x = fma(u, c1, v * c2), y = fma(u, c3, v * c4);
                return fma(x, x, y * y);
            };
            if (ty == T(1)) {
#pragma unroll
                for (int i = 0; i < PB; ++i)
#pragma unroll
                    for (int j = 0; j < QB; ++j) f[i][j] = gauss_of_scaled(r2_of(i, j));
            } else {
#pragma unroll
                for (int i = 0; i < PB; ++i)
#pragma unroll
                    for (int j = 0; j < QB; ++j) f[i][j] = sersic_of(r2_of(i, j));
            }
#pragma unroll
            for (int j = 0; j < QB; ++j) {
                const J4<T> Jq = Js[qj[j] * ct];
#pragma unroll
                for (int i = 0; i < PB; ++i) jacc_h(acc[i][j], jscale(Jp[i], f[i][j]), Jq);
            }
        }
    }
}

template <typename T, int PB, int QB>
__global__ void __launch_bounds__(pc_max_threads<T>())
k_forward_pc(RimeArgs<T> A, int cth) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const mbx_dims &d = A.d;
    const int NA = d.na.local_size;
    const int ct = A.ct, st = A.st;
    const int tid = threadIdx.x, nth = blockDim.x;
    int bc, ctile, t;
    if (!wg_coords(A, bc, ctile, t)) {       // padding workgroup
        if (threadIdx.x == 0) A.partial[blockIdx.x] = 0.0;
        return;
    }
    PcTile<T> L(smem_raw, NA, ct, st);
    const bool active = t < int(dext(d.ntime));
    const int c0 = ctile * ct;
    if (active) tile_setup<T>(A, L.uvw, L.pa, L.ch, t, c0, tid, nth);
    __syncthreads();
    const int ns_all = dext(d.nsrc), sb = A.s_begin;
    const int ntiles = (ns_all - sb + st - 1) / st;
    // The role is a per-wave (scalar) decision, so each wave runs exactly
    // one of the two loops below; both execute 1 + ntiles barriers.
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    double term_sum = 0.0;
    if (wave < (cth >> 6)) {
        const uint64_t NC = d.nchan.local_size, NBL = d.nbl.local_size;
        const int nc_ext = dext(d.nchan);
        const int blk_lo = bc * A.bchunk;
        const int nbk = max(0, min(A.bchunk, A.nblocks - blk_lo));
        const bool mine = active && tid < nbk * ct;
        const int blk = blk_lo + (mine ? tid / ct : 0);
        const int k = mine ? tid % ct : 0;
        const bool chan_ok = c0 + k < nc_ext;
        constexpr int W = 2 + PB * QB;
        const int32_t *plan = A.blocks + uint64_t(min(blk, A.nblocks - 1)) * W;
        int pi[PB], qj[QB];
        const int p0 = plan[0], q0 = plan[1];
#pragma unroll
        for (int i = 0; i < PB; ++i) pi[i] = min(max(p0 + i, 0), NA - 1);
#pragma unroll
        for (int j = 0; j < QB; ++j) qj[j] = min(max(q0 + j, 0), NA - 1);
        J4<T> acc[PB][QB];
#pragma unroll
        for (int i = 0; i < PB; ++i)
#pragma unroll
            for (int j = 0; j < QB; ++j) {
                acc[i][j] = jzero<T>();
                const int b = mine ? plan[2 + i * QB + j] : -1;
                if (b >= 0 && chan_ok && A.vin)
                    acc[i][j] = jload(A.vin + ((uint64_t(t) * NBL + b) * NC + c0 + k) * 4);
            }
        T up[PB], vp[PB], uq[QB], vq[QB];
#pragma unroll
        for (int i = 0; i < PB; ++i) { up[i] = T(L.uvw[3 * pi[i]]); vp[i] = T(L.uvw[3 * pi[i] + 1]); }
#pragma unroll
        for (int j = 0; j < QB; ++j) { uq[j] = T(L.uvw[3 * qj[j]]); vq[j] = T(L.uvw[3 * qj[j] + 1]); }
        __syncthreads();                       // tile 0 produced
        for (int it = 0; it < ntiles; ++it) {
            const int nst = min(st, ns_all - sb - it * st);
            if (mine) pc_consume<T, PB, QB>(A, L, it & 1, nst, k, pi, qj, up, vp, uq, vq, acc);
            __syncthreads();
        }
        const bool final_batch = d.nsrc.upper_extent >= d.nsrc.global_size;
        if (mine && chan_ok) {
#pragma unroll
            for (int i = 0; i < PB; ++i)
#pragma unroll
                for (int j = 0; j < QB; ++j) {
                    const int b = plan[2 + i * QB + j];
                    if (b >= 0)
                        term_sum += vis_out(A, t, b, c0 + k, pi[i], qj[j], acc[i][j], final_batch);
                    // one visibility at a time: keeps the epilogues' registers apart
                    __builtin_amdgcn_sched_barrier(0);
                }
        }
    } else {
        const int ptid = tid - cth, pth = nth - cth;
        if (active && ntiles > 0) pc_produce<T>(A, L, 0, t, c0, sb, min(st, ns_all - sb), ptid, pth);
        __syncthreads();
        for (int it = 0; it < ntiles; ++it) {
            const int s1 = sb + (it + 1) * st;
            if (active && it + 1 < ntiles)
                pc_produce<T>(A, L, (it + 1) & 1, t, c0, s1, min(st, ns_all - s1), ptid, pth);
            __syncthreads();
        }
    }
    wg_partial(term_sum, L.part, A.partial + blockIdx.x, tid, nth);
}

// ---------------------------------------------------------------------
// Launch planning.

struct FwdPlan {
    bool blocked, pc;
    int ct, st, ni, nbchunk, bchunk, nctile, threads, cth;
    uint64_t nwg;
    size_t lds_bytes;
};

constexpr size_t kLdsBudget = 64 * 1024;

template <typename T>
static int choose_st(int na, int ct, int nwaves) {
    int st = 32;
    while (st > 1 && tile_lds_bytes<T>(na, ct, st, nwaves) > kLdsBudget) st >>= 1;
    return st;
}

template <typename T>
static FwdPlan plan_forward(const mbx_dims &d, const mbx_pair_plan *pp) {
    FwdPlan p;
    const int na = d.na.local_size;
    const int nb = d.nbl.local_size;
    p.pc = false;
    p.cth = 0;
    if (pp && pp->blocks && pp->nblocks > 0 && pp->producer_waves >= 0) {
        // double-buffered producer / consumer kernel
        p.blocked = true;
        p.pc = true;
        p.ni = 0;
        const int maxth = pc_max_threads<T>();
        int pw = pp->producer_waves;
        const int max_cons = maxth - 64 * std::max(pw, 1);
        p.nbchunk = (pp->nblocks + max_cons - 1) / max_cons;
        p.bchunk = (pp->nblocks + p.nbchunk - 1) / p.nbchunk;
        int ct = (256 + p.bchunk - 1) / p.bchunk;
        ct = std::max(1, std::min(ct, 8));
        while (ct > 1 && p.bchunk * ct > max_cons) --ct;
        p.ct = std::min<int>(ct, d.nchan.local_size);
        p.cth = ((p.bchunk * p.ct + 63) / 64) * 64;
        // automatic: every wave the workgroup limit leaves, at most one
        // producer per consumer wave (the beam gathers are latency bound)
        if (pw == 0) pw = std::max(1, std::min((maxth - p.cth) / 64, p.cth / 64));
        p.threads = p.cth + 64 * pw;
        p.nctile = (d.nchan.local_size + p.ct - 1) / p.ct;
        int st = 32;
        while (st > 1 && pc_lds_bytes<T>(na, p.ct, st, p.threads / 64) > kLdsBudget) st >>= 1;
        p.st = st;
        p.lds_bytes = pc_lds_bytes<T>(na, p.ct, st, p.threads / 64);
        p.nwg = 8 * uint64_t(xcd_slots(uint64_t(d.ntime.local_size) * p.nbchunk *
                                        ((p.nctile + kChanGroup - 1) / kChanGroup) * kChanGroup));
        return p;
    }
    if (pp && pp->blocks && pp->nblocks > 0) {
        p.blocked = true;
        p.ni = 0;
        const int maxth = blk_max_threads<T>();
        p.nbchunk = (pp->nblocks + maxth - 1) / maxth;
        p.bchunk = (pp->nblocks + p.nbchunk - 1) / p.nbchunk;
        int ct = (256 + p.bchunk - 1) / p.bchunk;
        ct = std::max(1, std::min(ct, 8));
        while (ct > 1 && p.bchunk * ct > maxth) --ct;
        p.ct = std::min<int>(ct, d.nchan.local_size);
        p.threads = ((p.bchunk * p.ct + 63) / 64) * 64;
    } else {
        p.blocked = false;
        p.ct = 1;
        int need = (nb * p.ct + kGenThreads - 1) / kGenThreads;
        p.ni = need <= 1 ? 1 : need <= 2 ? 2 : need <= 4 ? 4 : 8;
        int per_chunk = p.ni * kGenThreads / p.ct;
        p.nbchunk = (nb + per_chunk - 1) / per_chunk;
        p.bchunk = (nb + p.nbchunk - 1) / p.nbchunk;
        p.threads = kGenThreads;
    }
    p.nctile = (d.nchan.local_size + p.ct - 1) / p.ct;
    p.st = choose_st<T>(na, p.ct, p.threads / 64);
    p.lds_bytes = tile_lds_bytes<T>(na, p.ct, p.st, p.threads / 64);
    p.nwg = 8 * uint64_t(xcd_slots(uint64_t(d.ntime.local_size) * p.nbchunk *
                                    ((p.nctile + kChanGroup - 1) / kChanGroup) * kChanGroup));
    return p;
}

// The epilogue alone, once the matrix-core passes have summed every
// source into A.vin: M = G_p V G_q^H, flags, model / residual and chi^2
// terms per visibility (vis_out), one thread per visibility of a time
// step (grid y strides over time, x over baselines x channels), one chi^2
// partial per workgroup in the grid's row-major order.
template <typename T>
__global__ void __launch_bounds__(kGenThreads) k_epilogue(RimeArgs<T> A) {
    __shared__ double part[kGenThreads / 64];
    const mbx_dims &d = A.d;
    const int NA = d.na.local_size;
    const uint64_t NBL = d.nbl.local_size, NC = d.nchan.local_size;
    const uint32_t nb = uint32_t(dext(d.nbl)), nc = uint32_t(dext(d.nchan));
    const uint32_t n = nb * nc, nt = uint32_t(dext(d.ntime));
    const bool final_batch = d.nsrc.upper_extent >= d.nsrc.global_size;
    double term = 0.0;
    for (uint32_t t = blockIdx.y; t < nt; t += gridDim.y)
    for (uint32_t i = blockIdx.x * kGenThreads + threadIdx.x; i < n; i += gridDim.x * kGenThreads) {
        const uint32_t b = i / nc, c = i - b * nc;
        const int p = min(max(A.a1[uint64_t(t) * NBL + b], 0), NA - 1);
        const int q = min(max(A.a2[uint64_t(t) * NBL + b], 0), NA - 1);
        const uint64_t vi = (uint64_t(t) * NBL + b) * NC + c;
        const J4<T> V = jload(A.vin + vi * 4);
        term += vis_out(A, int(t), int(b), int(c), p, q, V, final_batch);
    }
    wg_partial(term, part, A.partial + (uint64_t(blockIdx.y) * gridDim.x + blockIdx.x), threadIdx.x,
               kGenThreads);
}

// Upper bound on workgroups of any forward kernel (one float64 partial
// each): time x (channel tiles padded to kChanGroup) x baseline chunks
// (at least 64 pair blocks or baselines per chunk), padded to 8 XCDs.
static uint64_t max_partials(const mbx_dims &d) {
    const uint64_t nb = d.nbl.local_size;
    return uint64_t(d.ntime.local_size) * (d.nchan.local_size + kChanGroup) * ((nb + 63) / 64 + 1) + 8;
}

template <typename T>
static size_t beam_ws_bytes(const mbx_dims &d) {
    return beam_planes_bytes<T>(d);
}

template <typename T>
static size_t forward_ws_bytes(const mbx_dims &d) {
    return beam_ws_bytes<T>(d) + max_partials(d) * sizeof(double);
}

int validate_inputs(const mbx_dims *d, const mbx_inputs *in, const mbx_beam_extents *b,
                    const mbx_consts *k, int options, const char *who) {
    int r = check_dims(d, true);
    if (r) return r;
    if (!in || !b || !k) return set_error(MBX_E_NULL, "%s: NULL argument struct", who);
    if (!in->uvw || !in->frequency || !in->ref_frequency || !in->parallactic_angles ||
        !in->point_errors || !in->antenna_scaling || !in->E_beam || !in->antenna1 ||
        !in->antenna2 || !in->flag || !in->observed_vis || !in->G_term)
        return set_error(MBX_E_NULL, "%s: NULL input array", who);
    if (dext(d->nsrc) > 0 && (!in->lm || !in->stokes || !in->alpha))
        return set_error(MBX_E_NULL, "%s: NULL source array", who);
    if (dext(d->ngsrc) > 0 && !in->gauss_shape) return set_error(MBX_E_NULL, "%s: gauss_shape is NULL", who);
    if (dext(d->nssrc) > 0 && !in->sersic_shape) return set_error(MBX_E_NULL, "%s: sersic_shape is NULL", who);
    if ((options & MBX_APPLY_WEIGHTS) && !in->weight_vector)
        return set_error(MBX_E_NULL, "%s: weight_vector is NULL", who);
    if (options & ~(MBX_APPLY_WEIGHTS | MBX_RESIDUAL_OUTPUT))
        return set_error(MBX_E_UNSUPPORTED, "%s: unknown option bits 0x%x", who, options);
    if (!(b->ul != b->ll && b->um != b->lm && b->ufreq != b->lfreq))
        return set_error(MBX_E_DIMS, "%s: degenerate beam extents", who);
    return MBX_OK;
}

int validate_plan(const mbx_dims *d, const mbx_pair_plan *pp, bool is_double, const char *who) {
    if (!pp || !pp->blocks || pp->nblocks <= 0) return MBX_OK;
    (void)is_double;
    const bool ok = (pp->pb == 2 && pp->qb == 2) || (pp->pb == 2 && pp->qb == 4) ||
                    (pp->pb == 1 && pp->qb == 1);
    if (!ok) return set_error(MBX_E_UNSUPPORTED, "%s: pair blocks %dx%d not available", who, pp->pb, pp->qb);
    if (pp->nblocks > int64_t(d->nbl.local_size))
        return set_error(MBX_E_DIMS, "%s: %d pair blocks for %u baselines", who, pp->nblocks,
                         d->nbl.local_size);
    return MBX_OK;
}

template <typename T>
static void launch_blocked(const FwdPlan &p, const mbx_pair_plan *pp, const RimeArgs<T> &A,
                           hipStream_t stream) {
    dim3 grid((unsigned)p.nwg), block(p.threads);
    if (p.pc) {
        if (pp->pb == 2 && pp->qb == 2)
            hipLaunchKernelGGL((k_forward_pc<T, 2, 2>), grid, block, p.lds_bytes, stream, A, p.cth);
        else if (pp->pb == 2 && pp->qb == 4)
            hipLaunchKernelGGL((k_forward_pc<T, 2, 4>), grid, block, p.lds_bytes, stream, A, p.cth);
        else
            hipLaunchKernelGGL((k_forward_pc<T, 1, 1>), grid, block, p.lds_bytes, stream, A, p.cth);
        return;
    }
    if (pp->pb == 2 && pp->qb == 2)
        hipLaunchKernelGGL((k_forward_blk<T, 2, 2>), grid, block, p.lds_bytes, stream, A);
    else if (pp->pb == 2 && pp->qb == 4)
        hipLaunchKernelGGL((k_forward_blk<T, 2, 4>), grid, block, p.lds_bytes, stream, A);
    else
        hipLaunchKernelGGL((k_forward_blk<T, 1, 1>), grid, block, p.lds_bytes, stream, A);
}

template <typename T>
static int forward(const mbx_dims *d, const mbx_inputs *in, const mbx_outputs *out,
                   const mbx_beam_extents *b, const mbx_consts *k, const mbx_pair_plan *pp,
                   int options, void *ws, size_t ws_bytes, hipStream_t stream) {
    int r = validate_inputs(d, in, b, k, options, "forward");
    if (r) return r;
    if ((r = validate_plan(d, pp, sizeof(T) == 8, "forward"))) return r;
    if (!out) return set_error(MBX_E_NULL, "forward: NULL outputs");
    const bool final_batch = d->nsrc.upper_extent >= d->nsrc.global_size;
    if (final_batch && (!out->model_vis || !out->chi_sqrd_result || !out->X2))
        return set_error(MBX_E_NULL, "forward: NULL output array");
    if ((!final_batch || d->nsrc.lower_extent > 0) && !out->ekb_vis)
        return set_error(MBX_E_NULL, "forward: a partial source batch needs ekb_vis");
    size_t need = forward_ws_bytes<T>(*d);
    if (!ws || ws_bytes < need)
        return set_error(MBX_E_WORKSPACE, "forward: workspace needs %zu bytes (got %zu)", need, ws_bytes);

    FwdPlan p = plan_forward<T>(*d, pp);
    BeamCell<T> *beam = static_cast<BeamCell<T> *>(ws);
    double *partial = reinterpret_cast<double *>(static_cast<char *>(ws) + beam_ws_bytes<T>(*d));
    const double scale = (options & MBX_APPLY_WEIGHTS) ? 1.0 : 1.0 / k->sigma_sqrd;
    if (dext(d->ntime) == 0 || dext(d->nbl) == 0 || dext(d->nchan) == 0) {
        if (final_batch) launch_chi2_final(partial, 0, scale, out->X2, stream);
        return hip_status(hipGetLastError(), "forward (empty extents)");
    }
    launch_beam_planes<T>(*d, in->E_beam, in->frequency, b, beam, stream);

    RimeArgs<T> A;
    fill_args(A, d, in, b, k, options);
    A.beam = beam;
    A.model = (cx<T> *)out->model_vis; A.chi2 = (T *)out->chi_sqrd_result;
    A.ekb = (cx<T> *)out->ekb_vis; A.partial = partial;
    // running sum of earlier source batches (RimeSumCoherencies.py:183-190)
    A.vin = d->nsrc.lower_extent > 0 ? A.ekb : nullptr;
    const bool shaped = shaped_gemm_pays<T>(*d);
    if (pp && pp->baseline_map && pp->point_gemm && point_gemm_supported<T>(*d) &&
        (dext(d->npsrc) > 0 || (shaped && dext(d->nsrc) > dext(d->npsrc)))) {
        // point sources (and where it pays the gaussian and sersic ones)
        // on the matrix cores into the running-sum buffer; the kernel
        // below adds the remaining sources and applies gains, flags, chi^2
        // -- unless the float shaped pass, the last one, applied them
        cx<T> *vbuf = A.ekb ? A.ekb : A.model;
        if (gemm_workgroups(*d) > max_partials(*d))
            return set_error(MBX_E_UNSUPPORTED, "forward: %llu workgroups exceed the partial buffer",
                             (unsigned long long)gemm_workgroups(*d));
        uint64_t epi_wgs = 0;
        r = launch_point_gemm<T>(A, pp->baseline_map, vbuf, A.vin, shaped, stream, &epi_wgs);
        if (r) return r;
        if (epi_wgs) {
            if (final_batch) {
                launch_chi2_final(partial, unsigned(epi_wgs), scale, out->X2, stream);
                r = hip_status(hipGetLastError(), "forward chi2 launch");
            }
            return r;
        }
        A.vin = vbuf;
        A.s_begin = shaped ? dext(d->nsrc) : dext(d->npsrc);
        if (A.s_begin >= dext(d->nsrc)) {
            // every source summed: the epilogue alone, about four
            // visibilities per thread
            const uint64_t nbc = uint64_t(dext(d->nbl)) * dext(d->nchan);
            const uint64_t gx = std::max<uint64_t>(1, std::min<uint64_t>((nbc + 4 * kGenThreads - 1) /
                                                                          (4 * kGenThreads), 1024));
            const uint64_t gy = std::min<uint64_t>(dext(d->ntime), 65535);
            const uint64_t nwg = gx * gy;
            if (nwg > max_partials(*d) || nbc >= (uint64_t(1) << 32))
                return set_error(MBX_E_UNSUPPORTED, "forward epilogue: %llu workgroups",
                                 (unsigned long long)nwg);
            hipLaunchKernelGGL((k_epilogue<T>), dim3(unsigned(gx), unsigned(gy)),
                               dim3(kGenThreads), 0, stream, A);
            r = hip_status(hipGetLastError(), "forward epilogue launch");
            if (r) return r;
            if (final_batch) {
                launch_chi2_final(partial, unsigned(nwg), scale, out->X2, stream);
                r = hip_status(hipGetLastError(), "forward chi2 launch");
            }
            return r;
        }
    }
    A.ct = p.ct; A.st = p.st; A.nbchunk = p.nbchunk; A.bchunk = p.bchunk; A.nctile = p.nctile;
    A.per_xcd = uint32_t(p.nwg / 8);
    if (p.nwg > max_partials(*d))
        return set_error(MBX_E_UNSUPPORTED, "forward: %llu workgroups exceed the partial buffer",
                         (unsigned long long)p.nwg);
    if (p.blocked) {
        A.blocks = pp->blocks; A.nblocks = pp->nblocks; A.pb = pp->pb; A.qb = pp->qb;
        launch_blocked<T>(p, pp, A, stream);
    } else {
        dim3 grid((unsigned)p.nwg), block(kGenThreads);
        switch (p.ni) {
        case 1: hipLaunchKernelGGL((k_forward_gen<T, 1>), grid, block, p.lds_bytes, stream, A); break;
        case 2: hipLaunchKernelGGL((k_forward_gen<T, 2>), grid, block, p.lds_bytes, stream, A); break;
        case 4: hipLaunchKernelGGL((k_forward_gen<T, 4>), grid, block, p.lds_bytes, stream, A); break;
        default: hipLaunchKernelGGL((k_forward_gen<T, 8>), grid, block, p.lds_bytes, stream, A); break;
        }
    }
    r = hip_status(hipGetLastError(), "forward launch");
    if (r) return r;
    if (final_batch) {
        launch_chi2_final(partial, (unsigned)p.nwg, scale, out->X2, stream);
        r = hip_status(hipGetLastError(), "forward chi2 launch");
    }
    return r;
}

// ---- forward with a library-held workspace (no allocation) -------------
//
// The workspace of mbx_forward_static_* (beam planes, then the chi^2
// partials) lives in this file-scope device array: each device that runs
// the library holds its own copy until the process ends. Host code never
// passes the array's name to a kernel (on the host it is only a stand-in
// for the symbol): the device address comes from hipGetSymbolAddress, once
// per device with that device current, and is kept per device. Calls that
// overlap on one device share the array and race; every byte the forward
// reads from it is written earlier in the same call (beam planes by
// k_beam_planes, each partial by its own workgroup), so nothing needs
// clearing between calls.

static __device__ __align__(256) unsigned char scratch[MBX_STATIC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];

static int scratch_on(hipStream_t stream, void **out) {
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return hip_status(e, "forward_static: hipGetDevice");
    hipDevice_t sdev = cur;
    if (stream) {
        e = hipStreamGetDevice(stream, &sdev);
        if (e != hipSuccess) return hip_status(e, "forward_static: hipStreamGetDevice");
    }
    if (sdev != cur)
        return set_error(MBX_E_UNSUPPORTED, "forward_static: the stream is on device %d but device "
                         "%d is current; make the stream's device current", int(sdev), cur);
    if (cur < 0 || cur >= kMaxDevices)
        return set_error(MBX_E_UNSUPPORTED, "forward_static: device %d beyond %d", cur, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[cur]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_status(e, "forward_static: hipGetSymbolAddress(scratch)");
        if (!p) return set_error(MBX_E_NO_DEVICE, "forward_static: no device address for scratch");
        g_scratch_ptr[cur] = p;
    }
    *out = g_scratch_ptr[cur];
    return MBX_OK;
}

template <typename T>
static int forward_static(const mbx_dims *d, const mbx_inputs *in, const mbx_outputs *out,
                          const mbx_beam_extents *b, const mbx_consts *k, const mbx_pair_plan *pp,
                          int options, hipStream_t stream) {
    if (!d) return set_error(MBX_E_NULL, "forward_static: NULL dims");
    int r = check_dims(d, true);
    if (r) return r;
    const size_t need = forward_ws_bytes<T>(*d);
    if (need > size_t(MBX_STATIC_SCRATCH_BYTES))
        return set_error(MBX_E_WORKSPACE, "forward_static: the workspace of these dimensions is %zu "
                         "bytes, beyond the static scratch of %zu bytes; use mbx_forward_%s with a "
                         "caller workspace", need, size_t(MBX_STATIC_SCRATCH_BYTES),
                         sizeof(T) == 8 ? "f64" : "f32");
    void *ws = nullptr;
    if ((r = scratch_on(stream, &ws))) return r;
    return forward<T>(d, in, out, b, k, pp, options, ws, size_t(MBX_STATIC_SCRATCH_BYTES), stream);
}

}  // namespace mbx

using namespace mbx;

extern "C" {

int mbx_forward_static_f32(const mbx_dims *d, const mbx_inputs *in, const mbx_outputs *out,
                           const mbx_beam_extents *beam, const mbx_consts *k,
                           const mbx_pair_plan *plan, int options, void *s) {
    return forward_static<float>(d, in, out, beam, k, plan, options, (hipStream_t)s);
}
int mbx_forward_static_f64(const mbx_dims *d, const mbx_inputs *in, const mbx_outputs *out,
                           const mbx_beam_extents *beam, const mbx_consts *k,
                           const mbx_pair_plan *plan, int options, void *s) {
    return forward_static<double>(d, in, out, beam, k, plan, options, (hipStream_t)s);
}

size_t mbx_forward_workspace_bytes(const mbx_dims *d, int is_double) {
    if (!d) return 0;
    return is_double ? forward_ws_bytes<double>(*d) : forward_ws_bytes<float>(*d);
}

int mbx_forward_f32(const mbx_dims *d, const mbx_inputs *in, const mbx_outputs *out,
                    const mbx_beam_extents *beam, const mbx_consts *k, const mbx_pair_plan *plan,
                    int options, void *ws, size_t ws_bytes, void *s) {
    return forward<float>(d, in, out, beam, k, plan, options, ws, ws_bytes, (hipStream_t)s);
}
int mbx_forward_f64(const mbx_dims *d, const mbx_inputs *in, const mbx_outputs *out,
                    const mbx_beam_extents *beam, const mbx_consts *k, const mbx_pair_plan *plan,
                    int options, void *ws, size_t ws_bytes, void *s) {
    return forward<double>(d, in, out, beam, k, plan, options, ws, ws_bytes, (hipStream_t)s);
}

}  // extern "C"
