// Stage kernels: one HIP kernel per v4 pipeline node, so that partial
// pipelines (the reference's per-stage tests) run on the GPU exactly as the
// node-by-node reference does. The fused forward in mbx_fused.hip is the
// performance path; these keep the reference's scratch arrays (B_sqrt,
// jones) and its stage-by-stage semantics.
//
// All kernels use flat 1-D grids with 64-bit indices and a grid-stride
// loop, so any problem size fits the launch limits.

#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "mbx_common.hpp"

namespace mbx {

// ---------------------------------------------------------------------
// Error reporting and validation (host)

static thread_local char g_last_error[512] = "";

int set_error(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

int hip_status(hipError_t e, const char *where) {
    if (e == hipSuccess) return MBX_OK;
    return set_error((int)e, "%s: %s", where, hipGetErrorString(e));
}

static int check_dim(const mbx_dim &x, const char *name, bool zero_ok) {
    if (x.local_size > x.global_size)
        return set_error(MBX_E_DIMS, "%s: local_size %u > global_size %u", name,
                         x.local_size, x.global_size);
    if (x.lower_extent > x.upper_extent || x.upper_extent > x.global_size)
        return set_error(MBX_E_EXTENT, "%s: extents [%u, %u) outside global size %u", name,
                         x.lower_extent, x.upper_extent, x.global_size);
    if (x.upper_extent - x.lower_extent > x.local_size)
        return set_error(MBX_E_EXTENT, "%s: extent %u exceeds local size %u", name,
                         x.upper_extent - x.lower_extent, x.local_size);
    if (!zero_ok && x.local_size == 0)
        return set_error(MBX_E_DIMS, "%s: size is zero", name);
    return MBX_OK;
}

int check_dims(const mbx_dims *d, bool need_beam) {
    if (!d) return set_error(MBX_E_NULL, "dims is NULL");
    int r;
#define CHK(f, z) if ((r = check_dim(d->f, #f, z)) != MBX_OK) return r
    CHK(ntime, false); CHK(na, false); CHK(nbl, false); CHK(nchan, false);
    CHK(npol, false); CHK(npolchan, false);
    CHK(nsrc, true); CHK(npsrc, true); CHK(ngsrc, true); CHK(nssrc, true);
    if (need_beam) { CHK(beam_lw, false); CHK(beam_mh, false); CHK(beam_nud, false); }
#undef CHK
    if (d->npol.local_size != 4)
        return set_error(MBX_E_UNSUPPORTED, "npol must be 4 (got %u)", d->npol.local_size);
    if (d->npolchan.local_size != d->nchan.local_size * 4)
        return set_error(MBX_E_DIMS, "npolchan (%u) != 4*nchan (%u)",
                         d->npolchan.local_size, 4 * d->nchan.local_size);
    if (dext(d->nsrc) != dext(d->npsrc) + dext(d->ngsrc) + dext(d->nssrc))
        return set_error(MBX_E_EXTENT, "nsrc extent %u != point+gaussian+sersic extents %u",
                         dext(d->nsrc), dext(d->npsrc) + dext(d->ngsrc) + dext(d->nssrc));
    if (d->nsrc.local_size < d->npsrc.local_size + d->ngsrc.local_size + d->nssrc.local_size &&
        d->nsrc.local_size < dext(d->nsrc))
        return set_error(MBX_E_DIMS, "nsrc local size %u too small", d->nsrc.local_size);
    return MBX_OK;
}

static unsigned grid_for(uint64_t n, unsigned block) {
    uint64_t g = (n + block - 1) / block;
    if (g > 65536ull * 16) g = 65536ull * 16;
    if (g == 0) g = 1;
    return (unsigned)g;
}

// ---------------------------------------------------------------------
// B_sqrt: one work-item per (source, time, channel).

template <typename T>
__global__ void __launch_bounds__(256)
k_b_sqrt(mbx_dims d, const T *__restrict__ stokes, const T *__restrict__ alpha,
         const T *__restrict__ freq, const T *__restrict__ ref_freq, cx<T> *__restrict__ B_sqrt) {
    const uint64_t nc = dext(d.nchan), nt = dext(d.ntime), ns = dext(d.nsrc);
    const uint64_t NT = d.ntime.local_size, NC = d.nchan.local_size;
    const uint64_t total = nc * nt * ns;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t c = i % nc, t = (i / nc) % nt, s = i / (nc * nt);
        const T *S = stokes + (s * NT + t) * 4;
        J4<T> B = brightness_sqrt<T>(S[0], S[1], S[2], S[3], alpha[s * NT + t], freq[c], ref_freq[c]);
        jstore(B_sqrt + ((s * NT + t) * NC + c) * 4, B);
    }
}

// ---------------------------------------------------------------------
// E beam: one work-item per (source, time, antenna, channel).
// Grid coordinates are evaluated in float64 for both dtypes.

template <typename T>
__device__ __forceinline__ J4<T> e_beam_eval(const cx<T> *__restrict__ E, const mbx_dims &d,
                                             double l, double m, double gc0, double gc1,
                                             T chd, double lscale, double mscale,
                                             double ll, double lm) {
    const double LW1 = double(d.beam_lw.local_size - 1), MH1 = double(d.beam_mh.local_size - 1);
    double vl = fmin(LW1, fmax(0.0, lscale * (l - ll)));
    double gl0 = floor(vl), gl1 = fmin(gl0 + 1.0, LW1);
    T ld = T(vl - gl0);
    double vm = fmin(MH1, fmax(0.0, mscale * (m - lm)));
    double gm0 = floor(vm), gm1 = fmin(gm0 + 1.0, MH1);
    T md = T(vm - gm0);
    const uint64_t MH = d.beam_mh.local_size, NUD = d.beam_nud.local_size;
    const double gls[2] = {gl0, gl1}, gms[2] = {gm0, gm1}, gcs[2] = {gc0, gc1};
    const T wl[2] = {T(1) - ld, ld}, wm[2] = {T(1) - md, md}, wc[2] = {T(1) - chd, chd};
    cx<T> acc[4];
    T amp[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) { acc[p] = {T(0), T(0)}; amp[p] = T(0); }
#pragma unroll
    for (int kc = 0; kc < 2; ++kc)
#pragma unroll
        for (int km = 0; km < 2; ++km)
#pragma unroll
            for (int kl = 0; kl < 2; ++kl) {
                T w = wl[kl] * wm[km] * wc[kc];
                uint64_t base = ((uint64_t(gls[kl]) * MH + uint64_t(gms[km])) * NUD + uint64_t(gcs[kc])) * 4;
#pragma unroll
                for (int p = 0; p < 4; ++p) {
                    cx<T> e = E[base + p];
                    acc[p].re = fma(w, e.re, acc[p].re);
                    acc[p].im = fma(w, e.im, acc[p].im);
                    amp[p] = fma(w, sqrt(e.re * e.re + e.im * e.im), amp[p]);
                }
            }
    J4<T> r;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        T norm = rsqrt(acc[p].re * acc[p].re + acc[p].im * acc[p].im);
        if (!isfinite(norm)) norm = T(1);
        T g = norm * amp[p];
        r.v[p] = {acc[p].re * g, acc[p].im * g};
    }
    return r;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_e_beam(mbx_dims d, const T *__restrict__ lm, const T *__restrict__ pa, const T *__restrict__ pe,
         const T *__restrict__ ascale, const T *__restrict__ freq, const cx<T> *__restrict__ E,
         cx<T> *__restrict__ jones, mbx_beam_extents b) {
    const uint64_t nc = dext(d.nchan), nt = dext(d.ntime), ns = dext(d.nsrc), na = dext(d.na);
    const uint64_t NT = d.ntime.local_size, NA = d.na.local_size, NC = d.nchan.local_size;
    const uint64_t total = nc * na * nt * ns;
    const double lscale = double(d.beam_lw.local_size - 1) / (b.ul - b.ll);
    const double mscale = double(d.beam_mh.local_size - 1) / (b.um - b.lm);
    const double NUD1 = double(d.beam_nud.local_size - 1);
    const double fscale = NUD1 / (b.ufreq - b.lfreq);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t c = i % nc, a = (i / nc) % na, t = (i / (nc * na)) % nt, s = i / (nc * na * nt);
        double sn, cs;
        sincos(double(pa[t * NA + a]), &sn, &cs);
        double l0 = lm[2 * s], m0 = lm[2 * s + 1];
        const T *pet = pe + ((t * NA + a) * NC + c) * 2;
        const T *asc = ascale + (a * NC + c) * 2;
        double l = (l0 * cs - m0 * sn + double(pet[0])) * double(asc[0]);
        double m = (l0 * sn + m0 * cs + double(pet[1])) * double(asc[1]);
        double vc = fmin(NUD1, fmax(0.0, fscale * (double(freq[c]) - b.lfreq)));
        double gc0 = floor(vc), gc1 = fmin(gc0 + 1.0, NUD1);
        J4<T> Ej = e_beam_eval<T>(E, d, l, m, gc0, gc1, T(vc - gc0), lscale, mscale, b.ll, b.lm);
        jstore(jones + (((s * NT + t) * NA + a) * NC + c) * 4, Ej);
    }
}

// ---------------------------------------------------------------------
// EKB sqrt: jones <- jones . (K * B_sqrt), per (source, time, antenna, channel).

template <typename T>
__global__ void __launch_bounds__(256)
k_ekb_sqrt(mbx_dims d, const T *__restrict__ uvw, const T *__restrict__ lm,
           const T *__restrict__ freq, const cx<T> *__restrict__ B_sqrt, cx<T> *__restrict__ jones,
           double two_pi_over_c) {
    const uint64_t nc = dext(d.nchan), nt = dext(d.ntime), ns = dext(d.nsrc), na = dext(d.na);
    const uint64_t NT = d.ntime.local_size, NA = d.na.local_size, NC = d.nchan.local_size;
    const uint64_t total = nc * na * nt * ns;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t c = i % nc, a = (i / nc) % na, t = (i / (nc * na)) % nt, s = i / (nc * na * nt);
        double l = lm[2 * s], m = lm[2 * s + 1];
        double nm1 = sqrt(1.0 - l * l - m * m) - 1.0;
        const T *x = uvw + (t * NA + a) * 3;
        double sn, cs;
        phase_sincos(double(x[0]), double(x[1]), double(x[2]), l, m, nm1, double(freq[c]),
                     two_pi_over_c, &sn, &cs);
        cx<T> K = {T(cs), T(sn)};
        J4<T> B = jload(B_sqrt + ((s * NT + t) * NC + c) * 4);
        J4<T> KB;
#pragma unroll
        for (int p = 0; p < 4; ++p) KB.v[p] = cmul(K, B.v[p]);
        cx<T> *jp = jones + (((s * NT + t) * NA + a) * NC + c) * 4;
        jstore(jp, jmul(jload(jp), KB));
    }
}

// ---------------------------------------------------------------------
// Coherency sum, gains, flags, residual and chi^2 terms, one work-item per
// (time, baseline, channel).  RimeSumCoherencies.py:110-393.

template <typename T>
__device__ __forceinline__ T gauss_factor(T u, T v, T el, T em, T R, T scale) {
    T u1 = (u * em - v * el) * scale * R;
    T v1 = (u * el + v * em) * scale;
    return exp(-(u1 * u1 + v1 * v1));
}

template <typename T>
__device__ __forceinline__ T sersic_factor(T u, T v, T e1, T e2, T R, T scale) {
    T k = scale * R / (T(1) - e1 * e1 - e2 * e2);
    T u1 = (u * (T(1) + e1) + v * e2) * k;
    T v1 = (u * e2 + v * (T(1) - e1)) * k;
    T den = T(1) + u1 * u1 + v1 * v1;
    return T(1) / (den * sqrt(den));
}

// Epilogue shared with the fused kernel: M = G_p V G_q^H, flags, output,
// chi^2 term. Returns the term (unweighted by sigma^2).
template <typename T>
__device__ __forceinline__ T vis_epilogue(const J4<T> &V, const cx<T> *__restrict__ Gp,
                                          const cx<T> *__restrict__ Gq, const uint8_t *__restrict__ fl,
                                          const T *__restrict__ wv, const cx<T> *__restrict__ obs,
                                          cx<T> *__restrict__ out, int options) {
    J4<T> M = jmul_h(jmul(jload(Gp), V), jload(Gq));
    J4<T> O = jload(obs);
    uchar4 f = *reinterpret_cast<const uchar4 *>(fl);
    const uint8_t fa[4] = {f.x, f.y, f.z, f.w};
    J4<T> R;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        if (fa[p] > 0) { M.v[p] = {T(0), T(0)}; O.v[p] = {T(0), T(0)}; }
        R.v[p] = {O.v[p].re - M.v[p].re, O.v[p].im - M.v[p].im};
    }
    jstore(out, (options & MBX_RESIDUAL_OUTPUT) ? R : M);
    T term = T(0);
    if (options & MBX_APPLY_WEIGHTS) {
#pragma unroll
        for (int p = 0; p < 4; ++p)
            term += (R.v[p].re * R.v[p].re + R.v[p].im * R.v[p].im) * wv[p];
    } else {
#pragma unroll
        for (int p = 0; p < 4; ++p) term += R.v[p].re * R.v[p].re + R.v[p].im * R.v[p].im;
    }
    return term;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_sum_coherencies(mbx_dims d, const T *__restrict__ uvw, const T *__restrict__ gauss,
                  const T *__restrict__ sersic, const T *__restrict__ freq,
                  const int32_t *__restrict__ a1, const int32_t *__restrict__ a2,
                  const cx<T> *__restrict__ jones, const uint8_t *__restrict__ flag,
                  const T *__restrict__ wv, const cx<T> *__restrict__ obs,
                  const cx<T> *__restrict__ G, cx<T> *__restrict__ vis, T *__restrict__ chi2,
                  double gauss_scale, double two_pi_over_c, int options) {
    const uint64_t nc = dext(d.nchan), nt = dext(d.ntime), nb = dext(d.nbl);
    const uint64_t NT = d.ntime.local_size, NA = d.na.local_size, NC = d.nchan.local_size,
                   NBL = d.nbl.local_size;
    const uint64_t total = nc * nb * nt;
    const uint32_t np = dext(d.npsrc), ng = dext(d.ngsrc), nss = dext(d.nssrc);
    const uint64_t NG = d.ngsrc.local_size, NS = d.nssrc.local_size;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t c = i % nc, b = (i / nc) % nb, t = i / (nc * nb);
        uint64_t vi = (t * NBL + b) * NC + c;
        const int na_max = int(NA) - 1;
        int p = min(max(a1[t * NBL + b], 0), na_max), q = min(max(a2[t * NBL + b], 0), na_max);
        J4<T> acc = (d.nsrc.lower_extent > 0) ? jload(vis + vi * 4) : jzero<T>();
        const T *xp = uvw + (t * NA + p) * 3, *xq = uvw + (t * NA + q) * 3;
        T u = xq[0] - xp[0], v = xq[1] - xp[1];
        T f = freq[c];
        const uint64_t jstride = NT * NA * NC * 4;
        const cx<T> *jp = jones + ((t * NA + p) * NC + c) * 4;
        const cx<T> *jq = jones + ((t * NA + q) * NC + c) * 4;
        uint32_t s = 0;
        for (; s < np; ++s)
            jacc_h(acc, jload(jp + s * jstride), jload(jq + s * jstride));
        const T gs = T(gauss_scale) * f;
        for (uint32_t k = 0; k < ng; ++k, ++s) {
            T fac = gauss_factor(u, v, gauss[k], gauss[k + NG], gauss[k + 2 * NG], gs);
            jacc_h(acc, jscale(jload(jp + s * jstride), fac), jload(jq + s * jstride));
        }
        const T ss = T(two_pi_over_c) * f;
        for (uint32_t k = 0; k < nss; ++k, ++s) {
            T fac = sersic_factor(u, v, sersic[k], sersic[k + NS], sersic[k + 2 * NS], ss);
            jacc_h(acc, jscale(jload(jp + s * jstride), fac), jload(jq + s * jstride));
        }
        if (d.nsrc.upper_extent < d.nsrc.global_size) {
            jstore(vis + vi * 4, acc);
            continue;
        }
        chi2[vi] = vis_epilogue(acc, G + ((t * NA + p) * NC + c) * 4, G + ((t * NA + q) * NC + c) * 4,
                                flag + vi * 4, wv ? wv + vi * 4 : nullptr, obs + vi * 4,
                                vis + vi * 4, options);
    }
}

// ---------------------------------------------------------------------
// chi^2 reduction: float64 partial sums per block, then one block sums the
// partials in a fixed order (deterministic).

constexpr unsigned kChi2Blocks = 1024;

template <typename T>
__global__ void __launch_bounds__(256)
k_chi2_partial(mbx_dims d, const T *__restrict__ terms, double *__restrict__ partial) {
    const uint64_t nc = dext(d.nchan), nb = dext(d.nbl), nt = dext(d.ntime);
    const uint64_t NC = d.nchan.local_size, NBL = d.nbl.local_size;
    const uint64_t total = nc * nb * nt;
    double acc = 0.0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t c = i % nc, b = (i / nc) % nb, t = i / (nc * nb);
        acc += double(terms[(t * NBL + b) * NC + c]);
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (unsigned w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256)
k_chi2_final(const double *__restrict__ partial, unsigned n, double scale, double *__restrict__ X2) {
    __shared__ double red[256];
    double acc = 0.0;
    for (unsigned i = threadIdx.x; i < n; i += 256) acc += partial[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (unsigned w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *X2 = red[0] * scale;
}

void launch_chi2_final(const double *partial, unsigned n, double scale, double *X2,
                       hipStream_t st) {
    hipLaunchKernelGGL(k_chi2_final, dim3(1), dim3(256), 0, st, partial, n, scale, X2);
}

template <typename T>
static int chi2_reduce(const mbx_dims *d, const T *terms, double *X2, const mbx_consts *k,
                       int options, void *ws, size_t ws_bytes, hipStream_t st) {
    int r = check_dims(d, false);
    if (r) return r;
    if (!terms || !X2 || !k) return set_error(MBX_E_NULL, "chi2_reduce: NULL pointer");
    if (!ws || ws_bytes < kChi2Blocks * sizeof(double))
        return set_error(MBX_E_WORKSPACE, "chi2_reduce: workspace needs %zu bytes",
                         (size_t)kChi2Blocks * sizeof(double));
    double *partial = static_cast<double *>(ws);
    hipLaunchKernelGGL(k_chi2_partial<T>, dim3(kChi2Blocks), dim3(256), 0, st, *d, terms, partial);
    double scale = (options & MBX_APPLY_WEIGHTS) ? 1.0 : 1.0 / k->sigma_sqrd;
    hipLaunchKernelGGL(k_chi2_final, dim3(1), dim3(256), 0, st, partial, kChi2Blocks, scale, X2);
    return hip_status(hipGetLastError(), "chi2_reduce launch");
}

// ---------------------------------------------------------------------
// Entry point bodies.

template <typename T>
static int b_sqrt(const mbx_dims *d, const T *stokes, const T *alpha, const T *freq,
                  const T *rf, void *B, hipStream_t st) {
    int r = check_dims(d, false);
    if (r) return r;
    if (!stokes || !alpha || !freq || !rf || !B) return set_error(MBX_E_NULL, "b_sqrt: NULL pointer");
    uint64_t n = uint64_t(dext(d->nchan)) * dext(d->ntime) * dext(d->nsrc);
    if (n == 0) return MBX_OK;
    hipLaunchKernelGGL(k_b_sqrt<T>, dim3(grid_for(n, 256)), dim3(256), 0, st, *d, stokes, alpha,
                       freq, rf, static_cast<cx<T> *>(B));
    return hip_status(hipGetLastError(), "b_sqrt launch");
}

template <typename T>
static int e_beam(const mbx_dims *d, const T *lm, const T *pa, const T *pe, const T *as,
                  const T *freq, const void *E, void *jones, const mbx_beam_extents *b,
                  hipStream_t st) {
    int r = check_dims(d, true);
    if (r) return r;
    if (!lm || !pa || !pe || !as || !freq || !E || !jones || !b)
        return set_error(MBX_E_NULL, "e_beam: NULL pointer");
    if (!(b->ul != b->ll && b->um != b->lm && b->ufreq != b->lfreq))
        return set_error(MBX_E_DIMS, "e_beam: degenerate beam extents");
    uint64_t n = uint64_t(dext(d->nchan)) * dext(d->na) * dext(d->ntime) * dext(d->nsrc);
    if (n == 0) return MBX_OK;
    hipLaunchKernelGGL(k_e_beam<T>, dim3(grid_for(n, 256)), dim3(256), 0, st, *d, lm, pa, pe, as,
                       freq, static_cast<const cx<T> *>(E), static_cast<cx<T> *>(jones), *b);
    return hip_status(hipGetLastError(), "e_beam launch");
}

template <typename T>
static int ekb_sqrt(const mbx_dims *d, const T *uvw, const T *lm, const T *freq, const void *B,
                    void *jones, const mbx_consts *k, hipStream_t st) {
    int r = check_dims(d, false);
    if (r) return r;
    if (!uvw || !lm || !freq || !B || !jones || !k) return set_error(MBX_E_NULL, "ekb_sqrt: NULL pointer");
    uint64_t n = uint64_t(dext(d->nchan)) * dext(d->na) * dext(d->ntime) * dext(d->nsrc);
    if (n == 0) return MBX_OK;
    hipLaunchKernelGGL(k_ekb_sqrt<T>, dim3(grid_for(n, 256)), dim3(256), 0, st, *d, uvw, lm, freq,
                       static_cast<const cx<T> *>(B), static_cast<cx<T> *>(jones), k->two_pi_over_c);
    return hip_status(hipGetLastError(), "ekb_sqrt launch");
}

template <typename T>
static int sum_coherencies(const mbx_dims *d, const T *uvw, const T *gauss, const T *sersic,
                           const T *freq, const int32_t *a1, const int32_t *a2, const void *jones,
                           const uint8_t *flag, const T *wv, const void *obs, const void *G,
                           void *vis, T *chi2, const mbx_consts *k, int options, hipStream_t st) {
    int r = check_dims(d, false);
    if (r) return r;
    if (!uvw || !freq || !a1 || !a2 || !flag || !obs || !G || !vis || !chi2 || !k)
        return set_error(MBX_E_NULL, "sum_coherencies: NULL pointer");
    if (dext(d->nsrc) > 0 && !jones) return set_error(MBX_E_NULL, "sum_coherencies: jones is NULL");
    if (dext(d->ngsrc) > 0 && !gauss) return set_error(MBX_E_NULL, "sum_coherencies: gauss_shape is NULL");
    if (dext(d->nssrc) > 0 && !sersic) return set_error(MBX_E_NULL, "sum_coherencies: sersic_shape is NULL");
    if ((options & MBX_APPLY_WEIGHTS) && !wv)
        return set_error(MBX_E_NULL, "sum_coherencies: weight_vector is NULL");
    if (options & ~(MBX_APPLY_WEIGHTS | MBX_RESIDUAL_OUTPUT))
        return set_error(MBX_E_UNSUPPORTED, "sum_coherencies: unknown option bits 0x%x", options);
    uint64_t n = uint64_t(dext(d->nchan)) * dext(d->nbl) * dext(d->ntime);
    if (n == 0) return MBX_OK;
    hipLaunchKernelGGL(k_sum_coherencies<T>, dim3(grid_for(n, 256)), dim3(256), 0, st, *d, uvw,
                       gauss, sersic, freq, a1, a2, static_cast<const cx<T> *>(jones), flag, wv,
                       static_cast<const cx<T> *>(obs), static_cast<const cx<T> *>(G),
                       static_cast<cx<T> *>(vis), chi2, k->gauss_scale, k->two_pi_over_c, options);
    return hip_status(hipGetLastError(), "sum_coherencies launch");
}

}  // namespace mbx

using namespace mbx;

extern "C" {

int mbx_abi_version(void) { return MBX_ABI_VERSION; }
const char *mbx_last_error(void) { return g_last_error; }
int mbx_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int mbx_b_sqrt_f32(const mbx_dims *d, const float *stokes, const float *alpha,
                   const float *frequency, const float *ref_frequency, void *B_sqrt, void *s) {
    return b_sqrt<float>(d, stokes, alpha, frequency, ref_frequency, B_sqrt, (hipStream_t)s);
}
int mbx_b_sqrt_f64(const mbx_dims *d, const double *stokes, const double *alpha,
                   const double *frequency, const double *ref_frequency, void *B_sqrt, void *s) {
    return b_sqrt<double>(d, stokes, alpha, frequency, ref_frequency, B_sqrt, (hipStream_t)s);
}

int mbx_e_beam_f32(const mbx_dims *d, const float *lm, const float *pa, const float *pe,
                   const float *as, const float *frequency, const void *E_beam, void *jones,
                   const mbx_beam_extents *beam, void *s) {
    return e_beam<float>(d, lm, pa, pe, as, frequency, E_beam, jones, beam, (hipStream_t)s);
}
int mbx_e_beam_f64(const mbx_dims *d, const double *lm, const double *pa, const double *pe,
                   const double *as, const double *frequency, const void *E_beam, void *jones,
                   const mbx_beam_extents *beam, void *s) {
    return e_beam<double>(d, lm, pa, pe, as, frequency, E_beam, jones, beam, (hipStream_t)s);
}

int mbx_ekb_sqrt_f32(const mbx_dims *d, const float *uvw, const float *lm, const float *frequency,
                     const void *B_sqrt, void *jones, const mbx_consts *k, void *s) {
    return ekb_sqrt<float>(d, uvw, lm, frequency, B_sqrt, jones, k, (hipStream_t)s);
}
int mbx_ekb_sqrt_f64(const mbx_dims *d, const double *uvw, const double *lm, const double *frequency,
                     const void *B_sqrt, void *jones, const mbx_consts *k, void *s) {
    return ekb_sqrt<double>(d, uvw, lm, frequency, B_sqrt, jones, k, (hipStream_t)s);
}

int mbx_sum_coherencies_f32(const mbx_dims *d, const float *uvw, const float *gauss_shape,
                            const float *sersic_shape, const float *frequency,
                            const int32_t *antenna1, const int32_t *antenna2, const void *jones,
                            const uint8_t *flag, const float *weight_vector,
                            const void *observed_vis, const void *G_term, void *model_vis,
                            float *chi_sqrd_result, const mbx_consts *k, int options, void *s) {
    return sum_coherencies<float>(d, uvw, gauss_shape, sersic_shape, frequency, antenna1, antenna2,
                                  jones, flag, weight_vector, observed_vis, G_term, model_vis,
                                  chi_sqrd_result, k, options, (hipStream_t)s);
}
int mbx_sum_coherencies_f64(const mbx_dims *d, const double *uvw, const double *gauss_shape,
                            const double *sersic_shape, const double *frequency,
                            const int32_t *antenna1, const int32_t *antenna2, const void *jones,
                            const uint8_t *flag, const double *weight_vector,
                            const void *observed_vis, const void *G_term, void *model_vis,
                            double *chi_sqrd_result, const mbx_consts *k, int options, void *s) {
    return sum_coherencies<double>(d, uvw, gauss_shape, sersic_shape, frequency, antenna1, antenna2,
                                   jones, flag, weight_vector, observed_vis, G_term, model_vis,
                                   chi_sqrd_result, k, options, (hipStream_t)s);
}

size_t mbx_chi2_workspace_bytes(void) { return kChi2Blocks * sizeof(double); }

int mbx_chi2_reduce_f32(const mbx_dims *d, const float *chi_sqrd_result, double *X2,
                        const mbx_consts *k, int options, void *ws, size_t ws_bytes, void *s) {
    return chi2_reduce<float>(d, chi_sqrd_result, X2, k, options, ws, ws_bytes, (hipStream_t)s);
}
int mbx_chi2_reduce_f64(const mbx_dims *d, const double *chi_sqrd_result, double *X2,
                        const mbx_consts *k, int options, void *ws, size_t ws_bytes, void *s) {
    return chi2_reduce<double>(d, chi_sqrd_result, X2, k, options, ws, ws_bytes, (hipStream_t)s);
}

}  // extern "C"
