// Device machinery shared by the fused RIME kernels (forward, adjoint,
// batched): argument block, LDS tile layout, the per-source and
// per-antenna phases, shape factors and the visibility epilogue.
//
// A workgroup owns one time step t and a tile of `ct` channels. For each
// tile of `st` sources it stages in LDS
//   phase 0: source parameters (l, m, n-1, shape) and sqrt(B) per
//            (source, channel)                      [CPUSolver.py:235-317]
//   phase 1: per-antenna Jones J = E . (K sqrt(B)) per (source, antenna,
//            channel)               [CPUSolver.py:162-233, 319-475, 543-560]
// and the kernel-specific phase 2 consumes the tile.
#pragma once

#include "mbx_common.hpp"

namespace mbx {

// Prepared beam: the cube interpolated along frequency to each channel
// once per launch (the frequency weight of the trilinear lookup depends on
// the channel only), one cell per (channel, l, m) holding (re, im, |.|)
// of the four correlations. A Jones lookup then gathers four 48/96-byte
// cells instead of eight taps of four correlations.
template <typename T> struct alignas(4 * sizeof(T)) BeamCell { T v[12]; };

// Several samples in one launch (batched evaluation): workgroups
// [j wgs, (j + 1) wgs) belong to slot slot0 + j of the processing order,
// sample k = order[slot0 + j]. Sample k's sources are rows
// [src_offset[k], src_offset[k] + n_k) of lm / stokes / alpha and columns
// from gauss_offset[k] / sersic_offset[k] of the shape arrays (whose row
// stride stays the total column count); per-slot visibility buffers
// (running sums, pre-gain sums, Vb) lie vstride complex values apart, and
// the chi^2 partials of slot j from j wgs on.
struct SampleSpan {
    const int32_t *src_offset, *npsrc, *ngsrc, *nssrc, *gauss_offset, *sersic_offset, *order;
    uint32_t wgs;       // workgroups per sample; 0: one model
    int32_t slot0;
    uint64_t vstride;
};

template <typename T> struct RimeArgs {
    SampleSpan span;
    mbx_dims d;
    const T *uvw, *lm, *stokes, *alpha, *freq, *ref_freq, *pa, *pe, *ascale, *gauss, *sersic, *wv;
    const BeamCell<T> *beam;  // [nchan][beam_lw][beam_mh]
    const int32_t *a1, *a2;
    const uint8_t *flag;
    const cx<T> *obs, *G;
    cx<T> *model;
    T *chi2;
    cx<T> *ekb;
    const cx<T> *vin;         // running source sum to start from (or NULL)
    double *partial;
    double ll, lmin, lscale, mscale, lfreq, fscale;
    double gauss_scale, two_pi_over_c;
    int options;
    int ct, st, nbchunk, bchunk, nctile;
    int s_begin;              // first source the kernel sums (earlier ones are in vin)
    uint32_t per_xcd;         // workgroup slots per XCD (grid = 8 per_xcd)
    const int32_t *blocks;    // pair-block plan (blocked kernels)
    int nblocks, pb, qb;
};

// Cell (c, l, m): (1 - w) X[l, m, g0] + w X[l, m, g1] for X = value and
// |value| of each correlation, with g0, g1, w the channel's grid position
// (RimeEBeam.py:233-260: nu = (nud-1)(f - lfreq)/(ufreq - lfreq) clamped
// to [0, nud-1], g1 = min(g0 + 1, nud-1)).
template <typename T>
__global__ void __launch_bounds__(256)
k_beam_planes(const cx<T> *__restrict__ E, BeamCell<T> *__restrict__ out,
              const T *__restrict__ freq, uint32_t nchan, uint32_t lw, uint32_t mh, uint32_t nud,
              double lfreq, double fscale) {
    const uint64_t n = uint64_t(nchan) * lw * mh;
    const double NUD1 = double(nud - 1);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t lm = i % (uint64_t(lw) * mh);
        const uint32_t c = uint32_t(i / (uint64_t(lw) * mh));
        const double vc = fmin(NUD1, fmax(0.0, fscale * (double(freq[c]) - lfreq)));
        const double g0 = floor(vc);
        const uint32_t ic0 = uint32_t(g0), ic1 = uint32_t(fmin(g0 + 1.0, NUD1));
        const T w1 = T(vc - g0), w0 = T(1) - w1;
        const cx<T> *e0 = E + (lm * nud + ic0) * 4, *e1 = E + (lm * nud + ic1) * 4;
        BeamCell<T> cell;
#pragma unroll
        for (int x = 0; x < 4; ++x) {
            const cx<T> a = e0[x], b = e1[x];
            cell.v[3 * x] = w0 * a.re + w1 * b.re;
            cell.v[3 * x + 1] = w0 * a.im + w1 * b.im;
            cell.v[3 * x + 2] = w0 * sqrt(a.re * a.re + a.im * a.im) + w1 * sqrt(b.re * b.re + b.im * b.im);
        }
        out[i] = cell;
    }
}

template <typename T>
static inline size_t beam_planes_bytes(const mbx_dims &d) {
    const size_t n = size_t(d.nchan.local_size) * d.beam_lw.local_size * d.beam_mh.local_size;
    return (n * sizeof(BeamCell<T>) + 255) & ~size_t(255);
}

// Channel-interpolated beam planes of every channel, once per launch.
template <typename T>
static inline void launch_beam_planes(const mbx_dims &d, const void *E_beam, const void *freq,
                                      const mbx_beam_extents *b, BeamCell<T> *out,
                                      hipStream_t stream) {
    const uint64_t n = uint64_t(d.nchan.local_size) * d.beam_lw.local_size * d.beam_mh.local_size;
    const unsigned g = (unsigned)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536);
    const double fscale = double(d.beam_nud.local_size - 1) / (b->ufreq - b->lfreq);
    hipLaunchKernelGGL(k_beam_planes<T>, dim3(g), dim3(256), 0, stream,
                       static_cast<const cx<T> *>(E_beam), out, static_cast<const T *>(freq),
                       d.nchan.local_size, d.beam_lw.local_size, d.beam_mh.local_size,
                       d.beam_nud.local_size, b->lfreq, fscale);
}

// Argument block from the C-ABI structs (beam cube pointer, outputs and
// tiling are set by the caller).
template <typename T>
static inline void fill_args(RimeArgs<T> &A, const mbx_dims *d, const mbx_inputs *in,
                             const mbx_beam_extents *b, const mbx_consts *k, int options) {
    A.d = *d;
    A.uvw = (const T *)in->uvw; A.lm = (const T *)in->lm; A.stokes = (const T *)in->stokes;
    A.alpha = (const T *)in->alpha; A.freq = (const T *)in->frequency;
    A.ref_freq = (const T *)in->ref_frequency; A.pa = (const T *)in->parallactic_angles;
    A.pe = (const T *)in->point_errors; A.ascale = (const T *)in->antenna_scaling;
    A.gauss = (const T *)in->gauss_shape; A.sersic = (const T *)in->sersic_shape;
    A.wv = (const T *)in->weight_vector;
    A.a1 = in->antenna1; A.a2 = in->antenna2; A.flag = in->flag;
    A.obs = (const cx<T> *)in->observed_vis; A.G = (const cx<T> *)in->G_term;
    A.ll = b->ll; A.lmin = b->lm;
    A.lscale = double(d->beam_lw.local_size - 1) / (b->ul - b->ll);
    A.mscale = double(d->beam_mh.local_size - 1) / (b->um - b->lm);
    A.lfreq = b->lfreq;
    A.fscale = double(d->beam_nud.local_size - 1) / (b->ufreq - b->lfreq);
    A.gauss_scale = k->gauss_scale; A.two_pi_over_c = k->two_pi_over_c;
    A.options = options;
    A.model = nullptr; A.chi2 = nullptr; A.ekb = nullptr; A.vin = nullptr; A.partial = nullptr;
    A.s_begin = 0;
    A.per_xcd = 0;
    A.blocks = nullptr; A.nblocks = 0; A.pb = A.qb = 0;
    A.ct = A.st = A.nbchunk = A.bchunk = A.nctile = 1;
    A.span = SampleSpan{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0u, 0, 0ull};
}

// The workgroup's sample in a launch over a SampleSpan, and its block
// index within that sample's grid (blockIdx.x for one model).
struct SampleAt {
    uint32_t bx;        // block index within the sample's grid
    int j;              // slot within the launch
    int k;              // sample
    int src0, goff, soff;
};

// Points A at the workgroup's sample: source pointers and source extents
// become the sample's, the running / pre-gain sum buffers (vin, ekb) and
// the chi^2 partials the slot's.
template <typename T>
__device__ __forceinline__ SampleAt enter_sample(RimeArgs<T> &A) {
    SampleAt s{blockIdx.x, 0, 0, 0, 0, 0};
    const SampleSpan &S = A.span;
    if (S.wgs == 0) return s;
    s.j = int(blockIdx.x / S.wgs);
    s.bx = blockIdx.x - uint32_t(s.j) * S.wgs;
    s.k = S.order[S.slot0 + s.j];
    s.src0 = S.src_offset[s.k];
    s.goff = S.gauss_offset[s.k];
    s.soff = S.sersic_offset[s.k];
    const int np = S.npsrc[s.k], ng = S.ngsrc[s.k], nss = S.nssrc[s.k];
    const uint64_t NT = A.d.ntime.local_size;
    A.lm += 2 * uint64_t(s.src0);
    A.stokes += uint64_t(s.src0) * NT * 4;
    A.alpha += uint64_t(s.src0) * NT;
    if (A.gauss) A.gauss += s.goff;
    if (A.sersic) A.sersic += s.soff;
    A.d.npsrc.lower_extent = 0; A.d.npsrc.upper_extent = np;
    A.d.ngsrc.lower_extent = 0; A.d.ngsrc.upper_extent = ng;
    A.d.nssrc.lower_extent = 0; A.d.nssrc.upper_extent = nss;
    A.d.nsrc.lower_extent = 0; A.d.nsrc.upper_extent = np + ng + nss;
    A.d.nsrc.global_size = np + ng + nss;
    const uint64_t off = uint64_t(s.j) * S.vstride;
    if (A.ekb) A.ekb += off;
    if (A.vin) A.vin += off;
    if (A.partial) A.partial += uint64_t(s.j) * S.wgs;
    return s;
}

// enter_sample in the BATCHED instances of a kernel; one model otherwise
// (those instances carry no span code).
template <bool BATCHED, typename Args>
__device__ __forceinline__ SampleAt enter_sample_if(Args &A) {
    if constexpr (BATCHED) return enter_sample(A);
    else return SampleAt{blockIdx.x, 0, 0, 0, 0, 0};
}

// Workgroup order. The dispatcher deals workgroups to the eight XCDs
// round robin (workgroup x to XCD x mod 8), so slot w of XCD x is
// workgroup x + 8 w and the workgroups an XCD runs together are
// consecutive slots. Within an XCD's slots the channel varies fastest
// over a group of kChanGroup channels, then time: concurrent workgroups
// share kChanGroup channels, so their 32/64-byte visibility records fill
// whole cache lines, and touch few beam planes (the L2 is per XCD).
constexpr int kChanGroup = 4;

__host__ __device__ inline uint32_t xcd_slots(uint64_t nwg) { return uint32_t((nwg + 7) / 8); }

__device__ __forceinline__ uint64_t xcd_slot(uint32_t per_xcd, uint32_t bx = blockIdx.x) {
    return uint64_t(bx % 8) * per_xcd + bx / 8;
}

// (t, c) of a one-channel workgroup over nt x nc; false past the grid.
// bx: the block index within the grid (a sample's grid in a launch over
// several samples keeps bx mod 8 = blockIdx.x mod 8, its grid being a
// multiple of 8).
__device__ __forceinline__ bool tc_coords(uint32_t per_xcd, int nt, int nc, int &t, int &c,
                                          uint32_t bx = blockIdx.x) {
    uint64_t w = xcd_slot(per_xcd, bx);
    const int ci = int(w % kChanGroup);
    w /= kChanGroup;
    t = int(w % uint64_t(nt));
    c = int(w / uint64_t(nt)) * kChanGroup + ci;
    return c < nc;
}

// Grid of tc_coords: 8 slots per XCD, padded.
__host__ __device__ inline uint32_t tc_per_xcd(int nt, int nc) {
    return xcd_slots(uint64_t(nt) * ((nc + kChanGroup - 1) / kChanGroup) * kChanGroup);
}

// LDS carve-up of one tile.
constexpr int CH = 5;   // doubles per channel in the tile

template <typename T> struct Tile {
    double *uvw;     // [NA][3]
    double *pa;      // [NA][2]  sin, cos of the parallactic angle
    double *ch;      // [ct][CH] gc0, gc1, chd, frequency, log(frequency / ref)
    double *src;     // [st][3]  l, m, n - 1
    T *shape;        // [st][4]  p0, p1, p2, type (0 point, 1 gaussian, 2 sersic)
    J4<T> *sqB;      // [st][ct]
    J4<T> *J;        // [st][NA][ct]
    double *part;    // [nwaves]

    Tile() = default;
    __device__ __forceinline__ Tile(unsigned char *base, int NA, int ct, int st) {
        uvw = reinterpret_cast<double *>(base);
        pa = uvw + 3 * NA;
        ch = pa + 2 * NA;
        src = ch + CH * ct;
        shape = reinterpret_cast<T *>(src + 3 * st);
        // 16-byte aligned offset from the LDS base: the pointers stay
        // derived from the shared array, so their accesses compile to LDS
        // instructions (a round trip through an integer address would
        // make them generic flat accesses)
        const size_t off = (size_t(reinterpret_cast<unsigned char *>(shape + 4 * st) - base) + 15) & ~size_t(15);
        sqB = reinterpret_cast<J4<T> *>(base + off);
        J = sqB + st * ct;
        part = reinterpret_cast<double *>(J + st * NA * ct);
    }
};

template <typename T>
static inline size_t tile_lds_bytes(int na, int ct, int st, int nwaves) {
    return sizeof(double) * (5 * size_t(na) + CH * ct + 3 * st) + sizeof(T) * 4 * st + 16 +
           sizeof(J4<T>) * size_t(st) * ct * (1 + size_t(na)) + sizeof(double) * nwaves;
}

// exp(i phi) with phi evaluated in float64.
//
// Method (textbook, not a library's): Cody-Waite reduction of phi by pi/2
// in float64 with a two-term pi/2 (W. J. Cody, W. Waite, "Software Manual
// for the Elementary Functions", 1980) to y in [-pi/4, pi/4] and quadrant
// q, then the Taylor series of sin and cos on y (coefficients 1/n!), and a
// quadrant rotation. Degree 17/18 in float64 (truncation < 1e-19 relative),
// degree 9/10 in float (truncation < 2e-9). Keeps the kernels free of the
// general-argument paths of a library sincos. Accuracy is tested through
// the phase term against numpy (tests/test_gpu_*).
__device__ __forceinline__ double reduce_half_pi(double phi, int &quadrant) {
    const double two_over_pi = 0.6366197723675814;
    const double pio2_hi = 1.5707963267948966;       // pi/2 rounded to double
    const double pio2_lo = 6.123233995736766e-17;    // pi/2 - pio2_hi
    double q = rint(phi * two_over_pi);
    // |q| < 2^31 for |phi| < 3.3e9 rad (far past the ~1e5 rad to which the
    // two-term reduction keeps 1e-16 accuracy): a 32-bit conversion
    quadrant = int(q) & 3;
    return fma(-q, pio2_lo, fma(-q, pio2_hi, phi));
}

// A float64 constant placed in a scalar register pair where it is used:
// keeps loop-invariant polynomial coefficients from being hoisted into
// vector registers (and spilled) in register-bound kernels.
__device__ __forceinline__ double sconst(double v) {
    __asm__ volatile("" : "+s"(v));
    return v;
}

__device__ __forceinline__ void taylor_sincos(double y, double *s, double *c) {
    const double y2 = y * y;
    // sin y = y (1 - y^2/3! + y^4/5! - ... - y^16/17!)
    double ps = sconst(-1.0 / 355687428096000.0);            // -1/17!
    ps = fma(ps, y2, sconst(1.0 / 1307674368000.0));         //  1/15!
    ps = fma(ps, y2, sconst(-1.0 / 6227020800.0));           // -1/13!
    ps = fma(ps, y2, sconst(1.0 / 39916800.0));              //  1/11!
    ps = fma(ps, y2, sconst(-1.0 / 362880.0));               // -1/9!
    ps = fma(ps, y2, sconst(1.0 / 5040.0));                  //  1/7!
    ps = fma(ps, y2, sconst(-1.0 / 120.0));                  // -1/5!
    ps = fma(ps, y2, sconst(1.0 / 6.0));                     //  1/3!
    *s = fma(-y * y2, ps, y);
    // cos y = 1 - y^2/2! + y^4/4! - ... + y^18/18!
    double pc = sconst(1.0 / 6402373705728000.0);            //  1/18!
    pc = fma(pc, y2, sconst(-1.0 / 20922789888000.0));       // -1/16!
    pc = fma(pc, y2, sconst(1.0 / 87178291200.0));           //  1/14!
    pc = fma(pc, y2, sconst(-1.0 / 479001600.0));            // -1/12!
    pc = fma(pc, y2, sconst(1.0 / 3628800.0));               //  1/10!
    pc = fma(pc, y2, sconst(-1.0 / 40320.0));                // -1/8!
    pc = fma(pc, y2, sconst(1.0 / 720.0));                   //  1/6!
    pc = fma(pc, y2, sconst(-1.0 / 24.0));                   // -1/4!
    pc = fma(pc, y2, sconst(0.5));                           //  1/2!
    *c = fma(-y2, pc, 1.0);
}

__device__ __forceinline__ void taylor_sincos(float y, float *s, float *c) {
    const float y2 = y * y;
    float ps = 1.0f / 362880.0f;                     //  1/9!
    ps = fmaf(ps, y2, -1.0f / 5040.0f);
    ps = fmaf(ps, y2, 1.0f / 120.0f);
    ps = fmaf(ps, y2, -1.0f / 6.0f);
    *s = fmaf(y * y2, ps, y);
    float pc = -1.0f / 3628800.0f;                   // -1/10!
    pc = fmaf(pc, y2, 1.0f / 40320.0f);
    pc = fmaf(pc, y2, -1.0f / 720.0f);
    pc = fmaf(pc, y2, 1.0f / 24.0f);
    pc = fmaf(pc, y2, -0.5f);
    *c = fmaf(y2, pc, 1.0f);
}

// (sin, cos) of phi given the quadrant rotation of the reduced argument.
template <typename T>
__device__ __forceinline__ void sincos_phase(double phi, T *s, T *c) {
    int q;
    const double y = reduce_half_pi(phi, q);
    T sy, cy;
    taylor_sincos(T(y), &sy, &cy);
    const T ss = (q & 1) ? cy : sy;
    const T cc = (q & 1) ? sy : cy;
    *s = (q & 2) ? -ss : ss;
    *c = ((q + 1) & 2) ? -cc : cc;
}

template <typename T> struct PhaseCS {
    __device__ __forceinline__ static void eval(double phi, T *s, T *c) { sincos_phase<T>(phi, s, c); }
};

template <typename T>
__device__ __forceinline__ void tile_setup(const RimeArgs<T> &A, double *Luvw, double *Lpa,
                                           double *Lch, int t, int c0, int tid, int nth);

template <typename T>
__device__ __forceinline__ void tile_setup(const RimeArgs<T> &A, Tile<T> &L, int t, int c0,
                                           int tid, int nth) {
    tile_setup<T>(A, L.uvw, L.pa, L.ch, t, c0, tid, nth);
}

template <typename T>
__device__ __forceinline__ void tile_setup(const RimeArgs<T> &A, double *Luvw, double *Lpa,
                                           double *Lch, int t, int c0, int tid, int nth) {
    Tile<T> L;
    L.uvw = Luvw; L.pa = Lpa; L.ch = Lch;
    const mbx_dims &d = A.d;
    const int NA = d.na.local_size;
    const int nc_ext = dext(d.nchan);
    for (int a = tid; a < NA; a += nth) {
        const T *x = A.uvw + (uint64_t(t) * NA + a) * 3;
        L.uvw[3 * a] = x[0]; L.uvw[3 * a + 1] = x[1]; L.uvw[3 * a + 2] = x[2];
        double sn, cs;
        sincos_phase<double>(double(A.pa[uint64_t(t) * NA + a]), &sn, &cs);
        L.pa[2 * a] = sn; L.pa[2 * a + 1] = cs;
    }
    const double NUD1 = double(d.beam_nud.local_size - 1);
    for (int k = tid; k < A.ct; k += nth) {
        int c = min(c0 + k, nc_ext - 1);
        double f = A.freq[c];
        double vc = fmin(NUD1, fmax(0.0, A.fscale * (f - A.lfreq)));
        double g0 = floor(vc);
        L.ch[CH * k] = g0;
        L.ch[CH * k + 1] = fmin(g0 + 1.0, NUD1);
        L.ch[CH * k + 2] = vc - g0;
        L.ch[CH * k + 3] = f;
        L.ch[CH * k + 4] = log(f / double(A.ref_freq[c]));
    }
}

// Phase 0: parameters of sources [s0, s0 + nst) and sqrt(B) per channel.
template <typename T>
__device__ __forceinline__ void tile_phase0(const RimeArgs<T> &A, Tile<T> &L, int t, int c0,
                                            int s0, int nst, int tid, int nth) {
    const mbx_dims &d = A.d;
    const uint64_t NT = d.ntime.local_size;
    const int np = dext(d.npsrc), ng = dext(d.ngsrc);
    const uint64_t NG = d.ngsrc.local_size, NS = d.nssrc.local_size;
    const int nc_ext = dext(d.nchan);
    const int ct = A.ct;
    for (int i = tid; i < nst; i += nth) {
        int s = s0 + i;
        double l = A.lm[2 * uint64_t(s)], m = A.lm[2 * uint64_t(s) + 1];
        L.src[3 * i] = l;
        L.src[3 * i + 1] = m;
        L.src[3 * i + 2] = sqrt(1.0 - l * l - m * m) - 1.0;
        T p0 = T(0), p1 = T(0), p2 = T(0), ty = T(0);
        if (s >= np && s < np + ng) {
            int g = s - np;
            p0 = A.gauss[g]; p1 = A.gauss[g + NG]; p2 = A.gauss[g + 2 * NG]; ty = T(1);
        } else if (s >= np + ng) {
            int g = s - np - ng;
            p0 = A.sersic[g]; p1 = A.sersic[g + NS];
            p2 = A.sersic[g + 2 * NS] / (T(1) - p0 * p0 - p1 * p1); ty = T(2);
        }
        L.shape[4 * i] = p0; L.shape[4 * i + 1] = p1; L.shape[4 * i + 2] = p2;
        L.shape[4 * i + 3] = ty;
    }
    for (int i = tid; i < nst * ct; i += nth) {
        int si = i / ct, k = i % ct;
        int s = s0 + si;
        int c = min(c0 + k, nc_ext - 1);
        const T *S = A.stokes + (uint64_t(s) * NT + t) * 4;
        (void)c;
        const T power = T(exp(double(A.alpha[uint64_t(s) * NT + t]) * L.ch[CH * k + 4]));
        L.sqB[i] = brightness_sqrt_scaled<T>(S[0], S[1], S[2], S[3], power);
    }
}

// Beam Jones from the prepared planes (RimeEBeam.py:233-343 /
// CPUSolver.py:348-475 semantics: trilinear interpolation of the values
// and, separately, of their amplitudes, then E = value * amp / |value|)
// given the antenna's rotation (sin, cos of the parallactic angle),
// pointing error and scaling for this (time, channel c). Grid coordinates
// in float64. UT of the four bilinear taps are loaded together.
// The taps of one beam lookup: byte offsets in the channel's plane of the
// (0, 0) cell and of the +1 neighbours (0 at the last row / column), and
// the bilinear fractions.
template <typename T> struct BeamTapAddr {
    const char *plane;
    uint32_t o00, dl, dm;
    T ld, md;
    __device__ __forceinline__ const BeamCell<T> &cell(int tap) const {
        return *reinterpret_cast<const BeamCell<T> *>(plane + o00 + ((tap & 1) ? dl : 0u) +
                                                      ((tap & 2) ? dm : 0u));
    }
};

template <typename T>
__device__ __forceinline__ BeamTapAddr<T> beam_taps(const RimeArgs<T> &A, double ps, double pc,
                                                    double pel, double pem, double asl, double asm_,
                                                    int c, double l0, double m0) {
    const mbx_dims &d = A.d;
    double l = (l0 * pc - m0 * ps + pel) * asl;
    double m = (l0 * ps + m0 * pc + pem) * asm_;
    const double LW1 = double(d.beam_lw.local_size - 1), MH1 = double(d.beam_mh.local_size - 1);
    double vl = fmin(LW1, fmax(0.0, A.lscale * (l - A.ll)));
    double gl0 = floor(vl);
    double vm = fmin(MH1, fmax(0.0, A.mscale * (m - A.lmin)));
    double gm0 = floor(vm);
    BeamTapAddr<T> r;
    r.ld = T(vl - gl0);
    r.md = T(vm - gm0);
    const uint32_t MH = d.beam_mh.local_size;
    // a plane is a few MB at most: 32-bit offsets, 24-bit multiplies
    const uint32_t cell = uint32_t(sizeof(BeamCell<T>));
    r.o00 = uint32_t(__umul24(uint32_t(__umul24(uint32_t(gl0), MH)) + uint32_t(gm0), cell));
    r.dl = gl0 < LW1 ? uint32_t(__umul24(MH, cell)) : 0u;
    r.dm = gm0 < MH1 ? cell : 0u;
    r.plane = reinterpret_cast<const char *>(A.beam + uint64_t(c) * d.beam_lw.local_size * MH);
    return r;
}

// Weighted sums of the values and amplitudes of tap `tap` (taps in order 0..3).
template <typename T>
__device__ __forceinline__ void beam_accumulate(const BeamCell<T> &e, int tap, T ld, T md, cx<T> *acc, T *amp) {
    const bool hl = tap & 1, hm = (tap >> 1) & 1;
    const T w = (hl ? ld : T(1) - ld) * (hm ? md : T(1) - md);
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        acc[p].re = fma(w, e.v[3 * p], acc[p].re);
        acc[p].im = fma(w, e.v[3 * p + 1], acc[p].im);
        amp[p] = fma(w, e.v[3 * p + 2], amp[p]);
    }
}

// E = value * amp / |value| per correlation.
template <typename T>
__device__ __forceinline__ J4<T> beam_normalise(const cx<T> *acc, const T *amp) {
    J4<T> E;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        T n2 = acc[p].re * acc[p].re + acc[p].im * acc[p].im;
        T g = (n2 > T(0) && isfinite(n2)) ? amp[p] * rsqrt_acc(n2) : amp[p];
        E.v[p] = {acc[p].re * g, acc[p].im * g};
    }
    return E;
}

// Beam Jones from the prepared planes (RimeEBeam.py:233-343 /
// CPUSolver.py:348-475 semantics: trilinear interpolation of the values
// and, separately, of their amplitudes, then E = value * amp / |value|)
// given the antenna's rotation (sin, cos of the parallactic angle),
// pointing error and scaling for this (time, channel c). Grid coordinates
// in float64. UT of the four bilinear taps are loaded together.
template <typename T, int UT = 1>
__device__ __forceinline__ J4<T> beam_jones_at(const RimeArgs<T> &A, double ps, double pc,
                                               double pel, double pem, double asl, double asm_,
                                               int c, double l0, double m0) {
    const BeamTapAddr<T> ta = beam_taps<T>(A, ps, pc, pel, pem, asl, asm_, c, l0, m0);
    cx<T> acc[4];
    T amp[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) { acc[p] = {T(0), T(0)}; amp[p] = T(0); }
    static_assert(UT == 1 || UT == 2 || UT == 4, "taps per step");
#pragma unroll 1
    for (int tap0 = 0; tap0 < 4; tap0 += UT)
#pragma unroll
    for (int u = 0; u < UT; ++u) {
        const BeamCell<T> e = ta.cell(tap0 + u);
        beam_accumulate(e, tap0 + u, ta.ld, ta.md, acc, amp);
    }
    return beam_normalise(acc, amp);
}

// Beam Jones of one (source, antenna, channel), reading the antenna's
// pointing error and scaling from global memory.
template <typename T, int UT = 1>
__device__ __forceinline__ J4<T> beam_jones(const RimeArgs<T> &A, const double *Lpa,
                                            const double *Lch, int t, int a, int c, int k,
                                            double l0, double m0) {
    const mbx_dims &d = A.d;
    const uint64_t NA = d.na.local_size, NC = d.nchan.local_size;
    const T *pet = A.pe + ((uint64_t(t) * NA + a) * NC + c) * 2;
    const T *asc = A.ascale + (uint64_t(a) * NC + c) * 2;
    (void)Lch; (void)k;
    return beam_jones_at<T, UT>(A, Lpa[2 * a], Lpa[2 * a + 1], double(pet[0]), double(pet[1]),
                                double(asc[0]), double(asc[1]), c, l0, m0);
}

// Phase 1: J = E . (K sqrt(B)) for every (source, antenna, channel) of
// the tile.
template <typename T>
__device__ __forceinline__ void tile_phase1(const RimeArgs<T> &A, Tile<T> &L, int t, int c0,
                                            int nst, int tid, int nth) {
    const int NA = A.d.na.local_size;
    const int nc_ext = dext(A.d.nchan);
    const int ct = A.ct;
    const int n1 = nst * NA * ct;
    for (int i = tid; i < n1; i += nth) {
        const int k = i % ct;
        const int a = (i / ct) % NA;
        const int si = i / (ct * NA);
        const int c = min(c0 + k, nc_ext - 1);
        const double l0 = L.src[3 * si], m0 = L.src[3 * si + 1], nm1 = L.src[3 * si + 2];
        J4<T> E = beam_jones<T>(A, L.pa, L.ch, t, a, c, k, l0, m0);
        const double *x = L.uvw + 3 * a;
        double p = x[2] * nm1 + x[1] * m0 + x[0] * l0;
        double phi = -A.two_pi_over_c * L.ch[CH * k + 3] * p;
        T sn, cs;
        PhaseCS<T>::eval(phi, &sn, &cs);
        const cx<T> K = {cs, sn};
        const J4<T> &B = L.sqB[si * ct + k];
        J4<T> KB;
#pragma unroll
        for (int q = 0; q < 4; ++q) KB.v[q] = cmul(K, B.v[q]);
        L.J[i] = jmul(E, KB);
    }
}

template <typename T> __device__ __forceinline__ T exp_t(double x);
template <> __device__ __forceinline__ float exp_t<float>(double x) { return expf(float(x)); }
template <> __device__ __forceinline__ double exp_t<double>(double x) { return exp(x); }

// Per-antenna Jones of one (source s, antenna a, channel c = c0 + k) with
// every source quantity read from global memory (producer waves of the
// double-buffered kernels, which share no per-source LDS stage).
template <typename T, int UT = 1>
__device__ __forceinline__ J4<T> antenna_jones(const RimeArgs<T> &A, const double *Luvw,
                                               const double *Lpa, const double *Lch, int t, int s,
                                               int a, int c, int k) {
    const uint64_t NT = A.d.ntime.local_size;
    const double l0 = A.lm[2 * uint64_t(s)], m0 = A.lm[2 * uint64_t(s) + 1];
    const double nm1 = sqrt(1.0 - l0 * l0 - m0 * m0) - 1.0;
    const T *S = A.stokes + (uint64_t(s) * NT + t) * 4;
    const T power = exp_t<T>(double(A.alpha[uint64_t(s) * NT + t]) * Lch[CH * k + 4]);
    const J4<T> B = brightness_sqrt_scaled<T>(S[0], S[1], S[2], S[3], power);
    const J4<T> E = beam_jones<T, UT>(A, Lpa, Lch, t, a, c, k, l0, m0);
    const double *x = Luvw + 3 * a;
    const double phi = -A.two_pi_over_c * Lch[CH * k + 3] * (x[2] * nm1 + x[1] * m0 + x[0] * l0);
    T sn, cs;
    PhaseCS<T>::eval(phi, &sn, &cs);
    const cx<T> K = {cs, sn};
    J4<T> KB;
#pragma unroll
    for (int q = 0; q < 4; ++q) KB.v[q] = cmul(K, B.v[q]);
    return jmul(E, KB);
}

// Shape parameters of source s: p0, p1, p2 and type (0 point, 1 gaussian,
// 2 sersic; for sersic p2 = R / (1 - e1^2 - e2^2)).
template <typename T>
__device__ __forceinline__ void source_shape(const RimeArgs<T> &A, int s, T *out) {
    const mbx_dims &d = A.d;
    const int np = dext(d.npsrc), ng = dext(d.ngsrc);
    const uint64_t NG = d.ngsrc.local_size, NS = d.nssrc.local_size;
    T p0 = T(0), p1 = T(0), p2 = T(0), ty = T(0);
    if (s >= np && s < np + ng) {
        const int g = s - np;
        p0 = A.gauss[g]; p1 = A.gauss[g + NG]; p2 = A.gauss[g + 2 * NG]; ty = T(1);
    } else if (s >= np + ng) {
        const int g = s - np - ng;
        p0 = A.sersic[g]; p1 = A.sersic[g + NS];
        p2 = A.sersic[g + 2 * NS] / (T(1) - p0 * p0 - p1 * p1); ty = T(2);
    }
    out[0] = p0; out[1] = p1; out[2] = p2; out[3] = ty;
}

// LDS of the double-buffered (producer / consumer) kernels.
template <typename T> struct PcTile {
    double *uvw;   // [NA][3]
    double *pa;    // [NA][2]
    double *ch;    // [ct][CH]
    T *shape;      // [2][st][4]
    J4<T> *J;      // [2][st][NA][ct]
    double *part;  // [nwaves]
    int jstride;   // st * NA * ct

    __device__ __forceinline__ PcTile(unsigned char *base, int NA, int ct, int st) {
        uvw = reinterpret_cast<double *>(base);
        pa = uvw + 3 * NA;
        ch = pa + 2 * NA;
        shape = reinterpret_cast<T *>(ch + CH * ct);
        // aligned offset from the LDS base (see Tile)
        const size_t off = (size_t(reinterpret_cast<unsigned char *>(shape + 2 * 4 * st) - base) + 15) & ~size_t(15);
        J = reinterpret_cast<J4<T> *>(base + off);
        jstride = st * NA * ct;
        part = reinterpret_cast<double *>(J + 2 * jstride);
    }
};

template <typename T>
static inline size_t pc_lds_bytes(int na, int ct, int st, int nwaves) {
    return sizeof(double) * (5 * size_t(na) + CH * ct) + sizeof(T) * 8 * st + 16 +
           sizeof(J4<T>) * 2 * size_t(st) * na * ct + sizeof(double) * nwaves;
}

// Producer step: fill buffer `buf` with tile [s0, s0 + nst) using the
// producer threads ptid in [0, pth).
template <typename T>
__device__ __forceinline__ void pc_produce(const RimeArgs<T> &A, PcTile<T> &L, int buf, int t,
                                           int c0, int s0, int nst, int ptid, int pth) {
    const int NA = A.d.na.local_size;
    const int ct = A.ct;
    const int nc_ext = dext(A.d.nchan);
    J4<T> *J = L.J + buf * L.jstride;
    T *shape = L.shape + buf * 4 * A.st;
    for (int i = ptid; i < nst; i += pth) source_shape<T>(A, s0 + i, shape + 4 * i);
    const int n1 = nst * NA * ct;
    for (int i = ptid; i < n1; i += pth) {
        const int k = i % ct;
        const int a = (i / ct) % NA;
        const int si = i / (ct * NA);
        const int c = min(c0 + k, nc_ext - 1);
        J[i] = antenna_jones<T>(A, L.uvw, L.pa, L.ch, t, s0 + si, a, c, k);
    }
}

// exp(-r2) and (1 + r2)^(-3/2): float uses the hardware exp2 and
// reciprocal square root (about 1 ulp, well inside the float tolerance),
// float64 the correctly rounded library functions.
__device__ __forceinline__ float gauss_of(float r2) {
    return __builtin_amdgcn_exp2f(-1.4426950408889634f * r2);
}
__device__ __forceinline__ double gauss_of(double r2) { return exp(-r2); }
__device__ __forceinline__ float sersic_of(float r2) {
    const float r = __builtin_amdgcn_rsqf(1.0f + r2);
    return r * r * r;
}
__device__ __forceinline__ double sersic_of(double r2) {
    const double den = 1.0 + r2;
    return 1.0 / (den * sqrt(den));
}

// Gaussian factor from coordinates pre-scaled by kGaussExp2Scale: float
// exp2(-r2) on the hardware (coordinates scaled by sqrt(log2 e)), float64
// exp(-r2).
template <typename T> constexpr T kGaussExp2Scale() { return sizeof(T) == 4 ? T(1.2011224087864498) : T(1); }
__device__ __forceinline__ float gauss_of_scaled(float r2) { return __builtin_amdgcn_exp2f(-r2); }
__device__ __forceinline__ double gauss_of_scaled(double r2) { return exp(-r2); }

template <typename T>
__device__ __forceinline__ T gauss_f(T u, T v, T el, T em, T R, T k) {
    T x = (u * em - v * el) * k * R;
    T y = (u * el + v * em) * k;
    return gauss_of(x * x + y * y);
}

template <typename T>
__device__ __forceinline__ T sersic_f(T u, T v, T e1, T e2, T Rp, T k) {
    // Rp = R / (1 - e1^2 - e2^2), precomputed in phase 0
    T x = (u * (T(1) + e1) + v * e2) * k * Rp;
    T y = (u * e2 + v * (T(1) - e1)) * k * Rp;
    return sersic_of(x * x + y * y);
}

// Epilogue of one visibility (t, b, c): M = G_p V G_q^H, flags,
// model/residual output and the chi^2 term (RimeSumCoherencies.py:318-392).
// Also stores V to ekb when requested. Returns the chi^2 term.
template <typename T>
__device__ __forceinline__ double vis_out(const RimeArgs<T> &A, int t, int b, int c, int p, int q,
                                          const J4<T> &V, bool final_batch) {
    const mbx_dims &d = A.d;
    const uint64_t NA = d.na.local_size, NC = d.nchan.local_size, NBL = d.nbl.local_size;
    const uint64_t vi = (uint64_t(t) * NBL + b) * NC + c;
    if (A.ekb) jstore(A.ekb + vi * 4, V);
    if (!final_batch) return 0.0;
    const uint64_t gp = ((uint64_t(t) * NA + p) * NC + c) * 4;
    const uint64_t gq = ((uint64_t(t) * NA + q) * NC + c) * 4;
    J4<T> M = jmul_h(jmul(jload(A.G + gp), V), jload(A.G + gq));
    J4<T> O = jload(A.obs + vi * 4);
    uchar4 f4 = *reinterpret_cast<const uchar4 *>(A.flag + vi * 4);
    const uint8_t fa[4] = {f4.x, f4.y, f4.z, f4.w};
    J4<T> R;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (fa[k] > 0) { M.v[k] = {T(0), T(0)}; O.v[k] = {T(0), T(0)}; }
        R.v[k] = {O.v[k].re - M.v[k].re, O.v[k].im - M.v[k].im};
    }
    if (A.model) jstore(A.model + vi * 4, (A.options & MBX_RESIDUAL_OUTPUT) ? R : M);
    T term = T(0);
    if (A.options & MBX_APPLY_WEIGHTS) {
        const T *w = A.wv + vi * 4;
#pragma unroll
        for (int k = 0; k < 4; ++k) term += (R.v[k].re * R.v[k].re + R.v[k].im * R.v[k].im) * w[k];
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) term += R.v[k].re * R.v[k].re + R.v[k].im * R.v[k].im;
    }
    if (A.chi2) A.chi2[vi] = term;
    return double(term);
}

// Workgroup sum of a float64 value (wave64 shuffles, then LDS); thread 0
// writes it to partial[slot].
__device__ __forceinline__ void wg_partial(double v, double *part, double *out, int tid, int nth) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((tid & 63) == 0) part[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
        double s = 0.0;
        for (int w = 0; w < (nth + 63) / 64; ++w) s += part[w];
        *out = s;
    }
}

}  // namespace mbx
