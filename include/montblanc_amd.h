/*
 * montblanc_amd.h -- C-ABI of the MI355X (gfx950) v4 RIME hot path.
 *
 * Every entry point is stateless: dimensions come by pointer to an
 * mbx_dims (the counterpart of the reference's __constant__
 * rime_const_data struct, montblanc/util/const_data.py:39-169), arrays are
 * caller-owned DEVICE pointers in the reference's v4 layouts
 * (montblanc/impl/rime/v4/config.py:131-247), complex values are
 * interleaved (re, im), and work is enqueued on the given hipStream_t
 * (passed as void* so the header needs no HIP include).
 *
 * Return codes: 0 = success, < 0 = argument validation failure
 * (MBX_E_*), > 0 = a hipError_t from the launch.
 * mbx_last_error() returns a static string describing the last failure on
 * the calling thread.
 *
 * dtype suffix: _f32 = float/complex64 arrays (phase argument and beam
 * grid coordinates evaluated in float64), _f64 = double/complex128.
 */
#ifndef MONTBLANC_AMD_H
#define MONTBLANC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MBX_ABI_VERSION 3

enum {
    MBX_OK = 0,
    MBX_E_NULL = -1,        /* required pointer is NULL              */
    MBX_E_DIMS = -2,        /* inconsistent or zero dimensions       */
    MBX_E_EXTENT = -3,      /* lower/upper extents outside the sizes */
    MBX_E_WORKSPACE = -4,   /* workspace too small                   */
    MBX_E_UNSUPPORTED = -5, /* npol != 4 or an unsupported option    */
    MBX_E_NO_DEVICE = -6    /* no HIP device visible                 */
};

/* One dimension: montblanc/util/const_data.py:51-86 (dim_field). */
typedef struct mbx_dim {
    uint32_t global_size;
    uint32_t local_size;
    uint32_t lower_extent;
    uint32_t upper_extent;
} mbx_dim;

/* All dimensions the path reads (const_data.py:118-145; the nsrc extents
 * are the sums of the per-type extents, as RimeConstStruct.update sets). */
typedef struct mbx_dims {
    mbx_dim ntime, na, nbl, nchan, npol, npolchan;
    mbx_dim nsrc, npsrc, ngsrc, nssrc;
    mbx_dim beam_lw, beam_mh, beam_nud;
} mbx_dims;

/* Beam cube extents: solver properties beam_{ll,lm,lfreq,ul,um,ufreq}
 * (montblanc/impl/rime/v4/config.py:62-68). */
typedef struct mbx_beam_extents {
    double ll, lm, lfreq, ul, um, ufreq;
} mbx_beam_extents;

/* Options of the coherency sum: the stamped kernel variants
 * apply_weights x vis_output (RimeSumCoherencies.py:405-464). */
enum {
    MBX_APPLY_WEIGHTS = 1,    /* weight_vector multiplies chi^2 terms      */
    MBX_RESIDUAL_OUTPUT = 2   /* write obs - model instead of the model     */
};

/* Constants the kernels take from solver properties
 * (v4/config.py:55-57). */
typedef struct mbx_consts {
    double gauss_scale;     /* fwhm2int*sqrt(2)*pi/C */
    double two_pi_over_c;   /* 2*pi/C                */
    double sigma_sqrd;      /* chi^2 divisor when weights are off */
} mbx_consts;

int mbx_abi_version(void);
const char *mbx_last_error(void);
/* Number of visible HIP devices (0 when none); never fails. */
int mbx_device_count(void);

/* ---- Stage entry points (one per v4 kernel node) -------------------- */

/* B_sqrt[s,t,c,:] = sqrt(B(stokes[s,t], alpha[s,t]) * (f/rf)^alpha).
 * Replaces rime_jones_B_sqrt_{float,double}
 * (montblanc/impl/rime/v4/gpu/RimeBSqrt.py:130-148). */
int mbx_b_sqrt_f32(const mbx_dims *d, const float *stokes, const float *alpha,
                   const float *frequency, const float *ref_frequency,
                   void *B_sqrt, void *stream);
int mbx_b_sqrt_f64(const mbx_dims *d, const double *stokes, const double *alpha,
                   const double *frequency, const double *ref_frequency,
                   void *B_sqrt, void *stream);

/* jones[s,t,a,c,:] = E (trilinear beam-cube lookup, amplitude-normalised).
 * Replaces rime_jones_E_beam_{float,double} (RimeEBeam.py:348-366). */
int mbx_e_beam_f32(const mbx_dims *d, const float *lm, const float *parallactic_angles,
                   const float *point_errors, const float *antenna_scaling,
                   const float *frequency, const void *E_beam, void *jones,
                   const mbx_beam_extents *beam, void *stream);
int mbx_e_beam_f64(const mbx_dims *d, const double *lm, const double *parallactic_angles,
                   const double *point_errors, const double *antenna_scaling,
                   const double *frequency, const void *E_beam, void *jones,
                   const mbx_beam_extents *beam, void *stream);

/* jones <- jones . (K * B_sqrt), K = exp(-2 pi i f/C (ul + vm + w(n-1))).
 * Replaces rime_jones_EKBSqrt_{float,double} (RimeEKBSqrt.py:182-196). */
int mbx_ekb_sqrt_f32(const mbx_dims *d, const float *uvw, const float *lm,
                     const float *frequency, const void *B_sqrt, void *jones,
                     const mbx_consts *k, void *stream);
int mbx_ekb_sqrt_f64(const mbx_dims *d, const double *uvw, const double *lm,
                     const double *frequency, const void *B_sqrt, void *jones,
                     const mbx_consts *k, void *stream);

/* Per-baseline coherency sum with shapes, gains, flags, residual and
 * per-visibility chi^2 terms, honouring the source-batch extents
 * (read the accumulator when nsrc.lower_extent > 0, write the partial sum
 * and stop when nsrc.upper_extent < nsrc.global_size).
 * Replaces rime_sum_coherencies (RimeSumCoherencies.py:405-429).
 * gauss_shape / sersic_shape may be NULL when that type is empty;
 * weight_vector may be NULL without MBX_APPLY_WEIGHTS. */
int mbx_sum_coherencies_f32(const mbx_dims *d, const float *uvw,
                            const float *gauss_shape, const float *sersic_shape,
                            const float *frequency, const int32_t *antenna1,
                            const int32_t *antenna2, const void *jones,
                            const uint8_t *flag, const float *weight_vector,
                            const void *observed_vis, const void *G_term,
                            void *model_vis, float *chi_sqrd_result,
                            const mbx_consts *k, int options, void *stream);
int mbx_sum_coherencies_f64(const mbx_dims *d, const double *uvw,
                            const double *gauss_shape, const double *sersic_shape,
                            const double *frequency, const int32_t *antenna1,
                            const int32_t *antenna2, const void *jones,
                            const uint8_t *flag, const double *weight_vector,
                            const void *observed_vis, const void *G_term,
                            void *model_vis, double *chi_sqrd_result,
                            const mbx_consts *k, int options, void *stream);

/* X2 = sum of chi_sqrd_result over the active (time, baseline, channel)
 * extents, divided by sigma_sqrd unless MBX_APPLY_WEIGHTS; accumulated in
 * float64 and written to the device double *X2.
 * Replaces pycuda gpuarray.sum + set_X2 (RimeSumCoherencies.py:531-536)
 * and the v5 RimeReduction node (v5/gpu/RimeReduction.py:46-61).
 * workspace: >= mbx_chi2_workspace_bytes() bytes of device memory. */
size_t mbx_chi2_workspace_bytes(void);
int mbx_chi2_reduce_f32(const mbx_dims *d, const float *chi_sqrd_result, double *X2,
                        const mbx_consts *k, int options, void *workspace,
                        size_t workspace_bytes, void *stream);
int mbx_chi2_reduce_f64(const mbx_dims *d, const double *chi_sqrd_result, double *X2,
                        const mbx_consts *k, int options, void *workspace,
                        size_t workspace_bytes, void *stream);

/* ---- Fused forward (no reference counterpart: replaces the whole v4
 * pipeline RimeBSqrt -> RimeEBeam -> RimeEKBSqrt -> RimeSumCoherencies ->
 * reduction, v4/RimeSolver.py:36-40, without the jones / B_sqrt scratch) */

/* Inputs of the whole chain, all device pointers in v4 layouts. */
typedef struct mbx_inputs {
    const void *uvw, *lm, *stokes, *alpha, *frequency, *ref_frequency;
    const void *parallactic_angles, *point_errors, *antenna_scaling, *E_beam;
    const void *gauss_shape, *sersic_shape;     /* may be NULL if empty */
    const int32_t *antenna1, *antenna2;
    const uint8_t *flag;
    const void *weight_vector;                  /* may be NULL w/o weights */
    const void *observed_vis, *G_term;
} mbx_inputs;

/* Outputs. ekb_vis (optional, may be NULL) receives the source sum before
 * gains, (ntime, nbl, nchan, 4) complex: the adjoint needs it. */
typedef struct mbx_outputs {
    void *model_vis;           /* (ntime, nbl, nchan, 4) complex          */
    void *chi_sqrd_result;     /* (ntime, nbl, nchan) real                */
    double *X2;                /* device scalar                           */
    void *ekb_vis;             /* optional                                */
} mbx_outputs;

/* Launch plan built on the host from the antenna layout (no reference
 * counterpart: the reference pairs antennas per baseline in each thread,
 * RimeSumCoherencies.py:154-157).
 *
 * Pair blocks: the baselines grouped into blocks of pb x qb antenna
 * pairs (p0 + i, q0 + j) so that one thread reuses each per-antenna Jones
 * pb (or qb) times. blocks is a DEVICE array of nblocks entries of
 * (2 + pb*qb) int32: p0, q0, then the baseline index of slot i*qb + j or -1.
 * Supported shapes: 1x1, 2x2 and 2x4. blocks == NULL (or a NULL plan)
 * selects the generic kernel, which reads antenna1/2 per time step and so
 * also handles layouts that change with time.
 *
 * Point sources on the matrix cores: with baseline_map (a DEVICE copy of
 * mbx_build_baseline_map's output for the same, time-invariant layout)
 * and point_gemm != 0, the point sources' coherency sums are computed as
 * a Hermitian matrix product per (time, channel) for na <= 64; from 40
 * antennas (float) or 48 (float64) the gaussian and sersic sources also
 * run on the matrix cores (in float that pass applies the epilogue too),
 * otherwise the selected kernel adds the other source types. */
typedef struct mbx_pair_plan {
    const int32_t *blocks;
    int32_t nblocks, pb, qb;
    /* Kernel form: >0 = double-buffered kernel with that many producer
     * waves building the next source tile while the consumer waves
     * accumulate; 0 = the same with an automatic producer count;
     * <0 = single-role kernel (every wave does both phases). */
    int32_t producer_waves;
    const int32_t *baseline_map;
    int32_t point_gemm;
} mbx_pair_plan;

/* Host-side plan builder (no device work). antenna1/antenna2 are HOST
 * arrays (ntime x nbl int32). Writes at most `capacity` int32 into `out`
 * and returns the number of blocks, or MBX_E_UNSUPPORTED when the layout
 * differs between time steps (use the generic kernel), MBX_E_DIMS for an
 * antenna index outside [0, na), MBX_E_WORKSPACE when capacity is short.
 * mbx_pair_plan_capacity() is a sufficient capacity. No reference
 * counterpart: the reference pairs antennas per baseline in each thread
 * (RimeSumCoherencies.py:154-157). */
int64_t mbx_pair_plan_capacity(int32_t nbl, int32_t pb, int32_t qb);
int64_t mbx_build_pair_plan(const int32_t *antenna1, const int32_t *antenna2, int32_t ntime,
                            int32_t nbl, int32_t na, int32_t pb, int32_t qb, int32_t *out,
                            int64_t capacity);

size_t mbx_forward_workspace_bytes(const mbx_dims *d, int is_double);
int mbx_forward_f32(const mbx_dims *d, const mbx_inputs *in, const mbx_outputs *out,
                    const mbx_beam_extents *beam, const mbx_consts *k,
                    const mbx_pair_plan *plan, int options,
                    void *workspace, size_t workspace_bytes, void *stream);
int mbx_forward_f64(const mbx_dims *d, const mbx_inputs *in, const mbx_outputs *out,
                    const mbx_beam_extents *beam, const mbx_consts *k,
                    const mbx_pair_plan *plan, int options,
                    void *workspace, size_t workspace_bytes, void *stream);

/* The fused forward without a caller workspace. No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments and results as mbx_forward_f32/f64 minus the
 * workspace, which here is a library-held device array of
 * MBX_STATIC_SCRATCH_BYTES per device (each device that runs the library
 * keeps it until the process ends); dimensions whose
 * mbx_forward_workspace_bytes exceed it are rejected with
 * MBX_E_WORKSPACE before anything runs. The stream's device must be the
 * current one. Allocates nothing. Calls that overlap on one device share
 * the array and race: serialise them (one stream, or events). */
#define MBX_STATIC_SCRATCH_BYTES (64u << 20)
int mbx_forward_static_f32(const mbx_dims *d, const mbx_inputs *in, const mbx_outputs *out,
                           const mbx_beam_extents *beam, const mbx_consts *k,
                           const mbx_pair_plan *plan, int options, void *stream);
int mbx_forward_static_f64(const mbx_dims *d, const mbx_inputs *in, const mbx_outputs *out,
                           const mbx_beam_extents *beam, const mbx_consts *k,
                           const mbx_pair_plan *plan, int options, void *stream);

/* ---- Adjoint: d chi^2 / d(parameters), recomputing the Jones terms ----
 * No reference counterpart (the reference has no gradient). */

/* Gradient outputs, float64 whatever the input dtype (accumulated across
 * workgroups). Real parameters get d chi^2/d theta; the complex gains get
 * d chi^2/dRe + i d chi^2/dIm (torch's convention for complex tensors).
 * All are multiplied by grad_scale (the upstream d L / d chi^2). Any
 * pointer may be NULL to skip that gradient. */
typedef struct mbx_grads {
    double *lm;            /* (nsrc, 2)                      */
    double *stokes;        /* (nsrc, ntime, 4)               */
    double *alpha;         /* (nsrc, ntime)                  */
    double *gauss_shape;   /* (3, ngsrc)                     */
    double *sersic_shape;  /* (3, nssrc)                     */
    void *G_term;          /* (ntime, na, nchan, 4) complex128 */
} mbx_grads;

/* Host helper: (na x na) int32 map from an antenna pair to its baseline:
 * out[p*na + q] = b for baseline b = (p, q) with p = antenna1,
 * b | MBX_BL_REVERSED at [q*na + p], b | MBX_BL_AUTO for p == q, -1 where
 * no baseline. Fails (MBX_E_UNSUPPORTED) when the layout changes with
 * time or a pair occurs twice: the adjoint needs one baseline per pair. */
#define MBX_BL_REVERSED (1 << 30)
#define MBX_BL_AUTO (1 << 29)
int mbx_build_baseline_map(const int32_t *antenna1, const int32_t *antenna2, int32_t ntime,
                           int32_t nbl, int32_t na, int32_t *out);

/* Backward pass after a forward over the same inputs: ekb_vis is the
 * forward's pre-gain source sum (mbx_outputs.ekb_vis), baseline_map a
 * DEVICE copy of mbx_build_baseline_map's output. The source extents must
 * cover every source (lower 0, upper global): the adjoint needs the full
 * model. */
size_t mbx_adjoint_workspace_bytes(const mbx_dims *d, int is_double);
int mbx_adjoint_f32(const mbx_dims *d, const mbx_inputs *in, const void *ekb_vis,
                    const int32_t *baseline_map, const mbx_grads *grads,
                    const mbx_beam_extents *beam, const mbx_consts *k, int options,
                    double grad_scale, void *workspace, size_t workspace_bytes, void *stream);
int mbx_adjoint_f64(const mbx_dims *d, const mbx_inputs *in, const void *ekb_vis,
                    const int32_t *baseline_map, const mbx_grads *grads,
                    const mbx_beam_extents *beam, const mbx_consts *k, int options,
                    double grad_scale, void *workspace, size_t workspace_bytes, void *stream);

/* ---- Batched evaluation of many sky-model samples ---------------------
 * No reference counterpart. K samples, each with its own sources, against
 * ONE shared set of visibility-space inputs (uvw, antennas, frequencies,
 * beam, gains, flags, weights, observed visibilities) in one persistent
 * launch. The source arrays of all samples are concatenated in the
 * mbx_inputs (lm, stokes, alpha: rows; gauss_shape, sersic_shape: columns)
 * and mbx_dims' source sizes are the totals. All arrays below are DEVICE
 * int32 arrays of length nsamples (src_offset: nsamples + 1). */
typedef struct mbx_batch {
    int32_t nsamples;
    const int32_t *src_offset;     /* first row of sample k in lm/stokes/alpha  */
    const int32_t *npsrc, *ngsrc, *nssrc;   /* per-sample type counts          */
    const int32_t *gauss_offset;   /* first gauss_shape column of sample k     */
    const int32_t *sersic_offset;  /* first sersic_shape column of sample k    */
    const int32_t *order;          /* processing order of the samples (a
                                      permutation; largest first balances)    */
} mbx_batch;

/* Strategy and launch shape; zero fields take the defaults (also what
 * the autotuner in montblanc_amd/autotune.py varies).
 * strategy 0: one persistent launch on the vector units, work items of
 *   (sample, time, channel tile) taken from a queue.
 * strategy 1 (float, at most 64 antennas, baseline_map required): the
 *   matrix-core passes of the single-model forward and adjoint, each pass
 *   one launch over a group of samples_per_group samples (one grid of
 *   (time, channel) workgroups per sample, samples in `order`); the
 *   workspace then holds two visibility arrays per sample of a group. */
typedef struct mbx_batch_tuning {
    int32_t workgroups;          /* 0: persistent workgroups (default 2 per CU) */
    int32_t channels_per_item;   /* 0: channels per work item (default 1)       */
    int32_t sources_per_tile;    /* 0: forward source tile (default: LDS bound) */
    int32_t strategy;            /* 0 or 1, above                               */
    int32_t samples_per_group;   /* 1: samples per group of launches (> 0)      */
} mbx_batch_tuning;

/* X2[k] = chi^2 of sample k (device float64 [nsamples]). With grads
 * (lm/stokes/alpha/gauss_shape/sersic_shape shaped like the concatenated
 * inputs, float64; G_term must be NULL) the parameter gradients of every
 * sample's chi^2 are produced in the same call; baseline_map as for
 * mbx_adjoint_*. Model visibilities are not written. */
size_t mbx_batched_workspace_bytes(const mbx_dims *d, int is_double, const mbx_batch_tuning *tune);
int mbx_batched_f32(const mbx_dims *d, const mbx_inputs *in, const mbx_batch *batch, double *X2,
                    const mbx_grads *grads, const int32_t *baseline_map,
                    const mbx_beam_extents *beam, const mbx_consts *k,
                    const mbx_batch_tuning *tune, int options, void *workspace,
                    size_t workspace_bytes, void *stream);
int mbx_batched_f64(const mbx_dims *d, const mbx_inputs *in, const mbx_batch *batch, double *X2,
                    const mbx_grads *grads, const int32_t *baseline_map,
                    const mbx_beam_extents *beam, const mbx_consts *k,
                    const mbx_batch_tuning *tune, int options, void *workspace,
                    size_t workspace_bytes, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MONTBLANC_AMD_H */
