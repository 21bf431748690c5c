"""Algorithmic work model of the RIME forward and its roofline peaks.

Flop counts (add and multiply count 1 each, an FMA counts 2,
transcendentals count 0 and are reported separately):

  per source-visibility term (one source, time, baseline, channel, all
  four correlations):
    point     64   J_p J_q^H (8 complex multiplies at 6 + 4 complex adds
                   at 2 = 56) + accumulation (8)
    gaussian  83   point + shape factor (19) + 1 exp
    sersic    86   point + shape factor (22) + 1 sqrt, 1 divide
  per source-antenna term (one source, time, antenna, channel)   426
    phase (6, float64) + sincos; E beam (~340: coordinates, 8 taps x 4
    correlations of complex and amplitude interpolation, renormalisation);
    K sqrt(B) (24); E (K sqrt(B)) (56)
  per visibility (time, baseline, channel)                        143
    G_p M G_q^H (112), residual (8), weighted squares and sum (23)

FLOPs_fwd = 64 N_sv,pt + 83 N_sv,g + 86 N_sv,s + 426 N_sa + 143 N_vis.

HBM bytes of the fused forward are the compulsory array traffic: every
input read once, model_vis and chi_sqrd_result written once.
"""

PEAK_TFLOPS = {'f32': 157.3, 'f64': 78.6}    # MI355X dense FP32 / FP64 (matrix cores; packed-vector FP32 and vector FP64 are the same figures)
PEAK_TFLOPS_UNPACKED_F32 = 78.6              # FP32 without packed-math instructions
PEAK_HBM_GBS = 8000.0

FLOP_SV = {'point': 64, 'gaussian': 83, 'sersic': 86}
FLOP_SA = 426
FLOP_VIS = 143


def terms(na, ntime, nchan, nsrc, nbl=None):
    """Source-visibility terms (s, t, b, c)."""
    nbl = na * (na - 1) // 2 if nbl is None else nbl
    return nsrc * ntime * nbl * nchan


def forward_flops(na, ntime, nchan, npsrc, ngsrc, nssrc, nbl=None):
    nbl = na * (na - 1) // 2 if nbl is None else nbl
    nvis = ntime * nbl * nchan
    nsrc = npsrc + ngsrc + nssrc
    return (FLOP_SV['point'] * npsrc * nvis + FLOP_SV['gaussian'] * ngsrc * nvis
            + FLOP_SV['sersic'] * nssrc * nvis + FLOP_SA * nsrc * ntime * na * nchan
            + FLOP_VIS * nvis)


def forward_hbm_bytes(na, ntime, nchan, nsrc, ngsrc, nssrc, beam, itemsize, nbl=None):
    """Compulsory HBM bytes of one fused forward (inputs read once, outputs
    written once). itemsize = 4 (float) or 8 (double)."""
    nbl = na * (na - 1) // 2 if nbl is None else nbl
    nvis = ntime * nbl * nchan
    f, c = itemsize, 2 * itemsize
    vis_bytes = nvis * (4 * c      # observed_vis read
                        + 4 * c    # model_vis write
                        + 4        # flag
                        + f)       # chi_sqrd_result write
    ant_bytes = ntime * na * (3 * f + f) + ntime * na * nchan * (2 * f + 4 * c) + na * nchan * 2 * f
    src_bytes = nsrc * (2 * f + ntime * 5 * f) + 3 * f * (ngsrc + nssrc)
    beam_bytes = beam[0] * beam[1] * beam[2] * 4 * c
    misc = 2 * ntime * nbl * 4 + 2 * nchan * f
    return vis_bytes + ant_bytes + src_bytes + beam_bytes + misc
