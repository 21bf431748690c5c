"""Seeded synthetic inputs for the RIME workloads.

TEST INFRASTRUCTURE ONLY (tests/, smoke(), bench.py's cpu_baseline leg).

Distributions: uvw ~ U(+-3000 m) per (time, antenna, component);
frequency = linspace(1.40, 1.42 GHz), ref frequency 1.4 GHz;
l, m ~ U(+-0.01); I ~ lognormal(0, 1) per source, Q, U, V = I * U(+-0.1),
alpha ~ N(-0.7, 0.2), both repeated over time; gaussian el, em, R ~ U(+-1e-4);
sersic e1 = e2 = 0, scale ~ U(1e-5, 1e-4); beam cube re, im ~ U(0, 1) with
l, m extents +-0.02 and frequency extents 1.39-1.43 GHz; parallactic angle
~ U(0, pi); pointing error ~ U(+-5e-3); antenna scaling ~ U(0, 1);
G = identity + 0.01 * CN(0, 1); weights 1, sigma^2 1; flags 0 or 50 %;
observed = model + CN(0, 0.1) when a model is supplied, else CN(0, 0.1).
"""

import numpy as np

from . import rime_np

WORKLOADS = {
    # name: na, ntime, nchan, npsrc, ngsrc, nssrc, beam (lw, mh, nud)
    'cpu_parity_14ant': dict(na=14, ntime=10, nchan=16, npsrc=10, ngsrc=10, nssrc=0,
                             beam=(33, 33, 16)),
    'forward_64ant': dict(na=64, ntime=100, nchan=64, npsrc=1000, ngsrc=500, nssrc=100,
                          beam=(65, 65, 32)),
    'gradient_27ant': dict(na=27, ntime=200, nchan=128, npsrc=300, ngsrc=200, nssrc=0,
                           beam=(33, 33, 16)),
    'batched_64ant': dict(na=64, ntime=50, nchan=32, npsrc=0, ngsrc=0, nssrc=0,
                          beam=(65, 65, 32)),
    'large_64ant': dict(na=64, ntime=500, nchan=256, npsrc=10000, ngsrc=2000, nssrc=0,
                        beam=(129, 129, 64)),
}


def make_inputs(na, ntime, nchan, npsrc, ngsrc, nssrc, beam=(33, 33, 16), seed=1000,
                flag_fraction=0.0, auto_correlations=False, observed='noise', vis_arrays=True):
    """Return a dict of float64/complex128 numpy arrays in the v4 schema.

    observed: 'noise' -> CN(0, 0.1) only; 'model' -> oracle model + noise
    (only sensible for small shapes). vis_arrays=False leaves out the
    (ntime, nbl, nchan, ...) arrays (observed_vis, weight_vector, flag) for
    callers that generate them on the device.
    """
    rng = np.random.default_rng(seed)
    nsrc = npsrc + ngsrc + nssrc
    a1, a2 = rime_np.default_ant_pairs(na, ntime, auto_correlations=auto_correlations)
    nbl = a1.shape[1]
    lw, mh, nud = beam
    a = {}
    a['uvw'] = rng.uniform(-3000.0, 3000.0, (ntime, na, 3))
    a['antenna1'], a['antenna2'] = a1, a2
    a['frequency'] = np.linspace(1.40e9, 1.42e9, nchan)
    a['ref_frequency'] = np.full(nchan, 1.4e9)
    a['lm'] = rng.uniform(-0.01, 0.01, (nsrc, 2))
    I = rng.lognormal(0.0, 1.0, nsrc)
    frac = rng.uniform(-0.1, 0.1, (nsrc, 3))
    st = np.empty((nsrc, 4))
    st[:, 0] = I
    st[:, 1:] = I[:, None] * frac
    a['stokes'] = np.repeat(st[:, None, :], ntime, axis=1).copy()
    a['alpha'] = np.repeat(rng.normal(-0.7, 0.2, nsrc)[:, None], ntime, axis=1).copy()
    a['gauss_shape'] = rng.uniform(-1e-4, 1e-4, (3, ngsrc))
    ss = np.zeros((3, nssrc))
    ss[2] = rng.uniform(1e-5, 1e-4, nssrc)
    a['sersic_shape'] = ss
    a['E_beam'] = rng.uniform(0, 1, (lw, mh, nud, 4)) + 1j * rng.uniform(0, 1, (lw, mh, nud, 4))
    a['parallactic_angles'] = rng.uniform(0, np.pi, (ntime, na))
    a['point_errors'] = rng.uniform(-5e-3, 5e-3, (ntime, na, nchan, 2))
    a['antenna_scaling'] = rng.uniform(0, 1, (na, nchan, 2))
    G = np.zeros((ntime, na, nchan, 4), np.complex128)
    G[..., 0] = 1.0
    G[..., 3] = 1.0
    G += 0.01 * (rng.standard_normal(G.shape) + 1j * rng.standard_normal(G.shape))
    a['G_term'] = G
    a['beam_ll'], a['beam_lm'], a['beam_lfreq'] = -0.02, -0.02, 1.39e9
    a['beam_ul'], a['beam_um'], a['beam_ufreq'] = 0.02, 0.02, 1.43e9
    a['gauss_scale'] = rime_np.GAUSS_SCALE
    a['two_pi_over_c'] = rime_np.TWO_PI_OVER_C
    a['sigma_sqrd'] = 1.0
    a['_counts'] = (npsrc, ngsrc, nssrc)
    if not vis_arrays:
        return a
    a['weight_vector'] = np.ones((ntime, nbl, nchan, 4))
    if flag_fraction > 0:
        a['flag'] = (rng.uniform(0, 1, (ntime, nbl, nchan, 4)) < flag_fraction).astype(np.uint8)
    else:
        a['flag'] = np.zeros((ntime, nbl, nchan, 4), np.uint8)
    noise = 0.1 * (rng.standard_normal((ntime, nbl, nchan, 4))
                   + 1j * rng.standard_normal((ntime, nbl, nchan, 4)))
    if observed == 'model':
        clean = dict(a)
        clean['flag'] = np.zeros_like(a['flag'])
        clean['observed_vis'] = np.zeros((ntime, nbl, nchan, 4), np.complex128)
        model = rime_np.rime(clean, npsrc, ngsrc, nssrc)['model_vis']
        a['observed_vis'] = model + noise
    else:
        a['observed_vis'] = noise
    return a


def as_dtype(a, dtype):
    """Cast float arrays to float32/float64 (complex accordingly)."""
    ft = np.dtype(dtype)
    ct = np.complex64 if ft == np.float32 else np.complex128
    out = {}
    for k, v in a.items():
        if isinstance(v, np.ndarray):
            if v.dtype.kind == 'f':
                v = v.astype(ft)
            elif v.dtype.kind == 'c':
                v = v.astype(ct)
        out[k] = v
    return out


def upcast(a):
    """float32 inputs -> float64 with identical values (for the oracle)."""
    return as_dtype(a, np.float64)
