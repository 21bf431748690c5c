"""numpy float64 restatement of the reference's v4 CPU RIME.

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker, never by the product path.

Each function restates one ``compute_*`` method of
montblanc/impl/rime/v4/cpu/CPUSolver.py (line numbers cited per function),
with explicit arguments instead of a solver object and numpy in place of
numexpr. Everything is evaluated in float64/complex128 whatever the dtype of
the inputs (float32 inputs are up-cast value for value), so the same oracle
judges both the float and the double GPU paths.

Layout conventions (identical to the reference's v4 schema,
montblanc/impl/rime/v4/config.py:131-247):
  uvw (ntime, na, 3)             lm (nsrc, 2)
  stokes (nsrc, ntime, 4)        alpha (nsrc, ntime)
  frequency, ref_frequency (nchan,)
  parallactic_angles (ntime, na) point_errors (ntime, na, nchan, 2)
  antenna_scaling (na, nchan, 2) E_beam (beam_lw, beam_mh, beam_nud, 4)
  G_term (ntime, na, nchan, 4)   gauss_shape (3, ngsrc)  sersic_shape (3, nssrc)
  antenna1, antenna2 (ntime, nbl) int32
  flag (ntime, nbl, nchan, 4) u8 weight_vector (ntime, nbl, nchan, 4)
  observed_vis, model_vis (ntime, nbl, nchan, 4)
  chi_sqrd_result (ntime, nbl, nchan)
2x2 Jones matrices are stored row-major in the last axis: [XX, XY, YX, YY].
Sources are ordered point, then gaussian, then sersic
(montblanc/src_types.py:40-44).

Pinning (what checks each function independently of this file) is listed in
tests/test_oracle.py; every function here is pinned there.
"""

import numpy as np

# montblanc/__init__.py:52-59
C = 299792458.0
# montblanc/impl/rime/v4/config.py:51,55-56
FWHM2INT = 1.0 / np.sqrt(np.log(256.0))
GAUSS_SCALE = FWHM2INT * np.sqrt(2.0) * np.pi / C
TWO_PI_OVER_C = 2.0 * np.pi / C


def _f64(a):
    return np.asarray(a, dtype=np.float64)


def _c128(a):
    return np.asarray(a, dtype=np.complex128)


def default_ant_pairs(na, ntime, nbl=None, auto_correlations=False):
    """Per-time antenna pairs from the upper triangle.

    montblanc/impl/rime/v4/ant_pairs.py:23-37 and
    montblanc/solvers/rime_solver.py:92-99 (k=0 with auto-correlations).
    """
    a1, a2 = np.triu_indices(na, 0 if auto_correlations else 1)
    if nbl is not None:
        a1, a2 = a1[:nbl], a2[:nbl]
    return (np.tile(a1, (ntime, 1)).astype(np.int32),
            np.tile(a2, (ntime, 1)).astype(np.int32))


def baseline_uvw(uvw, antenna1, antenna2):
    """(ntime, nbl, 3) baseline coordinates uvw[q] - uvw[p].

    CPUSolver.py:69-79 (u = uq - up, per baseline from per antenna).
    """
    uvw = _f64(uvw)
    t = np.arange(uvw.shape[0])[:, None]
    return uvw[t, antenna2] - uvw[t, antenna1]


def gaussian_shape(uvw_bl, gauss_shape, frequency, gauss_scale=GAUSS_SCALE):
    """(ngsrc, ntime, nbl, nchan) gaussian visibility factors.

    CPUSolver.py:59-103: u1 = u*em - v*el, v1 = u*el + v*em,
    exp(-((u1*s*R)^2 + (v1*s)^2)) with s = gauss_scale * frequency.
    """
    el, em, R = (_f64(x)[:, None, None] for x in gauss_shape)
    u, v = uvw_bl[None, :, :, 0], uvw_bl[None, :, :, 1]
    u1 = u * em - v * el
    v1 = u * el + v * em
    s = gauss_scale * _f64(frequency)
    x = u1[..., None] * s * R[..., None]
    y = v1[..., None] * s
    return np.exp(-(x * x + y * y))


def sersic_shape(uvw_bl, sersic_shape, frequency, two_pi_over_c=TWO_PI_OVER_C):
    """(nssrc, ntime, nbl, nchan) sersic (exponential disk) factors.

    CPUSolver.py:105-160: u1 = u(1+e1) + v e2, v1 = u e2 + v(1-e1),
    den = 1 + (u1 s R')^2 + (v1 s R')^2 with R' = R/(1-e1^2-e2^2),
    factor = 1/(den*sqrt(den)).
    """
    e1, e2, R = (_f64(x) for x in sersic_shape)
    Rp = (R / (1.0 - e1 * e1 - e2 * e2))[:, None, None, None]
    e1, e2 = e1[:, None, None], e2[:, None, None]
    u, v = uvw_bl[None, :, :, 0], uvw_bl[None, :, :, 1]
    u1 = u * (1.0 + e1) + v * e2
    v1 = u * e2 + v * (1.0 - e1)
    s = two_pi_over_c * _f64(frequency)
    a = u1[..., None] * s * Rp
    b = v1[..., None] * s * Rp
    den = 1.0 + a * a + b * b
    return 1.0 / (den * np.sqrt(den))


_sersic_factor = sersic_shape


def k_jones(uvw, lm, frequency):
    """(nsrc, ntime, na, nchan) scalar phase term.

    CPUSolver.py:162-195: n = sqrt(1 - l^2 - m^2) - 1,
    K = exp(-2*pi*i*(u l + v m + w n) * f / C).
    """
    uvw, lm, f = _f64(uvw), _f64(lm), _f64(frequency)
    l, m = lm[:, 0], lm[:, 1]
    n = np.sqrt(1.0 - l * l - m * m) - 1.0
    p = (n[:, None, None] * uvw[None, :, :, 2]
         + m[:, None, None] * uvw[None, :, :, 1]
         + l[:, None, None] * uvw[None, :, :, 0])
    return np.exp(-2j * np.pi * p[..., None] * f / C)


def b_jones(stokes, alpha, frequency, ref_frequency):
    """(nsrc, ntime, nchan, 4) brightness matrix with spectral index.

    CPUSolver.py:235-265: B = [[I+Q, U+iV], [U-iV, I-Q]] * (f/rf)^alpha.
    """
    S = _f64(stokes)
    I, Q, U, V = S[..., 0], S[..., 1], S[..., 2], S[..., 3]
    B = np.stack([I + Q, U + 1j * V, U - 1j * V, I - Q], axis=-1)
    power = (_f64(frequency) / _f64(ref_frequency))[None, None, :] \
        ** _f64(alpha)[:, :, None]
    return B[:, :, None, :] * power[..., None]


def b_sqrt_jones(B):
    """Square root of (nsrc, ntime, nchan, 4) brightness matrices.

    CPUSolver.py:267-317: s = sqrt(det), t = sqrt(trace + 2s),
    sqrt(B) = (B + s*1)/t, with t = 1 where s == t == 0.
    Raises AssertionError on negative trace or determinant, as the
    reference does (CPUSolver.py:291-294).
    """
    B = _c128(B).copy()
    trace = (B[..., 0] + B[..., 3]).real
    det = (B[..., 0] * B[..., 3] - B[..., 1] * B[..., 2]).real
    assert np.all(trace >= 0.0), 'Negative brightness matrix trace'
    assert np.all(det >= 0.0), 'Negative brightness matrix determinant'
    s = np.sqrt(det)
    t = np.sqrt(trace + 2.0 * s)
    t[np.logical_and(s == 0.0, t == 0.0)] = 1.0
    B[..., 0] += s
    B[..., 3] += s
    B /= t[..., None]
    return B


def beam_grid_channel(frequency, beam_nud, beam_lfreq, beam_ufreq):
    """Clamped channel grid position, its floor/ceil taps and fraction.

    CPUSolver.py:408-414.
    """
    v = (beam_nud - 1) * (_f64(frequency) - beam_lfreq) / (beam_ufreq - beam_lfreq)
    v = np.clip(v, 0.0, beam_nud - 1)
    g0 = np.floor(v)
    g1 = np.minimum(g0 + 1.0, beam_nud - 1)
    return g0, g1, v - g0


def e_beam(lm, parallactic_angles, point_errors, antenna_scaling, frequency,
           E_beam, beam_ll, beam_lm, beam_lfreq, beam_ul, beam_um, beam_ufreq):
    """(nsrc, ntime, na, nchan, 4) direction-dependent beam Jones.

    CPUSolver.py:319-475: rotate (l, m) by the parallactic angle, add the
    pointing error, scale per antenna, trilinearly interpolate the complex
    cube and, separately, the cube's amplitudes, then E = sum*|.|sum/|sum|
    (normalisation 1 where 1/|sum| is not finite).
    """
    lm = _f64(lm)
    E_beam = _c128(E_beam)
    lw, mh, nud = E_beam.shape[:3]
    pa = _f64(parallactic_angles)
    sint, cost = np.sin(pa), np.cos(pa)
    l0 = lm[:, 0][:, None, None]
    m0 = lm[:, 1][:, None, None]
    l = l0 * cost[None] - m0 * sint[None]
    m = l0 * sint[None] + m0 * cost[None]
    pe = _f64(point_errors)
    l = l[..., None] + pe[None, ..., 0]
    m = m[..., None] + pe[None, ..., 1]
    sc = _f64(antenna_scaling)
    l = l * sc[None, None, :, :, 0]
    m = m * sc[None, None, :, :, 1]

    vl = np.clip((lw - 1) * (l - beam_ll) / (beam_ul - beam_ll), 0.0, lw - 1)
    gl0 = np.floor(vl)
    gl1 = np.minimum(gl0 + 1.0, lw - 1)
    ld = vl - gl0
    vm = np.clip((mh - 1) * (m - beam_lm) / (beam_um - beam_lm), 0.0, mh - 1)
    gm0 = np.floor(vm)
    gm1 = np.minimum(gm0 + 1.0, mh - 1)
    md = vm - gm0
    gc0, gc1, chd = beam_grid_channel(frequency, nud, beam_lfreq, beam_ufreq)
    chd = chd[None, None, None, :]

    absE = np.abs(E_beam)
    acc = np.zeros(l.shape + (4,), np.complex128)
    amp = np.zeros(l.shape + (4,), np.float64)
    for gl, wl in ((gl0, 1.0 - ld), (gl1, ld)):
        il = gl.astype(np.int64)
        for gm, wm in ((gm0, 1.0 - md), (gm1, md)):
            im = gm.astype(np.int64)
            for gc, wc in ((gc0, 1.0 - chd), (gc1, chd)):
                ic = gc.astype(np.int64)[None, None, None, :]
                w = (wl * wm * wc)[..., None]
                acc += w * E_beam[il, im, ic]
                amp += w * absE[il, im, ic]
    with np.errstate(divide='ignore', invalid='ignore'):
        norm = 1.0 / np.abs(acc)
    norm[~np.isfinite(norm)] = 1.0
    return acc * norm * amp


def jones_mul(A, B):
    """Row-major 2x2 product A.B on the last axis (CPUSolver.py:518-541)."""
    a0, a1, a2, a3 = A[..., 0], A[..., 1], A[..., 2], A[..., 3]
    b0, b1, b2, b3 = B[..., 0], B[..., 1], B[..., 2], B[..., 3]
    return np.stack([a0 * b0 + a1 * b2, a0 * b1 + a1 * b3,
                     a2 * b0 + a3 * b2, a2 * b1 + a3 * b3], axis=-1)


def jones_mul_h(A, B):
    """Row-major 2x2 product A.B^H on the last axis (CPUSolver.py:529-531)."""
    b0, b1, b2, b3 = (np.conj(B[..., i]) for i in range(4))
    a0, a1, a2, a3 = A[..., 0], A[..., 1], A[..., 2], A[..., 3]
    return np.stack([a0 * b0 + a1 * b1, a0 * b2 + a1 * b3,
                     a2 * b0 + a3 * b1, a2 * b2 + a3 * b3], axis=-1)


def ekb_sqrt_jones(E, K, Bsqrt):
    """Per-antenna Jones J = E.(K*sqrt(B)), (nsrc, ntime, na, nchan, 4).

    CPUSolver.py:217-233 (K*sqrt(B)) and :543-560 (E times that).
    """
    kb = K[..., None] * Bsqrt[:, :, None, :, :]
    return jones_mul(E, kb)


def coherency_sum(jones, antenna1, antenna2, npsrc, ngsrc, nssrc,
                  uvw=None, gauss_shape=None, sersic_shape=None, frequency=None,
                  gauss_scale=GAUSS_SCALE, two_pi_over_c=TWO_PI_OVER_C):
    """Sum over sources of f_s * J_ps . J_qs^H, (ntime, nbl, nchan, 4).

    CPUSolver.py:562-634 (per-baseline products with the gaussian and
    sersic shapes multiplied in, then summed over sources). The antenna
    pairs come from the antenna1/antenna2 arrays, as the GPU kernel reads
    them (RimeSumCoherencies.py:155-157).
    """
    jones = _c128(jones)
    nsrc = jones.shape[0]
    assert nsrc == npsrc + ngsrc + nssrc
    t = np.arange(jones.shape[1])[:, None]
    per_bl = jones_mul_h(jones[:, t, antenna1], jones[:, t, antenna2])
    if ngsrc or nssrc:
        uvw_bl = baseline_uvw(uvw, antenna1, antenna2)
    if ngsrc:
        g = gaussian_shape(uvw_bl, _f64(gauss_shape)[:, :ngsrc], frequency, gauss_scale)
        per_bl[npsrc:npsrc + ngsrc] *= g[..., None]
    if nssrc:
        s = _sersic_factor(uvw_bl, _f64(sersic_shape)[:, :nssrc], frequency, two_pi_over_c)
        per_bl[npsrc + ngsrc:] *= s[..., None]
    return per_bl.sum(axis=0)


def gekb_vis(ekb_vis, G_term, antenna1, antenna2, flag, observed_vis, residual):
    """Apply per-antenna gains and flags; model or residual output.

    CPUSolver.py:636-676: M = G_p . V . G_q^H; residual output is
    (obs - M) with flagged correlations zeroed, model output is M with
    flagged correlations zeroed.
    """
    G = _c128(G_term)
    t = np.arange(G.shape[0])[:, None]
    M = jones_mul_h(jones_mul(G[t, antenna1], _c128(ekb_vis)), G[t, antenna2])
    unflagged = np.asarray(flag) == 0
    if residual:
        return (_c128(observed_vis) - M) * unflagged
    M[~unflagged] = 0
    return M


def chi_sqrd_sum_terms(vis, observed_vis, flag, weight_vector, residual,
                       use_weight_vector):
    """(ntime, nbl, nchan) chi-squared terms. CPUSolver.py:678-722."""
    vis = _c128(vis)
    if residual:
        d = vis
    else:
        d = (_c128(observed_vis) - vis) * (np.asarray(flag) == 0)
    re2, im2 = d.real ** 2, d.imag ** 2
    if use_weight_vector:
        w = _f64(weight_vector)
        re2, im2 = re2 * w, im2 * w
    return re2.sum(axis=-1) + im2.sum(axis=-1)


def chi_sqrd(terms, sigma_sqrd, use_weight_vector):
    """Scalar chi-squared. CPUSolver.py:724-737."""
    total = np.sum(terms, dtype=np.float64)
    return float(total if use_weight_vector else total / sigma_sqrd)


BEAM_KEYS = ('beam_ll', 'beam_lm', 'beam_lfreq', 'beam_ul', 'beam_um', 'beam_ufreq')


def per_antenna_jones(a, t0=0, t1=None, s0=0, s1=None):
    """J = E.K.sqrt(B) for a time slice [t0,t1) and source slice [s0,s1).

    ``a`` is a dict keyed by the v4 schema names plus the beam extents.
    """
    t1 = a['uvw'].shape[0] if t1 is None else t1
    s1 = a['lm'].shape[0] if s1 is None else s1
    lm = a['lm'][s0:s1]
    E = e_beam(lm, a['parallactic_angles'][t0:t1], a['point_errors'][t0:t1],
               a['antenna_scaling'], a['frequency'], a['E_beam'],
               *(float(a[k]) for k in BEAM_KEYS))
    K = k_jones(a['uvw'][t0:t1], lm, a['frequency'])
    Bs = b_sqrt_jones(b_jones(a['stokes'][s0:s1, t0:t1], a['alpha'][s0:s1, t0:t1],
                              a['frequency'], a['ref_frequency']))
    return ekb_sqrt_jones(E, K, Bs)


def source_type_slices(npsrc, ngsrc, nssrc):
    """Global source index ranges of each type (montblanc/src_types.py:40-44)."""
    return ((0, npsrc), (npsrc, npsrc + ngsrc), (npsrc + ngsrc, npsrc + ngsrc + nssrc))


def coherencies(a, npsrc, ngsrc, nssrc, src_chunk=64, time_chunk=None):
    """Sum over all sources of f_s J_p J_q^H, (ntime, nbl, nchan, 4).

    Chunked over time and sources so that the per-baseline intermediate
    (the reference materialises (nsrc, ntime, nbl, nchan, 4) at once,
    CPUSolver.py:562-634) stays bounded.
    """
    ntime = a['uvw'].shape[0]
    nbl = a['antenna1'].shape[1]
    nchan = a['frequency'].shape[0]
    time_chunk = ntime if time_chunk is None else time_chunk
    out = np.zeros((ntime, nbl, nchan, 4), np.complex128)
    gscale = float(a.get('gauss_scale', GAUSS_SCALE))
    sscale = float(a.get('two_pi_over_c', TWO_PI_OVER_C))
    for t0 in range(0, ntime, time_chunk):
        t1 = min(ntime, t0 + time_chunk)
        a1, a2 = a['antenna1'][t0:t1], a['antenna2'][t0:t1]
        uvw = a['uvw'][t0:t1]
        for kind, (b, e) in enumerate(source_type_slices(npsrc, ngsrc, nssrc)):
            for s0 in range(b, e, src_chunk):
                s1 = min(e, s0 + src_chunk)
                J = per_antenna_jones(a, t0, t1, s0, s1)
                n = s1 - s0
                counts = [0, 0, 0]
                counts[kind] = n
                gs = ss = None
                if kind == 1:
                    gs = _f64(a['gauss_shape'])[:, s0 - b:s1 - b]
                if kind == 2:
                    ss = _f64(a['sersic_shape'])[:, s0 - b:s1 - b]
                out[t0:t1] += coherency_sum(J, a1, a2, *counts, uvw=uvw,
                                            gauss_shape=gs, sersic_shape=ss,
                                            frequency=a['frequency'],
                                            gauss_scale=gscale, two_pi_over_c=sscale)
    return out


def rime(a, npsrc, ngsrc, nssrc, residual=False, use_weight_vector=False,
         sigma_sqrd=1.0, src_chunk=64, time_chunk=None):
    """The composed v4 CPU chain, as montblanc/tests/test_rime_v4.py:205-222
    composes it (CPUSolver.solve() itself is broken at CPUSolver.py:744).

    Returns dict(ekb_vis, model_vis, chi_sqrd_result, X2).
    """
    V = coherencies(a, npsrc, ngsrc, nssrc, src_chunk, time_chunk)
    vis = gekb_vis(V, a['G_term'], a['antenna1'], a['antenna2'], a['flag'],
                   a['observed_vis'], residual)
    terms = chi_sqrd_sum_terms(vis, a['observed_vis'], a['flag'],
                               a['weight_vector'], residual, use_weight_vector)
    return dict(ekb_vis=V, model_vis=vis, chi_sqrd_result=terms,
                X2=chi_sqrd(terms, sigma_sqrd, use_weight_vector))
