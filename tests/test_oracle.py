"""Pin the numpy oracle (oracle/rime_np.py) independently of its own code.

Each oracle function is checked against something that does not retype its
formula: a library routine that the computation reduces to (scipy sqrtm,
scipy map_coordinates, numpy matmul), a closed form of the physics (the
Fourier relation of a point source), invariants the reference's own tests
use (test_rime_v4.py:311-315, 422-529), or brute force on tiny inputs. The
cases are chosen so that a plausible mistake -- a swapped antenna order, a
wrong phase sign, a transposed Jones product, a dropped shape term --
fails one of them.
"""

import cmath
import math

import numpy as np
import pytest
import scipy.linalg
import scipy.ndimage

from oracle import rime_np, synth

RNG = np.random.default_rng(7)


def as22(J):
    return J.reshape(J.shape[:-1] + (2, 2))


def rand_c(*shape):
    return RNG.standard_normal(shape) + 1j * RNG.standard_normal(shape)


# ---- Jones products --------------------------------------------------

def test_jones_mul_matches_matmul():
    A, B = rand_c(5, 7, 4), rand_c(5, 7, 4)
    np.testing.assert_allclose(as22(rime_np.jones_mul(A, B)), as22(A) @ as22(B), rtol=1e-14)
    ref = as22(A) @ np.conj(np.swapaxes(as22(B), -1, -2))
    np.testing.assert_allclose(as22(rime_np.jones_mul_h(A, B)), ref, rtol=1e-14)
    # not symmetric: a transposed operand would fail
    assert not np.allclose(rime_np.jones_mul(A, B), rime_np.jones_mul(B, A))


# ---- brightness ------------------------------------------------------

def positive_stokes(n):
    I = RNG.lognormal(0, 1, n)
    f = RNG.uniform(-0.3, 0.3, (n, 3))
    return np.concatenate([I[:, None], I[:, None] * f], axis=1)


def test_b_sqrt_matches_scipy_sqrtm_and_invariants():
    S = positive_stokes(12).reshape(3, 4, 4)          # (nsrc, ntime, 4)
    alpha = RNG.normal(-0.7, 0.2, (3, 4))
    f = np.linspace(1.3e9, 1.5e9, 5)
    rf = np.full(5, 1.4e9)
    B = rime_np.b_jones(S, alpha, f, rf)
    sq = rime_np.b_sqrt_jones(B)
    # library: principal square root
    for idx in np.ndindex(B.shape[:-1]):
        ref = scipy.linalg.sqrtm(B[idx].reshape(2, 2))
        np.testing.assert_allclose(sq[idx].reshape(2, 2), ref, rtol=1e-12, atol=1e-14)
    # invariants of test_rime_v4.py:311-315
    np.testing.assert_allclose(rime_np.jones_mul(sq, sq), B, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(as22(sq), np.conj(np.swapaxes(as22(sq), -1, -2)), atol=1e-14)


def test_b_jones_spectral_index_closed_form():
    S = np.array([[[2.0, 0.5, 0.25, -0.125]]])
    # f = rf: no scaling; f = 2 rf with alpha = 1: doubled; alpha = -1: halved
    f = np.array([1.4e9, 2.8e9, 2.8e9])
    rf = np.full(3, 1.4e9)
    B1 = rime_np.b_jones(S, np.array([[1.0]]), f, rf)[0, 0]
    want = np.array([2.5, 0.25 - 0.125j, 0.25 + 0.125j, 1.5])
    np.testing.assert_allclose(B1[0], want)
    np.testing.assert_allclose(B1[1], 2 * want)
    Bm = rime_np.b_jones(S, np.array([[-1.0]]), f, rf)[0, 0]
    np.testing.assert_allclose(Bm[2], want / 2)


def test_b_sqrt_zero_matrix_and_negative_det():
    Z = np.zeros((1, 1, 1, 4), np.complex128)
    np.testing.assert_array_equal(rime_np.b_sqrt_jones(Z), Z)
    bad = rime_np.b_jones(np.array([[[1.0, 2.0, 0, 0]]]), np.zeros((1, 1)), np.ones(1), np.ones(1))
    with pytest.raises(AssertionError):
        rime_np.b_sqrt_jones(bad)


# ---- phase -----------------------------------------------------------

def test_k_jones_brute_force_and_unit_modulus():
    uvw = RNG.uniform(-3000, 3000, (2, 3, 3))
    lm = RNG.uniform(-0.01, 0.01, (4, 2))
    f = np.array([1.4e9, 1.41e9])
    K = rime_np.k_jones(uvw, lm, f)
    for s, t, a, c in np.ndindex(K.shape):
        l, m = lm[s]
        n = math.sqrt(1 - l * l - m * m)
        u, v, w = uvw[t, a]
        want = cmath.exp(complex(0, -2 * math.pi * f[c] / 299792458.0 * (u * l + v * m + w * (n - 1))))
        assert abs(K[s, t, a, c] - want) < 1e-9
    np.testing.assert_allclose(np.abs(K), 1.0, rtol=1e-14)
    K0 = rime_np.k_jones(uvw, np.zeros((1, 2)), f)
    np.testing.assert_allclose(K0, 1.0)


def test_point_source_fourier_relation():
    """Unpolarised unit point source, identity beam and gains:
    V_pq = exp(+2 pi i f/C (u_pq l + v_pq m + w_pq (n-1))), u_pq = u_q - u_p.
    Pins the phase sign together with the (p, q) order of J_p J_q^H."""
    na, ntime, nchan = 4, 2, 3
    a = synth.make_inputs(na, ntime, nchan, 1, 0, 0, beam=(5, 5, 4), seed=3)
    a['E_beam'][:] = np.array([1, 0, 0, 1])
    a['G_term'][:] = np.array([1, 0, 0, 1])
    a['stokes'][:] = np.array([1.0, 0, 0, 0])
    a['alpha'][:] = 0.0
    V = rime_np.coherencies(a, 1, 0, 0)
    l, m = a['lm'][0]
    n1 = math.sqrt(1 - l * l - m * m) - 1
    for t in range(ntime):
        for b in range(a['antenna1'].shape[1]):
            p, q = a['antenna1'][t, b], a['antenna2'][t, b]
            u, v, w = a['uvw'][t, q] - a['uvw'][t, p]
            for c in range(nchan):
                ph = 2 * math.pi * a['frequency'][c] / 299792458.0 * (u * l + v * m + w * n1)
                want = cmath.exp(1j * ph)
                np.testing.assert_allclose(V[t, b, c], [want, 0, 0, want], atol=1e-9)


# ---- E beam ------------------------------------------------------------

def beam_case(lw=7, mh=6, nud=5, nsrc=3, ntime=2, na=3, nchan=4, seed=11):
    r = np.random.default_rng(seed)
    lm = r.uniform(-0.3, 0.3, (nsrc, 2))
    pa = r.uniform(0, np.pi, (ntime, na))
    pe = r.uniform(-0.05, 0.05, (ntime, na, nchan, 2))
    sc = r.uniform(0.5, 1.5, (na, nchan, 2))
    f = np.linspace(1.35e9, 1.45e9, nchan)
    E = r.uniform(-1, 1, (lw, mh, nud, 4)) + 1j * r.uniform(-1, 1, (lw, mh, nud, 4))
    ext = (-0.5, -0.4, 1.3e9, 0.5, 0.6, 1.5e9)
    return lm, pa, pe, sc, f, E, ext


def test_e_beam_matches_map_coordinates():
    """The reference's own cross-check (CPUSolver.py:477-500): scipy linear
    interpolation of re, im and |E| separately, then re-normalised."""
    lm, pa, pe, sc, f, E, ext = beam_case()
    got = rime_np.e_beam(lm, pa, pe, sc, f, E, *ext)
    lw, mh, nud = E.shape[:3]
    ll, lmn, lf, ul, um, uf = ext
    for s, t, a, c in np.ndindex(got.shape[:4]):
        ps, pc = math.sin(pa[t, a]), math.cos(pa[t, a])
        l0, m0 = lm[s]
        l = (l0 * pc - m0 * ps + pe[t, a, c, 0]) * sc[a, c, 0]
        m = (l0 * ps + m0 * pc + pe[t, a, c, 1]) * sc[a, c, 1]
        gl = (lw - 1) * (l - ll) / (ul - ll)
        gm = (mh - 1) * (m - lmn) / (um - lmn)
        gc = (nud - 1) * (f[c] - lf) / (uf - lf)
        for p in range(4):
            coords = np.array([[gl], [gm], [gc]])
            re = scipy.ndimage.map_coordinates(E[..., p].real, coords, order=1, mode='nearest')[0]
            im = scipy.ndimage.map_coordinates(E[..., p].imag, coords, order=1, mode='nearest')[0]
            amp = scipy.ndimage.map_coordinates(np.abs(E[..., p]), coords, order=1, mode='nearest')[0]
            want = complex(re, im) * amp / abs(complex(re, im))
            assert abs(got[s, t, a, c, p] - want) < 1e-12


def test_e_beam_constant_cube_and_grid_nodes():
    lm, pa, pe, sc, f, E, ext = beam_case()
    const = np.broadcast_to(np.array([0.3 + 0.4j, -1j, 2.0, 0.1 - 0.1j]), E.shape).copy()
    got = rime_np.e_beam(lm, pa, pe, sc, f, const, *ext)
    np.testing.assert_allclose(got, np.broadcast_to(const[0, 0, 0], got.shape), rtol=1e-13)
    # a source placed exactly on a grid node returns that node's value
    lw, mh, nud = E.shape[:3]
    ll, lmn, lf, ul, um, uf = ext
    il, im_, ic = 2, 3, 1
    l = ll + il * (ul - ll) / (lw - 1)
    m = lmn + im_ * (um - lmn) / (mh - 1)
    fc = lf + ic * (uf - lf) / (nud - 1)
    got = rime_np.e_beam(np.array([[l, m]]), np.zeros((1, 1)), np.zeros((1, 1, 1, 2)),
                         np.ones((1, 1, 2)), np.array([fc]), E, *ext)
    np.testing.assert_allclose(got[0, 0, 0, 0], E[il, im_, ic], rtol=1e-9)


def test_e_beam_degenerate_cube():
    """2x2x2 cube as in test_rime_v4.py:405-407 with every tap clamped."""
    _, _, _, _, f, E, ext = beam_case(lw=2, mh=2, nud=2)
    ll, lmn, _, ul, um, _ = ext
    far = np.array([[10.0, 10.0], [-10.0, 10.0], [10.0, -10.0], [-10.0, -10.0]])
    edge = np.array([[ul, um], [ll, um], [ul, lmn], [ll, lmn]])
    args = (np.zeros((1, 1)), np.zeros((1, 1, f.size, 2)), np.ones((1, f.size, 2)), f, E)
    got = rime_np.e_beam(far, *args, *ext)       # far outside: clamped to the corners
    want = rime_np.e_beam(edge, *args, *ext)
    assert np.all(np.isfinite(got))
    np.testing.assert_allclose(got, want, rtol=1e-13)
    # at a corner only the channel axis interpolates: compare with scipy
    lw, mh, nud = E.shape[:3]
    gc = (nud - 1) * (f - ext[2]) / (ext[5] - ext[2])
    for c in range(f.size):
        for p in range(4):
            co = np.array([[lw - 1], [mh - 1], [gc[c]]])
            re = scipy.ndimage.map_coordinates(E[..., p].real, co, order=1, mode='nearest')[0]
            im = scipy.ndimage.map_coordinates(E[..., p].imag, co, order=1, mode='nearest')[0]
            amp = scipy.ndimage.map_coordinates(np.abs(E[..., p]), co, order=1, mode='nearest')[0]
            z = complex(re, im)
            assert abs(got[0, 0, 0, c, p] - z * amp / abs(z)) < 1e-12


# ---- shapes ------------------------------------------------------------
#
# Both shape factors are checked against what they are meant to be: the
# normalised Fourier transform of an image-plane brightness, at the
# baseline's spatial frequency (u, v) * f / C. The brightness is built on a
# grid from its own definition (a Gaussian given by its full widths at half
# maximum; an exponential disk carried onto the sky by a shear matrix) and
# transformed by a direct sum, so the constants (gauss_scale, 2 pi / C,
# fwhm2int), the (el, em, R) axis convention and the (e1, e2) shear
# convention are all pinned without retyping the oracle's expression.

def _fourier_sum(image, x, ul, vl):
    """Normalised direct Fourier sum of an image on the square grid x * x
    (first axis l, second m) at spatial frequency (ul, vl) wavelengths."""
    ph = np.exp(-2j * np.pi * ul * x)[:, None] * np.exp(-2j * np.pi * vl * x)[None, :]
    return (image * ph).sum() / image.sum()


def _uvw_at(ul, vl, f):
    """One baseline whose (u, v) in metres is (ul, vl) wavelengths at f."""
    return np.array([[[ul * rime_np.C / f, vl * rime_np.C / f, 0.0]]])


def _elliptical_gaussian_image(el, em, R, x):
    """Gaussian brightness with full width at half maximum |(el, em)| along
    the direction (el, em) and R |(el, em)| across it."""
    a = math.hypot(el, em)
    para = np.array([el, em]) / a
    perp = np.array([-para[1], para[0]])
    L, M = np.meshgrid(x, x, indexing='ij')
    xa = L * para[0] + M * para[1]
    xp = L * perp[0] + M * perp[1]
    # exp(-4 ln2 (x / FWHM)^2) is 1/2 at x = FWHM / 2
    return np.exp(-4.0 * math.log(2.0) * ((xa / a) ** 2 + (xp / (R * a)) ** 2))


def test_gaussian_shape_is_fourier_transform_of_elliptical_gaussian():
    f = 1.4e9
    rng = np.random.default_rng(11)
    for el, em, R in ((6e-5, -8e-5, 0.45), (-3e-5, 2e-5, 1.7)):
        a = math.hypot(el, em)
        w = a * max(R, 1.0)
        x = (np.arange(601) - 300) * (5 * w / 300)
        img = _elliptical_gaussian_image(el, em, R, x)
        gs = np.array([[el], [em], [R]])
        for _ in range(5):
            k = rng.uniform(0.1, 2.5) / w
            th = rng.uniform(0, 2 * np.pi)
            ul, vl = k * math.cos(th), k * math.sin(th)
            ft = _fourier_sum(img, x, ul, vl)
            got = rime_np.gaussian_shape(_uvw_at(ul, vl, f), gs, np.array([f]))[0, 0, 0, 0]
            assert abs(ft.imag) < 1e-12
            assert abs(got - ft.real) < 1e-12, (el, em, R, ul, vl, got, ft.real)


def _rot(th):
    return np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])


def test_gaussian_shape_rotation_covariance_and_axes():
    """(u, v) and (el, em) rotated together leave the factor unchanged; at
    R = 1 it depends on |(el, em)| |(u, v)| only; along (el, em) the width is
    set by |(el, em)| and across it by R |(el, em)|."""
    rng = np.random.default_rng(12)
    f = np.array([1.4e9, 1.42e9])
    uvw = rng.uniform(-3000, 3000, (2, 5, 3))
    gs = np.array([[7e-5], [-4e-5], [0.6]])
    g0 = rime_np.gaussian_shape(uvw, gs, f)
    for th in (0.3, 1.9, -2.4):
        Rm = _rot(th)
        uvr = uvw.copy()
        uvr[..., :2] = uvw[..., :2] @ Rm.T
        gsr = gs.copy()
        gsr[:2, 0] = Rm @ gs[:2, 0]
        np.testing.assert_allclose(rime_np.gaussian_shape(uvr, gsr, f), g0, rtol=1e-12)
    # R = 1: a round Gaussian, the same for every direction of (el, em) of
    # the same length
    round1 = rime_np.gaussian_shape(uvw, np.array([[7e-5], [-4e-5], [1.0]]), f)
    round2 = rime_np.gaussian_shape(uvw, np.array([[0.0], [math.hypot(7e-5, 4e-5)], [1.0]]), f)
    np.testing.assert_allclose(round1, round2, rtol=1e-12)
    # a baseline along (el, em) does not see R; one across it sees R * |a|
    el, em = 7e-5, -4e-5
    para = np.array([el, em]) / math.hypot(el, em)
    perp = np.array([-para[1], para[0]])
    b = 2000.0
    for R in (0.2, 3.0):
        shape = np.array([[el], [em], [R]])
        along = rime_np.gaussian_shape(np.array([[[*(b * para), 0.0]]]), shape, f)
        across = rime_np.gaussian_shape(np.array([[[*(b * perp), 0.0]]]), shape, f)
        ref_round = rime_np.gaussian_shape(np.array([[[*(b * para), 0.0]]]),
                                           np.array([[el], [em], [1.0]]), f)
        np.testing.assert_allclose(along, ref_round, rtol=1e-12)
        # across: the round factor at R times the baseline length
        across_ref = rime_np.gaussian_shape(np.array([[[*(R * b * para), 0.0]]]),
                                            np.array([[el], [em], [1.0]]), f)
        np.testing.assert_allclose(across, across_ref, rtol=1e-12)


def test_gaussian_shape_limits():
    uvw_bl = RNG.uniform(-6000, 6000, (2, 3, 3))
    f = np.array([1.4e9, 1.42e9])
    gs = np.array([[1e-4, 0.0], [-5e-5, 0.0], [0.7, 0.0]])   # second source: zero shape
    g = rime_np.gaussian_shape(uvw_bl, gs, f)
    np.testing.assert_allclose(g[1], 1.0)
    # zero baseline: every shape gives 1
    np.testing.assert_allclose(rime_np.gaussian_shape(0 * uvw_bl, gs, f), 1.0)
    # symmetric in (u, v) -> (-u, -v): the baseline direction does not matter
    np.testing.assert_allclose(g, rime_np.gaussian_shape(-uvw_bl, gs, f), rtol=1e-14)
    # the constant: fwhm2int*sqrt(2)*pi/C with fwhm2int = 1/sqrt(ln 256)
    k = rime_np.GAUSS_SCALE
    assert abs(k - math.sqrt(2) * math.pi / 299792458.0 / math.sqrt(8 * math.log(2))) < 1e-25


def _shear(e1, e2):
    return np.array([[1.0 + e1, e2], [e2, 1.0 - e1]])


def _sheared_disk_image(e1, e2, R, x):
    """A round exponential disk exp(-|y| / R') carried onto the sky by the
    symmetric shear x = S y, S = [[1 + e1, e2], [e2, 1 - e1]], with
    R' = R / det S."""
    S = _shear(e1, e2)
    Rp = R / np.linalg.det(S)
    Si = np.linalg.inv(S)
    L, M = np.meshgrid(x, x, indexing='ij')
    y0 = Si[0, 0] * L + Si[0, 1] * M
    y1 = Si[1, 0] * L + Si[1, 1] * M
    return np.exp(-np.hypot(y0, y1) / Rp), Rp


def test_sersic_shape_is_fourier_transform_of_sheared_exponential_disk():
    """The exponential disk's transform has a cusp-limited grid error of
    ~2e-6 on this grid (40 R' each side, 2500 points); a flipped or dropped
    shear term moves the factor by > 3e-2 at these frequencies."""
    f = 1.4e9
    rng = np.random.default_rng(13)
    for e1, e2, R in ((0.3, -0.2, 5e-5), (-0.15, 0.35, 3e-5)):
        Rp = R / np.linalg.det(_shear(e1, e2))
        x = (np.arange(2500) - 1249.5) * (80 * Rp * 1.4 / 2500)
        img, _ = _sheared_disk_image(e1, e2, R, x)
        ss = np.array([[e1], [e2], [R]])
        for _ in range(3):
            k = rng.uniform(0.2, 2.0) / (2 * np.pi * Rp)
            th = rng.uniform(0, 2 * np.pi)
            ul, vl = k * math.cos(th), k * math.sin(th)
            ft = _fourier_sum(img, x, ul, vl)
            got = rime_np.sersic_shape(_uvw_at(ul, vl, f), ss, np.array([f]))[0, 0, 0, 0]
            assert abs(got - ft.real) < 1e-5, (e1, e2, R, ul, vl, got, ft.real)


def test_sersic_shape_spin2_rotation_and_reflection():
    """Rotating (u, v) by th and (e1, e2) as a spin-2 quantity by 2 th leaves
    the factor unchanged, and so does reflecting v -> -v with e2 -> -e2."""
    rng = np.random.default_rng(14)
    f = np.array([1.4e9, 1.41e9])
    uvw = rng.uniform(-3000, 3000, (2, 5, 3))
    ss = np.array([[0.25], [-0.3], [4e-5]])
    s0 = rime_np.sersic_shape(uvw, ss, f)
    for th in (0.4, -1.3, 2.8):
        uvr = uvw.copy()
        uvr[..., :2] = uvw[..., :2] @ _rot(th).T
        ssr = ss.copy()
        ssr[:2, 0] = _rot(2 * th) @ ss[:2, 0]
        np.testing.assert_allclose(rime_np.sersic_shape(uvr, ssr, f), s0, rtol=1e-12)
    ref = uvw.copy()
    ref[..., 1] *= -1
    ssf = ss.copy()
    ssf[1] *= -1
    np.testing.assert_allclose(rime_np.sersic_shape(ref, ssf, f), s0, rtol=1e-12)


def test_sersic_shape_limits():
    uvw_bl = RNG.uniform(-6000, 6000, (2, 3, 3))
    f = np.array([1.4e9])
    ss = np.array([[0.0, 0.0, 0.2], [0.0, 0.0, 0.1], [0.0, 1e-4, 1e-4]])
    g = rime_np.sersic_shape(uvw_bl, ss, f)
    np.testing.assert_allclose(g[0], 1.0)
    np.testing.assert_allclose(rime_np.sersic_shape(0 * uvw_bl, ss, f), 1.0)
    # e1 = e2 = 0: the round disk's transform, (1 + (2 pi R k)^2)^(-3/2)
    k = 2 * math.pi / 299792458.0 * f[0] * 1e-4
    for t, b in np.ndindex(2, 3):
        q2 = (uvw_bl[t, b, 0] ** 2 + uvw_bl[t, b, 1] ** 2) * k * k
        assert abs(g[1, t, b, 0] - (1 + q2) ** -1.5) < 1e-14
    np.testing.assert_allclose(g, rime_np.sersic_shape(-uvw_bl, ss, f), rtol=1e-14)


# ---- per-baseline products, gains, flags, chi^2 -------------------------

def brute_force_rime(a, npsrc, ngsrc, nssrc, residual, use_w):
    """Explicit loops with numpy matrices on every (t, b, c)."""
    J = rime_np.per_antenna_jones(a)
    ntime, nbl = a['antenna1'].shape
    nchan = a['frequency'].shape[0]
    uvw_bl = rime_np.baseline_uvw(a['uvw'], a['antenna1'], a['antenna2'])
    g = rime_np.gaussian_shape(uvw_bl, a['gauss_shape'], a['frequency']) if ngsrc else None
    s = rime_np.sersic_shape(uvw_bl, a['sersic_shape'], a['frequency']) if nssrc else None
    model = np.zeros((ntime, nbl, nchan, 4), complex)
    terms = np.zeros((ntime, nbl, nchan))
    for t, b, c in np.ndindex(ntime, nbl, nchan):
        p, q = a['antenna1'][t, b], a['antenna2'][t, b]
        V = np.zeros((2, 2), complex)
        for k in range(npsrc + ngsrc + nssrc):
            f = 1.0
            if npsrc <= k < npsrc + ngsrc:
                f = g[k - npsrc, t, b, c]
            elif k >= npsrc + ngsrc:
                f = s[k - npsrc - ngsrc, t, b, c]
            Jp = J[k, t, p, c].reshape(2, 2)
            Jq = J[k, t, q, c].reshape(2, 2)
            V += f * Jp @ Jq.conj().T
        Gp = a['G_term'][t, p, c].reshape(2, 2)
        Gq = a['G_term'][t, q, c].reshape(2, 2)
        M = (Gp @ V @ Gq.conj().T).reshape(4)
        O = a['observed_vis'][t, b, c].copy()
        fl = a['flag'][t, b, c] > 0
        M[fl] = 0
        O[fl] = 0
        R = O - M
        model[t, b, c] = R if residual else M
        w = a['weight_vector'][t, b, c] if use_w else 1.0
        terms[t, b, c] = np.sum(w * (R.real ** 2 + R.imag ** 2))
    return model, terms


@pytest.mark.parametrize('residual,use_w', [(False, False), (True, True)])
def test_rime_brute_force(residual, use_w):
    a = synth.make_inputs(4, 2, 3, 2, 2, 1, beam=(6, 6, 4), seed=5, flag_fraction=0.5)
    a['weight_vector'] = RNG.uniform(0.5, 2, a['weight_vector'].shape)
    a['sersic_shape'][:] = np.array([[0.1], [-0.2], [3e-5]])
    r = rime_np.rime(a, 2, 2, 1, residual=residual, use_weight_vector=use_w, sigma_sqrd=2.0,
                     src_chunk=2)
    model, terms = brute_force_rime(a, 2, 2, 1, residual, use_w)
    np.testing.assert_allclose(r['model_vis'], model, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(r['chi_sqrd_result'], terms, rtol=1e-12, atol=1e-13)
    want = terms.sum() if use_w else terms.sum() / 2.0
    assert abs(r['X2'] - want) <= 1e-12 * abs(want)


def test_sqrt_multiply_identity():
    """(J_p sqrtB)(J_q sqrtB)^H == J_p B J_q^H (test_rime_v4.py:422-503)."""
    S = positive_stokes(1).reshape(1, 1, 4)
    B = rime_np.b_jones(S, np.zeros((1, 1)), np.ones(1), np.ones(1))[0, 0, 0]
    sq = rime_np.b_sqrt_jones(B[None, None, None])[0, 0, 0]
    Jp, Jq = rand_c(4), rand_c(4)
    lhs = rime_np.jones_mul_h(rime_np.jones_mul(Jp, sq), rime_np.jones_mul(Jq, sq))
    rhs = rime_np.jones_mul_h(rime_np.jones_mul(Jp, B), Jq)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-12)


def test_residual_and_model_modes_agree_on_chi2():
    a = synth.make_inputs(5, 2, 3, 2, 1, 0, beam=(6, 6, 4), seed=9, flag_fraction=0.3)
    m = rime_np.rime(a, 2, 1, 0, residual=False)
    r = rime_np.rime(a, 2, 1, 0, residual=True)
    assert abs(m['X2'] - r['X2']) < 1e-12 * m['X2']
    np.testing.assert_allclose(r['model_vis'], (a['observed_vis'] - m['model_vis']) * (a['flag'] == 0))


def test_chunking_is_invisible():
    a = synth.make_inputs(5, 3, 4, 5, 3, 2, beam=(6, 6, 4), seed=13)
    r1 = rime_np.coherencies(a, 5, 3, 2, src_chunk=1, time_chunk=1)
    r2 = rime_np.coherencies(a, 5, 3, 2, src_chunk=64)
    np.testing.assert_allclose(r1, r2, rtol=1e-13, atol=1e-13)


def _fails(check):
    try:
        check()
    except AssertionError:
        return True
    return False


def test_planted_mistakes_are_caught(monkeypatch):
    """Plant plausible mistakes in the oracle (on the CPU, in this process
    only) and check that the pinning checks above fail on each."""
    # 1. a dropped term: the w (n - 1) part of the phase
    good_k = rime_np.k_jones

    def k_no_w(uvw, lm, frequency):
        flat = uvw.copy()
        flat[..., 2] = 0.0
        return good_k(flat, lm, frequency)
    monkeypatch.setattr(rime_np, 'k_jones', k_no_w)
    assert _fails(test_k_jones_brute_force_and_unit_modulus)
    monkeypatch.setattr(rime_np, 'k_jones', good_k)

    # 2. a wrong sign: sqrt(B) = (B - s 1) / t
    good_sqrt = rime_np.b_sqrt_jones

    def sqrt_sign(B):
        r = good_sqrt(B)
        det = B[..., 0] * B[..., 3] - B[..., 1] * B[..., 2]
        s_ = np.sqrt(det)
        t_ = np.sqrt(B[..., 0] + B[..., 3] + 2 * s_)
        t_ = np.where(t_ == 0, 1.0, t_)
        r = B / t_[..., None]
        r[..., 0] -= s_ / t_
        r[..., 3] -= s_ / t_
        return r
    monkeypatch.setattr(rime_np, 'b_sqrt_jones', sqrt_sign)
    assert _fails(test_b_sqrt_matches_scipy_sqrtm_and_invariants)
    monkeypatch.setattr(rime_np, 'b_sqrt_jones', good_sqrt)

    # 3. a transposed operand: J_p J_q^T instead of J_p J_q^H
    good_mul_h = rime_np.jones_mul_h

    def mul_t(A, B):
        return good_mul_h(A, np.conj(B))
    monkeypatch.setattr(rime_np, 'jones_mul_h', mul_t)
    assert _fails(lambda: test_rime_brute_force(False, False))
    monkeypatch.setattr(rime_np, 'jones_mul_h', good_mul_h)

    # 4. swapped antennas (V_qp instead of V_pq): the Fourier relation fails
    good_cs = rime_np.coherency_sum

    def swapped(jones, antenna1, antenna2, *args, **kw):
        return good_cs(jones, antenna2, antenna1, *args, **kw)
    monkeypatch.setattr(rime_np, 'coherency_sum', swapped)
    assert _fails(test_point_source_fourier_relation)
    monkeypatch.setattr(rime_np, 'coherency_sum', good_cs)

    # 5. el and em swapped in the gaussian shape
    good_g = rime_np.gaussian_shape

    def g_swapped(uvw_bl, gauss_shape, frequency, *args, **kw):
        return good_g(uvw_bl, np.asarray(gauss_shape)[[1, 0, 2]], frequency, *args, **kw)
    monkeypatch.setattr(rime_np, 'gaussian_shape', g_swapped)
    assert _fails(test_gaussian_shape_is_fourier_transform_of_elliptical_gaussian)
    assert _fails(test_gaussian_shape_rotation_covariance_and_axes)
    monkeypatch.setattr(rime_np, 'gaussian_shape', good_g)

    # 6. the sign of e2 flipped in u1: u1 = u (1 + e1) - v e2
    good_s = rime_np.sersic_shape

    def s_flipped(uvw_bl, sersic_shape, frequency, two_pi_over_c=rime_np.TWO_PI_OVER_C):
        e1, e2, R = (np.asarray(x, np.float64)[:, None, None, None] for x in sersic_shape)
        u, v = uvw_bl[None, ..., 0, None], uvw_bl[None, ..., 1, None]
        s = two_pi_over_c * np.asarray(frequency) * R / (1 - e1 * e1 - e2 * e2)
        den = 1 + ((u * (1 + e1) - v * e2) * s) ** 2 + ((u * e2 + v * (1 - e1)) * s) ** 2
        return den ** -1.5
    monkeypatch.setattr(rime_np, 'sersic_shape', s_flipped)
    assert _fails(test_sersic_shape_is_fourier_transform_of_sheared_exponential_disk)
    assert _fails(test_sersic_shape_spin2_rotation_and_reflection)
    monkeypatch.setattr(rime_np, 'sersic_shape', good_s)

    # 7. a dropped term: R' = R without the 1 / (1 - e1^2 - e2^2) scaling
    def s_no_det(uvw_bl, sersic_shape, frequency, *args, **kw):
        e1, e2, R = np.asarray(sersic_shape, np.float64)
        return good_s(uvw_bl, np.stack([e1, e2, R * (1 - e1 * e1 - e2 * e2)]), frequency,
                      *args, **kw)
    monkeypatch.setattr(rime_np, 'sersic_shape', s_no_det)
    assert _fails(test_sersic_shape_is_fourier_transform_of_sheared_exponential_disk)
    monkeypatch.setattr(rime_np, 'sersic_shape', good_s)

    # and the unplanted oracle passes all of them
    test_k_jones_brute_force_and_unit_modulus()
    test_b_sqrt_matches_scipy_sqrtm_and_invariants()
    test_rime_brute_force(False, False)
    test_point_source_fourier_relation()
    test_gaussian_shape_is_fourier_transform_of_elliptical_gaussian()
    test_gaussian_shape_rotation_covariance_and_axes()
    test_sersic_shape_is_fourier_transform_of_sheared_exponential_disk()
    test_sersic_shape_spin2_rotation_and_reflection()


def test_parallel_oracle_matches_single_process():
    """The multi-core CPU baseline (oracle/parallel.py): source ranges over
    two worker processes, including ranges that straddle the type
    boundaries, give the single-process oracle's outputs."""
    from oracle import parallel
    counts = (5, 4, 3)
    a = synth.make_inputs(6, 2, 3, *counts, beam=(9, 8, 6), seed=3, flag_fraction=0.2)
    assert parallel.source_ranges(12, 5) == [(0, 2), (2, 5), (5, 7), (7, 10), (10, 12)]
    ref = rime_np.rime(a, *counts)
    p = parallel.ParallelRime(a, *counts, processes=2)
    try:
        out = p()
    finally:
        p.close()
    np.testing.assert_allclose(out['model_vis'], ref['model_vis'], rtol=1e-12, atol=1e-12)
    assert abs(out['X2'] - ref['X2']) <= 1e-12 * ref['X2']
