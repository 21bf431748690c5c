"""The v4 array and property schema.

Names, shapes, dtypes and the 'default' / 'test' initialisers follow
montblanc/impl/rime/v4/config.py:48-247 (array list A, property list P and
the random test-data generators). The 'test' generators draw from numpy's
global RNG like the reference's do, so callers seed with np.random.seed.
"""

import numpy as np

C = 299792458.0                       # montblanc/__init__.py:52-59
FWHM2INT = 1.0 / np.sqrt(np.log(256.0))  # v4/config.py:51


def random_like(ary):
    """Uniform [0, 1) values shaped like ary; complex gets re and im draws
    (montblanc/util/__init__.py random_like)."""
    if np.iscomplexobj(ary):
        return np.random.random(ary.shape) + 1j * np.random.random(ary.shape)
    return np.random.random(ary.shape)


# ---- 'test' initialisers (v4/config.py:71-115) ------------------------

def rand_uvw(slvr, ary):
    ntime, na = slvr.dim_local_size('ntime', 'na')
    distance = 10
    ant_angles = 2 * np.pi * np.arange(na) / slvr.ft(na)
    time_angle = np.arange(ntime) / slvr.ft(ntime)
    ta = time_angle[:, None] * ant_angles[None, :]
    ary[:, :, 0] = distance * np.sin(ta)
    ary[:, :, 1] = distance * np.sin(ta)
    ary[:, :, 2] = np.random.random(size=(ntime, na)) * 0.1
    ary[:, 0, :] = 0
    return ary


def rand_stokes(slvr, ary):
    I, Q, U, V = ary[:, :, 0], ary[:, :, 1], ary[:, :, 2], ary[:, :, 3]
    noise = np.random.random(size=I.shape) * 0.1
    Q[:] = np.random.random(size=Q.shape) - 0.5
    U[:] = np.random.random(size=U.shape) - 0.5
    V[:] = np.random.random(size=V.shape) - 0.5
    I[:] = np.sqrt(Q ** 2 + U ** 2 + V ** 2 + noise)
    return ary


def rand_gauss_shape(slvr, ary):
    for i in range(3):
        ary[i, :] = np.random.random(size=ary.shape[1])
    return ary


def rand_sersic_shape(slvr, ary):
    ary[0, :] = 0
    ary[1, :] = 0
    ary[2, :] = np.pi / 648000   # 1 arcsec
    return ary


IDENTITY4 = np.array([1, 0, 0, 1])

# name, shape, dtype, default, test, scratch
ARRAYS = [
    ('uvw', ('ntime', 'na', 3), 'ft', 0, rand_uvw, False),
    ('antenna1', ('ntime', 'nbl'), np.int32,
     lambda s, a: s.default_ant_pairs()[0], lambda s, a: s.default_ant_pairs()[0], False),
    ('antenna2', ('ntime', 'nbl'), np.int32,
     lambda s, a: s.default_ant_pairs()[1], lambda s, a: s.default_ant_pairs()[1], False),
    ('frequency', ('nchan',), 'ft',
     lambda s, a: np.linspace(1e9, 2e9, s.dim_local_size('nchan')),
     lambda s, a: np.linspace(1e9, 2e9, s.dim_local_size('nchan')), False),
    ('ref_frequency', ('nchan',), 'ft', 1.4e9, 1.4e9, False),
    ('parallactic_angles', ('ntime', 'na'), 'ft', 0, lambda s, a: random_like(a) * np.pi, False),
    ('point_errors', ('ntime', 'na', 'nchan', 2), 'ft', 0,
     lambda s, a: (random_like(a) - 0.5) * 1e-2, False),
    ('antenna_scaling', ('na', 'nchan', 2), 'ft', 1, lambda s, a: random_like(a), False),
    ('E_beam', ('beam_lw', 'beam_mh', 'beam_nud', 4), 'ct',
     IDENTITY4[None, None, None, :], lambda s, a: random_like(a), False),
    ('G_term', ('ntime', 'na', 'nchan', 4), 'ct',
     IDENTITY4[None, None, None, :], lambda s, a: random_like(a), False),
    ('lm', ('nsrc', 2), 'ft', 0, lambda s, a: (random_like(a) - 0.5) * 1e-1, False),
    ('stokes', ('nsrc', 'ntime', 4), 'ft', np.array([1, 0, 0, 0])[None, None, :], rand_stokes, False),
    ('alpha', ('nsrc', 'ntime'), 'ft', 0.8, lambda s, a: random_like(a) * 0.1, False),
    ('gauss_shape', (3, 'ngsrc'), 'ft', np.array([0, 0, 1])[:, None], rand_gauss_shape, False),
    ('sersic_shape', (3, 'nssrc'), 'ft', np.array([0, 0, 0])[:, None], rand_sersic_shape, False),
    ('flag', ('ntime', 'nbl', 'nchan', 4), np.uint8, 0,
     lambda s, a: np.random.randint(0, 2, size=a.shape), False),
    ('weight_vector', ('ntime', 'nbl', 'nchan', 4), 'ft', 1, lambda s, a: random_like(a), False),
    ('observed_vis', ('ntime', 'nbl', 'nchan', 4), 'ct', 0, lambda s, a: random_like(a), False),
    ('B_sqrt', ('nsrc', 'ntime', 'nchan', 4), 'ct', None, None, True),
    ('jones', ('nsrc', 'ntime', 'na', 'nchan', 4), 'ct', None, None, True),
    ('model_vis', ('ntime', 'nbl', 'nchan', 4), 'ct', None, None, False),
    ('chi_sqrd_result', ('ntime', 'nbl', 'nchan'), 'ft', None, None, False),
]

# name, dtype, default (v4/config.py:54-69)
PROPERTIES = [
    ('gauss_scale', 'ft', FWHM2INT * np.sqrt(2) * np.pi / C),
    ('two_pi_over_c', 'ft', 2 * np.pi / C),
    ('sigma_sqrd', 'ft', 1.0),
    ('X2', 'ft', 0.0),
    ('beam_ll', 'ft', -0.5),
    ('beam_lm', 'ft', -0.5),
    ('beam_lfreq', 'ft', 1.3e9),
    ('beam_ul', 'ft', 0.5),
    ('beam_um', 'ft', 0.5),
    ('beam_ufreq', 'ft', 1.5e9),
]
