"""Load tests/golden/rime_small.npz (data only, no pickles)."""

import os

import numpy as np

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'rime_small.npz')
GRAD_NAMES = ('lm', 'stokes', 'alpha', 'gauss_shape', 'sersic_shape', 'G_term')


def load():
    z = np.load(PATH)
    a = {k: z[k] for k in z.files if not k.startswith('expect_') and k != 'counts'}
    for k in ('beam_ll', 'beam_lm', 'beam_lfreq', 'beam_ul', 'beam_um', 'beam_ufreq'):
        a[k] = float(a[k])
    expect = {k[len('expect_'):]: z[k] for k in z.files if k.startswith('expect_')}
    return a, tuple(int(c) for c in z['counts']), expect
