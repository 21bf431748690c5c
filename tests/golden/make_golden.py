"""Write tests/golden/rime_small.npz: seeded inputs of a small v4 problem
and the float64 oracle's outputs and gradients for them.

The reference holds no stored RIME values (its tests draw fresh random
data each run), so these fixtures come from the oracle, pinned
independently in tests/test_oracle.py and tests/test_oracle_grad.py. They
freeze that pinned state: a later change to the oracle or to the HIP path
that moves any value shows up against them.

Run from the repository root:  python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import rime_np, rime_torch, synth  # noqa: E402

COUNTS = (3, 2, 1)
SHAPE = dict(na=6, ntime=3, nchan=4, beam=(7, 6, 5))
SEED = 2024


def inputs():
    a = synth.make_inputs(SHAPE['na'], SHAPE['ntime'], SHAPE['nchan'], *COUNTS,
                          beam=SHAPE['beam'], seed=SEED, flag_fraction=0.25, observed='model')
    r = np.random.default_rng(SEED)
    a['sersic_shape'][:2] = r.uniform(-0.3, 0.3, (2, COUNTS[2]))
    a['weight_vector'] = r.uniform(0.5, 2.0, a['weight_vector'].shape)
    return a


def main():
    a = inputs()
    out = {k: v for k, v in a.items() if isinstance(v, np.ndarray)}
    for k in rime_np.BEAM_KEYS:
        out[k] = np.float64(a[k])
    out['counts'] = np.array(COUNTS)
    m = rime_np.rime(a, *COUNTS, use_weight_vector=True)
    r = rime_np.rime(a, *COUNTS, residual=True, sigma_sqrd=1.7)
    out['expect_model_vis'] = m['model_vis']
    out['expect_chi_sqrd_result'] = m['chi_sqrd_result']
    out['expect_X2_weighted'] = np.float64(m['X2'])
    out['expect_residual_vis'] = r['model_vis']
    out['expect_X2_sigma'] = np.float64(r['X2'])
    g = rime_torch.gradients(a, *COUNTS, use_weight_vector=True)
    for k in ('lm', 'stokes', 'alpha', 'gauss_shape', 'sersic_shape', 'G_term'):
        out['expect_grad_' + k] = g[k]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'rime_small.npz')
    np.savez_compressed(path, **out)
    print('wrote', path, os.path.getsize(path), 'bytes')


if __name__ == '__main__':
    main()
