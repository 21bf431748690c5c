"""The committed fixtures (tests/golden/rime_small.npz, written by
tests/golden/make_golden.py) against the oracle: the oracle's pinned state
is frozen, and the GPU tests compare the HIP path with the same file."""

import numpy as np

from golden_util import GRAD_NAMES, load
from oracle import rime_np, rime_torch


def test_oracle_reproduces_fixtures():
    a, counts, e = load()
    m = rime_np.rime(a, *counts, use_weight_vector=True)
    np.testing.assert_allclose(m['model_vis'], e['model_vis'], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(m['chi_sqrd_result'], e['chi_sqrd_result'], rtol=1e-13, atol=1e-13)
    assert abs(m['X2'] - e['X2_weighted']) <= 1e-13 * e['X2_weighted']
    r = rime_np.rime(a, *counts, residual=True, sigma_sqrd=1.7)
    np.testing.assert_allclose(r['model_vis'], e['residual_vis'], rtol=1e-13, atol=1e-13)
    assert abs(r['X2'] - e['X2_sigma']) <= 1e-13 * e['X2_sigma']


def test_gradient_oracle_reproduces_fixtures():
    a, counts, e = load()
    g = rime_torch.gradients(a, *counts, use_weight_vector=True)
    for k in GRAD_NAMES:
        np.testing.assert_allclose(g[k], e['grad_' + k], rtol=1e-12, atol=1e-12 * np.abs(e['grad_' + k]).max())
