"""Pin the gradient oracle (oracle/rime_torch.py).

Its forward must equal the numpy oracle's, and its autograd gradient must
equal central finite differences of the numpy oracle's chi^2 -- two
independent implementations of the chain, so neither a dropped term nor a
wrong sign in the torch twin survives.
"""

import numpy as np
import pytest

from oracle import rime_np, rime_torch, synth

COUNTS = (2, 2, 1)


def problem(seed=3, weights=False, flags=0.3):
    a = synth.make_inputs(5, 2, 3, *COUNTS, beam=(7, 6, 4), seed=seed, flag_fraction=flags,
                          observed='model')
    r = np.random.default_rng(seed)
    a['sersic_shape'][:2] = r.uniform(-0.3, 0.3, (2, COUNTS[2]))
    a['observed_vis'] = a['observed_vis'] * (1 + 0.05 * r.standard_normal(a['observed_vis'].shape))
    if weights:
        a['weight_vector'] = r.uniform(0.5, 2.0, a['weight_vector'].shape)
    return a


@pytest.mark.parametrize('residual,weights', [(False, False), (True, True)])
def test_torch_forward_matches_numpy(residual, weights):
    a = problem(weights=weights)
    t = rime_torch.to_torch(a, requires_grad=False)
    x2, vis = rime_torch.chi2(t, *COUNTS, residual=residual, use_weight_vector=weights,
                              sigma_sqrd=1.3)
    ref = rime_np.rime(a, *COUNTS, residual=residual, use_weight_vector=weights, sigma_sqrd=1.3)
    np.testing.assert_allclose(vis.numpy(), ref['model_vis'], rtol=1e-12, atol=1e-12)
    assert abs(float(x2) - ref['X2']) < 1e-12 * ref['X2']


def fd_grad(a, key, idx, h, imag=False, **kw):
    """Central difference with one Richardson step (cancels the h^2 term)."""
    def f(delta):
        b = dict(a)
        b[key] = a[key].copy()
        b[key][idx] += 1j * delta if imag else delta
        return rime_np.rime(b, *COUNTS, **kw)['X2']
    d1 = (f(h) - f(-h)) / (2 * h)
    d2 = (f(h / 2) - f(-h / 2)) / h
    return (4 * d2 - d1) / 3


# finite-difference steps per parameter (scale of the parameter's effect);
# shape rows: gaussian (el, em, R), sersic (e1, e2, scale ~ 1e-5..1e-4)
STEP = {'lm': 1e-7, 'stokes': 1e-4, 'alpha': 1e-4, 'G_term': 1e-5,
        ('gauss_shape', 0): 1e-8, ('gauss_shape', 1): 1e-8, ('gauss_shape', 2): 1e-6,
        ('sersic_shape', 0): 1e-4, ('sersic_shape', 1): 1e-4, ('sersic_shape', 2): 1e-8}


def step(key, idx):
    return STEP.get((key, idx[0]), STEP.get(key))


@pytest.mark.parametrize('weights', [False, True])
def test_torch_gradient_matches_finite_differences(weights):
    kw = dict(use_weight_vector=weights, sigma_sqrd=1.3)
    a = problem(seed=11, weights=weights)
    g = rime_torch.gradients(a, *COUNTS, **kw)
    r = np.random.default_rng(0)
    checks = [('lm', (0, 0)), ('lm', (3, 1)), ('lm', (4, 0)),
              ('stokes', (0, 1, 0)), ('stokes', (2, 0, 1)), ('stokes', (3, 1, 2)), ('stokes', (4, 0, 3)),
              ('alpha', (1, 1)), ('alpha', (4, 0)),
              ('gauss_shape', (0, 0)), ('gauss_shape', (1, 1)), ('gauss_shape', (2, 0)),
              ('sersic_shape', (0, 0)), ('sersic_shape', (1, 0)), ('sersic_shape', (2, 0))]
    for key, idx in checks:
        fd = fd_grad(a, key, idx, step(key, idx), **kw)
        got = g[key][idx]
        assert abs(got - fd) <= 1e-6 * max(abs(fd), 1e-3 * abs(g['X2'])), (key, idx, got, fd)
    # complex gains: torch's convention d/dRe + i d/dIm
    for idx in [(0, 1, 0, 0), (1, 3, 2, 1), (1, 0, 1, 3)]:
        fr = fd_grad(a, 'G_term', idx, STEP['G_term'], **kw)
        fi = fd_grad(a, 'G_term', idx, STEP['G_term'], imag=True, **kw)
        got = g['G_term'][idx]
        scale = max(abs(fr), abs(fi), 1e-4 * g['X2'])
        assert abs(got.real - fr) <= 1e-6 * scale and abs(got.imag - fi) <= 1e-6 * scale, \
            (idx, got, fr, fi)
