"""A catalog file through the forward and the batched evaluation: the
source arrays read from the reference's example catalog format give the
oracle's model and chi^2."""

import numpy as np
import pytest

from oracle import rime_np, synth

pytestmark = pytest.mark.gpu

CATALOG = ('# format npsrc: l, m, I, Q, U, V, alpha\n'
           '10e-4, 20e-4, 0.8, 0.3, 0.2, 0.4, -0.8\n'
           '30e-4, 10e-4, 0.9, 0.33, 0.21, 0.42, -0.8\n'
           '# format ngsrc: l, m, I, Q, U, V, alpha, el, em, eR\n'
           '10e-4, 20e-4, 0.8, 0.3, 0.2, 0.4, -0.8, 1e-4, 2e-4, 0.9\n'
           '30e-4, 10e-4, 0.9, 0.33, 0.21, 0.42, -0.8, 1e-4, 2e-4, 0.9\n'
           '# format nssrc: l, m, I, e1, e2, eS\n'
           '-20e-4, 5e-4, 1.1, 0.1, -0.2, 3e-5\n')


def test_catalog_feeds_forward_and_batched(tmp_path):
    from montblanc_amd import batched, rime, sky_model
    p = tmp_path / 'sky.txt'
    p.write_text(CATALOG)
    ntime = 3
    src = sky_model.source_arrays(sky_model.parse_sky_model(str(p)), ntime=ntime)
    counts = src['counts']
    assert counts == (2, 2, 1)
    a = synth.make_inputs(9, ntime, 4, *counts, beam=(9, 8, 6), seed=12, flag_fraction=0.2)
    a.update({k: src[k] for k in ('lm', 'stokes', 'alpha', 'gauss_shape', 'sersic_shape')})
    ref = rime_np.rime(a, *counts)
    print('catalog forward', counts, flush=True)
    out = rime.forward_numpy(a, counts)
    scale = np.abs(ref['model_vis']).max()
    assert np.abs(out['model_vis'] - ref['model_vis']).max() <= 1e-10 * scale
    assert abs(out['X2'] - ref['X2']) <= 1e-10 * ref['X2']
    # the same dict is one batched sample
    shared = {k: v for k, v in a.items() if k not in ('lm', 'stokes', 'alpha', 'gauss_shape', 'sersic_shape')}
    b = batched.Batch(shared, [src, src])
    X2 = batched.evaluate(b, gradients=False, tune=None)
    np.testing.assert_allclose(X2.cpu().numpy(), [ref['X2'], ref['X2']], rtol=1e-10)
