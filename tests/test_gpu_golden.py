"""HIP path against the committed fixtures (tests/golden/rime_small.npz):
float64 forward, residual mode and adjoint within 1e-10 relative."""

import numpy as np
import pytest

from golden_util import GRAD_NAMES, load

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(np.asarray(a) - b).max() / np.abs(b).max())


@pytest.mark.parametrize('pair_block', ['generic', (1, 1, -1), (2, 2, 0)])
def test_forward_matches_fixtures(pair_block):
    from montblanc_amd import rime
    a, counts, e = load()
    print('golden forward', pair_block, flush=True)
    out = rime.forward_numpy(a, counts, use_weight_vector=True, pair_block=pair_block)
    assert rel(out['model_vis'], e['model_vis']) < 1e-10
    assert rel(out['chi_sqrd_result'], e['chi_sqrd_result']) < 1e-10
    assert abs(out['X2'] - e['X2_weighted']) < 1e-10 * e['X2_weighted']
    r = rime.forward_numpy(a, counts, residual=True, sigma_sqrd=1.7, pair_block=pair_block)
    assert rel(r['model_vis'], e['residual_vis']) < 1e-10
    assert abs(r['X2'] - e['X2_sigma']) < 1e-10 * e['X2_sigma']


def test_adjoint_matches_fixtures():
    import torch
    from montblanc_amd import rime
    a, counts, e = load()
    print('golden adjoint', flush=True)
    arrays = rime.to_device(a)
    out = rime.forward(arrays, counts, use_weight_vector=True, keep_ekb=True)
    g = rime.adjoint(arrays, counts, out['ekb_vis'], use_weight_vector=True)
    torch.cuda.synchronize()
    for k in GRAD_NAMES:
        assert rel(g[k].cpu().numpy(), e['grad_' + k]) < 1e-10, k
