"""GPU adjoint (hand-written backward) against the reference gradient.

The reference gradient is torch float64 autograd of oracle/rime_torch.py,
itself pinned to central finite differences of the numpy oracle
(tests/test_oracle_grad.py). float64 gradients must agree to 1e-10
relative per parameter array (max-norm); float32 ones (float64 phase and
beam coordinates) within the derived first-order bound of
tests/grad_bound.py: the float32 model's measured error and the Jones
recomputation's error carried through chi^2's derivative, by
Cauchy-Schwarz over the visibilities, element by element.
"""

import numpy as np
import pytest
import torch

from grad_bound import check_f32_gradients
from oracle import rime_np, rime_torch, synth

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-10}
NAMES = ('lm', 'stokes', 'alpha', 'gauss_shape', 'sersic_shape', 'G_term')


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def problem(na, ntime, nchan, counts, seed, weights=False, flags=0.3, auto=False):
    a = synth.make_inputs(na, ntime, nchan, *counts, beam=(9, 8, 6), seed=seed,
                          flag_fraction=flags, observed='model', auto_correlations=auto)
    r = np.random.default_rng(seed)
    if counts[2]:
        a['sersic_shape'][:2] = r.uniform(-0.3, 0.3, (2, counts[2]))
    a['observed_vis'] = a['observed_vis'] * (1 + 0.05 * r.standard_normal(a['observed_vis'].shape))
    if weights:
        a['weight_vector'] = r.uniform(0.5, 2.0, a['weight_vector'].shape)
    a['sigma_sqrd'] = 1.3
    return a


def gpu_grads(a, counts, dtype, use_w):
    from montblanc_amd import rime
    print('adjoint', np.dtype(dtype).name, a['uvw'].shape, counts, use_w, flush=True)
    arrays = rime.to_device(synth.as_dtype(a, dtype))
    out = rime.forward(arrays, counts, use_weight_vector=use_w, keep_ekb=True)
    g = rime.adjoint(arrays, counts, out['ekb_vis'], use_weight_vector=use_w)
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in g.items()}
    got['model_vis'] = out['model_vis'].cpu().numpy()
    return got, float(out['X2'].item())


def check_grads(a, counts, dtype, got, ref, use_w=False, sigma_sqrd=1.3, names=NAMES):
    """float64: 1e-10 relative per array; float32: the derived bound."""
    if dtype == np.float32:
        up = synth.upcast(synth.as_dtype(a, dtype))
        print(check_f32_gradients(up, counts, got, ref, got['model_vis'], use_weight_vector=use_w,
                                  sigma_sqrd=sigma_sqrd, names=names))
        return
    for name in names:
        if ref[name] is None or ref[name].size == 0:
            continue
        e = rel(got[name], ref[name])
        print(name, e)
        assert e < TOL[dtype], (name, e)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('counts,weights', [((3, 2, 1), False), ((2, 3, 0), True),
                                            ((0, 0, 2), False), ((4, 0, 0), True)])
def test_adjoint_matches_reference_gradient(dtype, counts, weights):
    a = problem(7, 3, 5, counts, seed=sum(counts) + 10 * weights, weights=weights)
    ref = rime_torch.gradients(synth.upcast(synth.as_dtype(a, dtype)), *counts,
                               use_weight_vector=weights, sigma_sqrd=1.3)
    got, x2 = gpu_grads(a, counts, dtype, weights)
    if dtype == np.float64:
        assert abs(x2 - ref['X2']) < TOL[dtype] * ref['X2']
    else:
        # |X2~ - X2| <= (2 w_max / s)(||R|| ||dM|| + ||dM||^2 / 2) + 8 u X2
        up = synth.upcast(synth.as_dtype(a, dtype))
        keep = up['flag'] == 0
        M = rime_np.rime(up, *counts)['model_vis'] * keep
        dM = np.linalg.norm((got['model_vis'] - M) * keep)
        R = np.linalg.norm((up['observed_vis'] - M) * keep)
        wmax, s = (up['weight_vector'].max(), 1.0) if weights else (1.0, 1.3)
        assert abs(x2 - ref['X2']) <= 2 * wmax / s * (R * dM + dM * dM / 2) + 8 * 2.0 ** -24 * ref['X2']
    check_grads(a, counts, dtype, got, ref, use_w=weights)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_adjoint_autocorrelations_and_single_channel(dtype):
    counts = (2, 2, 0)
    a = problem(5, 2, 1, counts, seed=41, auto=True)
    ref = rime_torch.gradients(synth.upcast(synth.as_dtype(a, dtype)), *counts, sigma_sqrd=1.3)
    got, _ = gpu_grads(a, counts, dtype, False)
    check_grads(a, counts, dtype, got, ref, names=('lm', 'stokes', 'alpha', 'gauss_shape', 'G_term'))


@pytest.mark.parametrize('dtype,na,counts', [
    (np.float64, 9, (40, 3, 0)),      # three point tiles of 16, pipelined
    (np.float64, 27, (300, 0, 0)),    # 19 tiles: several gradient windows
    (np.float64, 33, (20, 0, 0)),     # 2 na > 64: the single-phase kernel
    (np.float32, 27, (35, 2, 0)),
    (np.float32, 64, (70, 0, 1)),     # tiles of 8 sources, 8 matrix waves
])
def test_adjoint_many_point_tiles(dtype, na, counts):
    """Point sources spread over many source tiles: the pipelined adjoint
    (X of tile p, product of tile p - 1 and backward of tile p - 2 in one
    phase) and its gradient windows."""
    a = problem(na, 2, 3, counts, seed=na + counts[0])
    ref = rime_torch.gradients(synth.upcast(synth.as_dtype(a, dtype)), *counts, sigma_sqrd=1.3)
    got, _ = gpu_grads(a, counts, dtype, False)
    check_grads(a, counts, dtype, got, ref)


@pytest.mark.parametrize('dtype,na', [(np.float64, 27), (np.float32, 27), (np.float64, 40),
                                      (np.float32, 40), (np.float32, 64)])
def test_adjoint_many_shaped_tiles(dtype, na):
    """More gaussian and sersic sources than one gradient window holds (up
    to 32 source tiles): the shaped pass's window flushes, and in float its
    sixteen-wave layout with the pair loop split over two threads. Odd and
    even antenna counts take the two record-row paddings of the shaped
    pass's X records."""
    counts = (4, 600, 30) if na < 64 else (2, 150, 30)
    a = problem(na, 1, 2, counts, seed=77 + na)
    ref = rime_torch.gradients(synth.upcast(synth.as_dtype(a, dtype)), *counts, sigma_sqrd=1.3)
    got, _ = gpu_grads(a, counts, dtype, False)
    check_grads(a, counts, dtype, got, ref)


def test_autograd_function_and_gpu_finite_differences():
    """The torch.autograd.Function path gives the same gradients as the
    adjoint, and they match central differences (h = 1e-7 relative) of the
    float64 GPU forward itself."""
    from montblanc_amd import autograd, rime
    counts = (2, 1, 1)
    a = problem(6, 2, 3, counts, seed=5)
    arrays = rime.to_device(a)
    params = {}
    for n in autograd.PARAMS:
        params[n] = arrays[n].clone().requires_grad_(True)
    arr = dict(arrays, **params)
    x2 = autograd.chi2(arr, counts)
    x2.backward()
    ref = rime_torch.gradients(a, *counts, sigma_sqrd=1.3)
    for n in ('lm', 'stokes', 'alpha', 'gauss_shape', 'sersic_shape', 'G_term'):
        assert rel(params[n].grad.cpu().numpy(), ref[n]) < 1e-10, n

    def x2_of(name, idx, delta, imag=False):
        b = dict(arrays)
        t = arrays[name].clone()
        t[idx] += (1j * delta) if imag else delta
        b[name] = t
        return rime.forward(b, counts)['X2'].item()

    for name, idx in [('lm', (0, 1)), ('stokes', (1, 0, 2)), ('alpha', (3, 1)),
                      ('gauss_shape', (0, 0)), ('sersic_shape', (1, 0))]:
        x = float(arrays[name][idx])
        h = 1e-7 * max(abs(x), 1e-3)
        fd = (x2_of(name, idx, h) - x2_of(name, idx, -h)) / (2 * h)
        got = float(params[name].grad[idx])
        assert abs(got - fd) <= 1e-5 * max(abs(fd), 1e-3 * x2.item()), (name, idx, got, fd)
    idx = (1, 2, 0, 1)
    h = 1e-7
    fr = (x2_of('G_term', idx, h) - x2_of('G_term', idx, -h)) / (2 * h)
    fi = (x2_of('G_term', idx, h, True) - x2_of('G_term', idx, -h, True)) / (2 * h)
    g = complex(params['G_term'].grad[idx])
    assert abs(g - complex(fr, fi)) <= 1e-5 * max(abs(complex(fr, fi)), 1e-3 * x2.item())


def test_adjoint_rejects_partial_source_batch_and_duplicates():
    from montblanc_amd import _lib, rime
    counts = (2, 0, 0)
    a = problem(4, 1, 2, counts, seed=3)
    arrays = rime.to_device(a)
    out = rime.forward(arrays, counts, keep_ekb=True)
    dup = dict(arrays)
    a1 = arrays['antenna1'].clone()
    a2 = arrays['antenna2'].clone()
    a1[:, 1], a2[:, 1] = a1[:, 0], a2[:, 0]
    dup['antenna1'], dup['antenna2'] = a1, a2
    # raised on the host before any launch (the layout builder's
    # MBX_E_UNSUPPORTED, as a ValueError naming the layout rule)
    with pytest.raises(ValueError, match='one baseline per antenna pair'):
        rime.adjoint(dup, counts, out['ekb_vis'])
