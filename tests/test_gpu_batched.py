"""Batched evaluation (the matrix-core passes over groups of samples, one
persistent launch over all samples, or the per-sample graph) and HIP-graph
replay.

Each sample's chi^2 and gradients must equal the numpy / torch float64
oracle evaluated on that sample alone (float64: 1e-10 relative), and the
launch-shape variants the autotuner picks from must all agree.
"""

import numpy as np
import pytest
import torch

from grad_bound import check_f32_gradients
from oracle import rime_np, rime_torch, synth

pytestmark = pytest.mark.gpu

COUNTS = [(3, 1, 0), (0, 2, 0), (7, 2, 1), (1, 0, 0), (2, 3, 2)]


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def make_batch(na=7, ntime=3, nchan=4, counts=COUNTS, seed=8, dtype=np.float64, weights=False):
    shared = synth.make_inputs(na, ntime, nchan, 0, 0, 0, beam=(9, 8, 6), seed=seed,
                               flag_fraction=0.3)
    r = np.random.default_rng(seed)
    if weights:
        shared['weight_vector'] = r.uniform(0.5, 2.0, shared['weight_vector'].shape)
    shared['sigma_sqrd'] = 1.3
    samples = []
    for i, c in enumerate(counts):
        s = synth.make_inputs(na, ntime, nchan, *c, beam=(2, 2, 2), seed=seed + 100 + i)
        s['sersic_shape'][:2] = r.uniform(-0.3, 0.3, (2, c[2]))
        samples.append({k: s[k] for k in ('lm', 'stokes', 'alpha', 'gauss_shape', 'sersic_shape')})
        samples[-1]['counts'] = c
    # observed: the first sample's model plus noise, so chi^2 values differ
    full = dict(shared)
    full.update({k: samples[0][k] for k in ('lm', 'stokes', 'alpha', 'gauss_shape', 'sersic_shape')})
    full['flag'] = np.zeros_like(shared['flag'])
    full['observed_vis'] = np.zeros_like(shared['observed_vis'])
    m = rime_np.rime(full, *counts[0])['model_vis']
    shared['observed_vis'] = m + 0.1 * (r.standard_normal(m.shape) + 1j * r.standard_normal(m.shape))
    return shared, samples


def sample_inputs(shared, s):
    a = dict(shared)
    a.update({k: s[k] for k in ('lm', 'stokes', 'alpha', 'gauss_shape', 'sersic_shape')})
    return a


MATRIX = {'strategy': 'matrix'}
MATRIX2 = {'strategy': 'matrix', 'samples_per_group': 2}


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('weights', [False, True])
@pytest.mark.parametrize('strategy', [None, {'strategy': 'per_sample'}, MATRIX, MATRIX2])
def test_batched_matches_oracle_per_sample(dtype, weights, strategy):
    """Every strategy -- the persistent launch (library default shape), the
    per-sample matrix-core sequence replayed as a HIP graph, and the
    matrix-core passes launched over all samples at once or over groups of
    two (the last group partial) -- against the float64 reference gradients
    of every sample. The samples include one without point sources and one
    without gaussian or sersic sources. The matrix strategy is float only:
    float64 is rejected before any launch."""
    from montblanc_amd import _lib, batched
    shared, samples = make_batch(dtype=dtype, weights=weights)
    print('batched', np.dtype(dtype).name, weights, strategy, flush=True)
    b = batched.Batch(shared, samples, dtype=dtype)
    if dtype == np.float64 and strategy in (MATRIX, MATRIX2):
        with pytest.raises(_lib.MontblancHipError, match='float only'):
            batched.evaluate(b, gradients=True, use_weight_vector=weights, tune=strategy)
        return
    X2, g = batched.evaluate(b, gradients=True, use_weight_vector=weights, tune=strategy)
    if strategy is not None:   # a second replay of the captured graph gives the same
        X2 = X2.clone()
        g = {k: v.clone() for k, v in g.items()}
        X2b, gb = batched.evaluate(b, gradients=True, use_weight_vector=weights, tune=strategy)
        torch.cuda.synchronize()
        assert torch.equal(X2b, X2)                 # fixed-order reduction
        for k in g:                                 # float64 atomics: order varies
            assert rel(gb[k].cpu().numpy(), g[k].cpu().numpy()) < 1e-12, k
    torch.cuda.synchronize()
    X2 = X2.cpu().numpy()
    g = {k: v.cpu().numpy() for k, v in g.items()}
    for k, s in enumerate(samples):
        a = synth.upcast(synth.as_dtype(sample_inputs(shared, s), dtype))
        ref = rime_torch.gradients(a, *s['counts'], use_weight_vector=weights, sigma_sqrd=1.3)
        o0, o1 = b.src_offset[k], b.src_offset[k + 1]
        got = {name: g[name][o0:o1] for name in ('lm', 'stokes', 'alpha')}
        if s['counts'][1]:
            got['gauss_shape'] = b.split_shapes(torch.from_numpy(g['gauss_shape']), 'gauss')[k].numpy()
        if s['counts'][2]:
            got['sersic_shape'] = b.split_shapes(torch.from_numpy(g['sersic_shape']), 'sersic')[k].numpy()
        if dtype == np.float64:
            assert abs(X2[k] - ref['X2']) <= 1e-10 * ref['X2']
            for name, v in got.items():
                assert rel(v, ref[name]) < 1e-10, (k, name)
        else:
            # X2: the float32 model's error (at most 1e-5 ||M||, the forward's
            # bound) through the residual, |dX2| <= 2 w_max ||R|| dM / s + dM^2 / s
            keep = a['flag'] == 0
            M = rime_np.rime(a, *s['counts'])['model_vis'] * keep
            dM = 1e-5 * np.linalg.norm(M)
            R = np.linalg.norm((a['observed_vis'] - M) * keep)
            wmax, sg = (a['weight_vector'].max(), 1.0) if weights else (1.0, 1.3)
            assert abs(X2[k] - ref['X2']) <= 2 * wmax / sg * (R * dM + dM * dM) + 1e-6 * ref['X2']
            print(k, check_f32_gradients(a, s['counts'], got, ref, None, use_weight_vector=weights,
                                         sigma_sqrd=1.3, names=tuple(got)))


@pytest.mark.parametrize('strategy,dtype', [(None, np.float64), ({'strategy': 'per_sample'}, np.float64),
                                            (MATRIX, np.float32)])
def test_batched_sample_without_sources(strategy, dtype):
    """A sample with no sources at all: its chi^2 is that of a zero model
    (the oracle's), and the other samples' chi^2 and gradients are those of
    the same batch without it."""
    from montblanc_amd import batched
    counts = [(3, 1, 0), (0, 0, 0), (2, 2, 1)]
    shared, samples = make_batch(na=9, ntime=2, nchan=3, counts=counts, seed=61)
    print('empty sample', strategy, np.dtype(dtype).name, flush=True)
    b = batched.Batch(shared, samples, dtype=dtype)
    X2, g = batched.evaluate(b, gradients=True, tune=strategy)
    torch.cuda.synchronize()
    X2 = X2.cpu().numpy()
    g = {k: v.cpu().numpy() for k, v in g.items()}
    tol = 1e-10 if dtype == np.float64 else 1e-5
    ref = rime_np.rime(synth.upcast(synth.as_dtype(sample_inputs(shared, samples[1]), dtype)), 0, 0, 0,
                       sigma_sqrd=shared['sigma_sqrd'])
    assert abs(X2[1] - ref['X2']) <= tol * ref['X2']
    kept = [samples[0], samples[2]]
    b2 = batched.Batch(shared, kept, dtype=dtype)
    X2b, gb = batched.evaluate(b2, gradients=True, tune=strategy)
    torch.cuda.synchronize()
    assert rel(X2[[0, 2]], X2b.cpu().numpy()) < 1e-12
    for name in ('lm', 'stokes', 'alpha', 'gauss_shape', 'sersic_shape'):
        assert rel(g[name], gb[name].cpu().numpy()) < (1e-11 if dtype == np.float64 else 1e-5), name


@pytest.mark.parametrize('tune', [None, {'workgroups': 1}, {'workgroups': 3, 'channels_per_item': 2},
                                  {'sources_per_tile': 2, 'workgroups': 17}])
def test_batched_launch_shapes_agree(tune):
    """Every launch shape (down to one persistent workgroup working through
    the whole queue) gives the same results."""
    from montblanc_amd import batched
    shared, samples = make_batch(na=9, ntime=2, nchan=5, seed=21)
    b = batched.Batch(shared, samples)
    X2, g = batched.evaluate(b, gradients=True, tune=None)
    X2 = X2.clone()
    g = {k: v.clone() for k, v in g.items()}
    X2b, gb = batched.evaluate(b, gradients=True, tune=tune)
    assert rel(X2b.cpu(), X2.cpu()) < 1e-12
    for k in g:
        assert rel(gb[k].cpu(), g[k].cpu()) < 1e-11, k


def test_batched_64_antennas_matches_single_model_path():
    """At 64 antennas (2016 baselines) the batched float32 results and the
    single-model float32 forward + adjoint run per sample both agree with
    the float64 reference gradient within the derived float32 bound
    (tests/grad_bound.py): the gradient carries the residual R = O - M,
    whose float32 error is that of the model, so near a fit (|M| / |R| up
    to ~120 here) the model's error dominates; the bound carries it through
    by Cauchy-Schwarz, element by element."""
    from montblanc_amd import batched, rime
    counts = [(40, 10, 0), (5, 2, 0), (200, 60, 0)]
    shared, samples = make_batch(na=64, ntime=2, nchan=3, counts=counts, seed=33)
    b = batched.Batch(shared, samples, dtype=np.float32)
    X2, g = batched.evaluate(b, gradients=True, tune=None)
    X2, g = X2.clone(), {n: v.clone() for n, v in g.items()}
    # the matrix-core passes over all three samples at once
    X2m, gm = batched.evaluate(b, gradients=True, tune=MATRIX)
    torch.cuda.synchronize()
    for k, s in enumerate(samples):
        a32 = synth.as_dtype(sample_inputs(shared, s), np.float32)
        ref = rime_torch.gradients(synth.upcast(a32), *s['counts'], sigma_sqrd=shared['sigma_sqrd'])
        arrays = rime.to_device(a32)
        out = rime.forward(arrays, s['counts'], keep_ekb=True)
        gg = rime.adjoint(arrays, s['counts'], out['ekb_vis'])
        torch.cuda.synchronize()
        assert abs(X2[k].item() - out['X2'].item()) < 1e-6 * out['X2'].item()
        # the matrix strategy runs the single-model kernels' arithmetic
        assert abs(X2m[k].item() - out['X2'].item()) < 1e-9 * out['X2'].item()
        o0, o1 = b.src_offset[k], b.src_offset[k + 1]
        up = synth.upcast(a32)
        names = ('lm', 'stokes', 'alpha', 'gauss_shape')
        batch = {n: g[n][o0:o1].cpu().numpy() for n in names[:3]}
        batch['gauss_shape'] = b.split_shapes(g['gauss_shape'], 'gauss')[k].cpu().numpy()
        matrix = {n: gm[n][o0:o1].cpu().numpy() for n in names[:3]}
        matrix['gauss_shape'] = b.split_shapes(gm['gauss_shape'], 'gauss')[k].cpu().numpy()
        single = {n: gg[n].cpu().numpy() for n in names}
        for n in names:
            assert rel(matrix[n], single[n]) < 1e-5, (k, n)
        print('sample', k, 'batched', check_f32_gradients(up, s['counts'], batch, ref, None,
                                                          sigma_sqrd=shared['sigma_sqrd'], names=names[:3]),
              flush=True)
        print('sample', k, 'matrix', check_f32_gradients(up, s['counts'], matrix, ref, None,
                                                         sigma_sqrd=shared['sigma_sqrd'], names=names),
              flush=True)
        print('sample', k, 'single', check_f32_gradients(up, s['counts'], single, ref,
                                                         out['model_vis'].cpu().numpy(),
                                                         sigma_sqrd=shared['sigma_sqrd'], names=names),
              flush=True)


def test_autotuner_picks_and_caches(tmp_path, monkeypatch):
    from montblanc_amd import autotune, batched
    monkeypatch.setenv('MBX_AUTOTUNE_CACHE', str(tmp_path / 'tune.json'))
    autotune._MEM.clear()
    shared, samples = make_batch(na=9, ntime=4, nchan=3, seed=5)
    b = batched.Batch(shared, samples)
    choice = autotune.batched_tuning(b, gradients=False)
    # a persistent launch shape, or the per-sample graph (float64: no
    # matrix strategy), decided by a final timing on the full extent
    assert (set(choice) == {'workgroups', 'channels_per_item', 'sources_per_tile'}
            or choice == {'strategy': 'per_sample'}), choice
    import json
    rec = json.load(open(tmp_path / 'tune.json'))
    (entry,) = rec.values()
    assert len(entry['final_seconds']) == 2 and json.dumps(choice) in entry['final_seconds']
    autotune._MEM.clear()
    assert autotune.batched_tuning(b, gradients=False) == choice   # from disk, no timing
    X2 = batched.evaluate(b, tune='auto').clone()
    X2n = batched.evaluate(b, tune=None)
    assert rel(X2.cpu(), X2n.cpu()) < 1e-12


def test_graph_replay_matches_eager():
    """Forward and batched launches captured in a HIP graph: every replay,
    with eager work and input changes between replays, equals the eager
    result for the current inputs."""
    from montblanc_amd import batched, graph, rime
    counts = (3, 2, 1)
    a = synth.make_inputs(8, 3, 4, *counts, beam=(9, 8, 6), seed=12, flag_fraction=0.2)
    arrays = rime.to_device(a)
    out = rime.alloc_outputs(arrays)
    g = graph.Graphed(lambda: rime.forward(arrays, counts, outputs=out))
    for rep in range(3):
        arrays['stokes'].mul_(1.0 + 0.1 * rep)                 # eager change of an input
        eager = rime.forward(arrays, counts)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert rel(out['model_vis'].cpu(), eager['model_vis'].cpu()) < 1e-13
        assert abs(out['X2'].item() - eager['X2'].item()) < 1e-12 * eager['X2'].item()
    shared, samples = make_batch(seed=40)
    b = batched.Batch(shared, samples)
    gb = graph.Graphed(lambda: batched.evaluate(b, gradients=True, tune=None))
    for rep in range(3):
        b.arrays['stokes'].mul_(1.0 + 0.05 * rep)
        X2e, ge = batched.evaluate(b, gradients=True, tune=None)
        X2e, lme = X2e.clone(), ge['lm'].clone()
        torch.cuda.synchronize()
        gb.replay()
        torch.cuda.synchronize()
        assert rel(b.X2.cpu(), X2e.cpu()) < 1e-12
        assert rel(b.grads_buffers()['lm'].cpu(), lme.cpu()) < 1e-11


def test_matrix_strategy_graph_replay():
    """The matrix strategy captured in a HIP graph (its clears are fill
    kernels): every replay, with eager work and input changes between
    replays, equals the eager result for the current inputs."""
    from montblanc_amd import batched, graph
    shared, samples = make_batch(na=11, ntime=2, nchan=5, seed=44)
    b = batched.Batch(shared, samples, dtype=np.float32)
    tune = {'strategy': 'matrix', 'samples_per_group': 3}
    gb = graph.Graphed(lambda: batched.evaluate(b, gradients=True, tune=tune))
    for rep in range(3):
        b.arrays['stokes'].mul_(1.0 + 0.05 * rep)
        X2e, ge = batched.evaluate(b, gradients=True, tune=tune)
        X2e, lme = X2e.clone(), ge['lm'].clone()
        torch.cuda.synchronize()
        gb.replay()
        torch.cuda.synchronize()
        assert torch.equal(b.X2.cpu(), X2e.cpu())         # fixed-order reductions
        assert rel(b.grads_buffers()['lm'].cpu(), lme.cpu()) < 1e-5
