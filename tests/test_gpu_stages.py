"""GPU parity of the per-node stage kernels and of the drop-in solver.

Mirrors the structure of montblanc/tests/test_rime_v4.py: each kernel node
is checked in isolation against the oracle (B_sqrt, E beam, EKB sqrt, and
the coherency sum driven with random jones for every source-type
permutation and both weight settings), then the solver runs the staged and
the fused pipelines through rime_solver(...).
"""

import itertools

import numpy as np
import pytest
import torch

from oracle import rime_np, synth

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-10, np.float32: 1e-5}
DTYPES = [np.float64, np.float32]


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def small(dtype, counts=(3, 2, 1), seed=5, **kw):
    a = synth.make_inputs(7, 4, 6, *counts, beam=(9, 8, 5), seed=seed, **kw)
    a['sersic_shape'][:2] = np.random.default_rng(seed).uniform(-0.3, 0.3, (2, counts[2]))
    return synth.as_dtype(a, dtype)


@pytest.mark.parametrize('dtype', DTYPES)
def test_b_sqrt(dtype):
    from montblanc_amd import rime
    a = small(dtype)
    out = rime.b_sqrt(rime.to_device(a), (3, 2, 1)).cpu().numpy()
    up = synth.upcast(a)
    ref = rime_np.b_sqrt_jones(rime_np.b_jones(up['stokes'], up['alpha'], up['frequency'],
                                               up['ref_frequency']))
    assert rel(out, ref) < TOL[dtype]
    # invariants of test_rime_v4.py:311-315 on the device result
    np.testing.assert_allclose(rime_np.jones_mul(out, out),
                               rime_np.b_jones(up['stokes'], up['alpha'], up['frequency'],
                                               up['ref_frequency']),
                               rtol=20 * TOL[dtype], atol=20 * TOL[dtype])


@pytest.mark.parametrize('dtype', DTYPES)
@pytest.mark.parametrize('beam', [(9, 8, 5), (2, 2, 2)])
def test_e_beam(dtype, beam):
    from montblanc_amd import rime
    a = synth.as_dtype(synth.make_inputs(7, 3, 6, 4, 0, 0, beam=beam, seed=7), dtype)
    out = rime.e_beam(rime.to_device(a), (4, 0, 0)).cpu().numpy()
    up = synth.upcast(a)
    ref = rime_np.e_beam(up['lm'], up['parallactic_angles'], up['point_errors'],
                         up['antenna_scaling'], up['frequency'], up['E_beam'],
                         *(float(up[k]) for k in rime_np.BEAM_KEYS))
    assert rel(out, ref) < TOL[dtype]
    # non-trivial: most entries non-zero (test_rime_v4.py:367-372)
    assert np.count_nonzero(out) > 0.85 * out.size


@pytest.mark.parametrize('dtype', DTYPES)
def test_ekb_sqrt(dtype):
    from montblanc_amd import rime
    a = small(dtype)
    arrays = rime.to_device(a)
    jones = rime.e_beam(arrays, (3, 2, 1))
    B = rime.b_sqrt(arrays, (3, 2, 1))
    rime.ekb_sqrt(arrays, (3, 2, 1), B, jones)
    ref = rime_np.per_antenna_jones(synth.upcast(a))
    assert rel(jones.cpu().numpy(), ref) < TOL[dtype]


PERMS = [c for c in itertools.product((0, 3), (0, 2), (0, 2)) if sum(c) > 0]


@pytest.mark.parametrize('dtype', DTYPES)
@pytest.mark.parametrize('counts', PERMS)
@pytest.mark.parametrize('use_w', [False, True])
def test_sum_coherencies_random_jones(dtype, counts, use_w):
    """test_rime_v4.py:224-264: random jones into the coherency kernel."""
    from montblanc_amd import rime
    a = small(dtype, counts=counts, seed=sum(counts) + 3 * use_w, flag_fraction=0.5)
    nsrc = sum(counts)
    r = np.random.default_rng(11)
    ntime, na = a['uvw'].shape[:2]
    nchan = a['frequency'].shape[0]
    ct = np.complex64 if dtype == np.float32 else np.complex128
    J = (r.standard_normal((nsrc, ntime, na, nchan, 4)) +
         1j * r.standard_normal((nsrc, ntime, na, nchan, 4))).astype(ct)
    a['weight_vector'] = r.uniform(0.5, 2, a['weight_vector'].shape).astype(dtype)
    arrays = rime.to_device(a)
    jones = torch.from_numpy(J).cuda()
    mv = torch.empty(a['observed_vis'].shape, dtype=jones.dtype, device='cuda')
    chi = torch.empty(a['observed_vis'].shape[:3], dtype=arrays['uvw'].dtype, device='cuda')
    X2 = torch.zeros((), dtype=torch.float64, device='cuda')
    rime.sum_coherencies(arrays, counts, jones, mv, chi, use_weight_vector=use_w)
    rime.chi2_reduce(arrays, counts, chi, X2, use_weight_vector=use_w, sigma_sqrd=2.0)
    up = synth.upcast(a)
    V = rime_np.coherency_sum(J.astype(np.complex128), up['antenna1'], up['antenna2'], *counts,
                              uvw=up['uvw'], gauss_shape=up['gauss_shape'],
                              sersic_shape=up['sersic_shape'], frequency=up['frequency'])
    M = rime_np.gekb_vis(V, up['G_term'], up['antenna1'], up['antenna2'], up['flag'],
                         up['observed_vis'], False)
    terms = rime_np.chi_sqrd_sum_terms(M, up['observed_vis'], up['flag'], up['weight_vector'],
                                       False, use_w)
    x2 = rime_np.chi_sqrd(terms, 2.0, use_w)
    assert rel(mv.cpu().numpy(), M) < TOL[dtype]
    assert rel(chi.cpu().numpy(), terms) < TOL[dtype]
    assert abs(X2.item() - x2) < TOL[dtype] * abs(x2)


def make_solver(dtype, pipeline, counts=(3, 2, 1), residual=False, weights=False, na=7):
    import montblanc_amd as mb
    cfg = mb.rime_solver_cfg(na=na, ntime=4, nchan=6, dtype=dtype,
                             sources=mb.sources(point=counts[0], gaussian=counts[1],
                                                sersic=counts[2]),
                             beam_lw=9, beam_mh=8, beam_nud=5, data_source='test',
                             weight_vector=weights,
                             vis_output='residuals' if residual else 'model',
                             pipeline=pipeline)
    return mb.rime_solver(cfg)


def solver_inputs(slvr):
    a = {}
    for name in ('uvw', 'antenna1', 'antenna2', 'frequency', 'ref_frequency', 'parallactic_angles',
                 'point_errors', 'antenna_scaling', 'E_beam', 'G_term', 'lm', 'stokes', 'alpha',
                 'gauss_shape', 'sersic_shape', 'flag', 'weight_vector', 'observed_vis'):
        a[name] = getattr(slvr, 'retrieve_' + name)()
    for k in rime_np.BEAM_KEYS + ('gauss_scale', 'two_pi_over_c', 'sigma_sqrd'):
        a[k] = float(getattr(slvr, k))
    return synth.upcast(a)


@pytest.mark.parametrize('dtype', ['double', 'float'])
@pytest.mark.parametrize('residual,weights', [(False, False), (True, True)])
def test_solver_staged_and_fused_pipelines(dtype, residual, weights):
    import montblanc_amd as mb
    np.random.seed(3)
    ft = np.float64 if dtype == 'double' else np.float32
    results = {}
    for kind in ('staged', 'fused'):
        pipe = mb.staged_pipeline() if kind == 'staged' else mb.default_pipeline()
        np.random.seed(3)
        slvr = make_solver(dtype, pipe, residual=residual, weights=weights)
        slvr.set_beam_ll(-0.06); slvr.set_beam_lm(-0.06); slvr.set_beam_ul(0.06)
        slvr.set_beam_um(0.06); slvr.set_sigma_sqrd(1.5)
        with slvr:
            slvr.solve()
            assert slvr.pipeline.last_error is None, slvr.pipeline.last_error
            results[kind] = (slvr.retrieve_model_vis(), slvr.retrieve_chi_sqrd_result(), slvr.X2)
            a = solver_inputs(slvr)
    ref = rime_np.rime(a, 3, 2, 1, residual=residual, use_weight_vector=weights, sigma_sqrd=1.5)
    for kind, (mv, chi, x2) in results.items():
        assert rel(mv, ref['model_vis']) < TOL[ft], kind
        assert rel(chi, ref['chi_sqrd_result']) < TOL[ft], kind
        assert abs(float(x2) - ref['X2']) < TOL[ft] * abs(ref['X2']), kind


def test_solver_transfer_retrieve_and_validation():
    import montblanc_amd as mb
    slvr = make_solver('double', mb.default_pipeline())
    x = np.random.default_rng(0).uniform(size=slvr.retrieve_uvw().shape)
    slvr.transfer_uvw(x)
    np.testing.assert_array_equal(slvr.retrieve_uvw(), x)
    with pytest.raises(ValueError):
        slvr.transfer_uvw(x[:1])
    with pytest.raises(ValueError):
        slvr.transfer_uvw(x.astype(np.float32))
    assert slvr.dim_local_size('nbl') == 21
    assert slvr.dim_local_size('ntime', 'na') == (4, 7)


def test_pipeline_error_is_logged_and_swallowed():
    import montblanc_amd as mb
    slvr = make_solver('double', mb.default_pipeline())
    slvr.set_beam_ul(float(slvr.beam_ll))      # degenerate beam -> library rejects
    with slvr:
        assert slvr.solve() is False
        assert isinstance(slvr.pipeline.last_error, mb._lib.MontblancHipError)
