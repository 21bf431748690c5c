"""Point sources on the matrix cores (mbx_gemm.hip) against the oracle and
against the vector-unit path, over the layouts the pair -> baseline map
handles: baselines in either antenna order, auto-correlations, antenna
counts that do and do not fill the 16-row tiles, point counts that do and
do not fill a source tile, and point sources mixed with the other types."""

import numpy as np
import pytest

from oracle import rime_np, synth

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-10, np.float32: 1e-5}


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def run(a, counts, dtype, **kw):
    from montblanc_amd import rime
    print('point gemm', np.dtype(dtype).name, a['uvw'].shape, counts, kw, flush=True)
    return rime.forward_numpy(synth.as_dtype(a, dtype), counts, **kw)


def check(a, counts, dtype, pair_block=None):
    ref = rime_np.rime(synth.upcast(synth.as_dtype(a, dtype)), *counts)
    got = run(a, counts, dtype, pair_block=pair_block)
    vec = run(a, counts, dtype, pair_block=pair_block, point_gemm=False)
    assert rel(got['model_vis'], ref['model_vis']) < TOL[dtype]
    assert rel(got['model_vis'], vec['model_vis']) < TOL[dtype]
    if dtype == np.float64:
        assert abs(got['X2'] - ref['X2']) < 1e-10 * ref['X2']


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('na', [2, 8, 9, 27, 64])
@pytest.mark.parametrize('counts', [(1, 0, 0), (13, 0, 0), (41, 3, 2)])
def test_point_gemm_matches_oracle(dtype, na, counts):
    a = synth.make_inputs(na, 2, 3, *counts, beam=(9, 8, 6), seed=na + sum(counts))
    check(a, counts, dtype)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('pair_block', ['generic', (2, 2, 0), (1, 1, -1)])
def test_reversed_baselines_and_autocorrelations(dtype, pair_block):
    counts = (7, 2, 0)
    a = synth.make_inputs(12, 2, 4, *counts, beam=(9, 8, 6), seed=5, auto_correlations=True,
                          flag_fraction=0.2)
    # every third baseline with its antennas the other way round
    a1, a2 = a['antenna1'].copy(), a['antenna2'].copy()
    flip = np.zeros(a1.shape[1], bool)
    flip[::3] = True
    a['antenna1'] = np.where(flip, a2, a1).astype(a1.dtype)
    a['antenna2'] = np.where(flip, a1, a2).astype(a2.dtype)
    check(a, counts, dtype, pair_block)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_too_many_antennas_falls_back(dtype):
    counts = (3, 1, 0)
    a = synth.make_inputs(65, 1, 2, *counts, beam=(9, 8, 6), seed=8)
    check(a, counts, dtype)


@pytest.mark.parametrize('dtype,na', [(np.float32, 41), (np.float32, 64), (np.float64, 49), (np.float64, 64)])
def test_shaped_matrix_pass_many_sources(dtype, na):
    """Gaussian and sersic sources on the matrix cores (float from 40
    antennas, float64 from 48): several source tiles, the gaussian / sersic
    boundary inside a tile, a partial last antenna quad (41) or an odd
    antenna count (49: the last lane pair shares a clamped antenna), point
    sources before them and a flagged fraction."""
    counts = (3, 37, 29)
    a = synth.make_inputs(na, 2, 2, *counts, beam=(9, 8, 6), seed=na, flag_fraction=0.1)
    check(a, counts, dtype)


@pytest.mark.parametrize('residual,use_w', [(False, False), (True, False), (False, True)])
def test_shaped_matrix_pass_applies_epilogue(residual, use_w):
    """In float from 40 antennas the matrix-core kernel applies the
    epilogue itself (here k_forward_mx: point sources, their sums staged in
    LDS, then the gaussian and sersic sources): gains, flags, residual
    mode, weights, chi^2 terms and their sum, with baselines in either
    antenna order (the baseline's own order picks the gains) and
    auto-correlations."""
    from test_gpu_forward import check_outputs, run_forward
    counts = (5, 6, 4)
    a = synth.make_inputs(40, 2, 3, *counts, beam=(9, 8, 6), seed=77, flag_fraction=0.3,
                          auto_correlations=True, observed='model')
    a1, a2 = a['antenna1'].copy(), a['antenna2'].copy()
    flip = np.zeros(a1.shape[1], bool)
    flip[::3] = True
    a['antenna1'] = np.where(flip, a2, a1).astype(a1.dtype)
    a['antenna2'] = np.where(flip, a1, a2).astype(a2.dtype)
    if use_w:
        a['weight_vector'] = np.random.default_rng(3).uniform(0.5, 2.0, a['weight_vector'].shape)
    a['sigma_sqrd'] = 1.7
    ref = rime_np.rime(synth.upcast(synth.as_dtype(a, np.float32)), *counts, residual=residual,
                       use_weight_vector=use_w, sigma_sqrd=1.7)
    out = run_forward(a, counts, np.float32, residual=residual, use_weight_vector=use_w)
    check_outputs(out, ref, a, np.float32, residual, use_w, 1.7)


@pytest.mark.parametrize('residual,use_w,keep', [(False, False, True), (True, False, False),
                                                 (False, True, False)])
def test_float64_epilogue_kernel(residual, use_w, keep):
    """Float64 from 48 antennas: the point and shaped matrix-core passes sum
    every source, then k_epilogue (one thread per visibility) applies gains,
    flags, residual mode, weights and the chi^2 terms, with baselines in
    either antenna order and auto-correlations; with keep_ekb the pre-gain
    sums stay in ekb_vis."""
    from test_gpu_forward import check_outputs, run_forward
    counts = (5, 6, 4)
    a = synth.make_inputs(48, 2, 3, *counts, beam=(9, 8, 6), seed=81, flag_fraction=0.3,
                          auto_correlations=True, observed='model')
    a1, a2 = a['antenna1'].copy(), a['antenna2'].copy()
    flip = np.zeros(a1.shape[1], bool)
    flip[::3] = True
    a['antenna1'] = np.where(flip, a2, a1).astype(a1.dtype)
    a['antenna2'] = np.where(flip, a1, a2).astype(a2.dtype)
    if use_w:
        a['weight_vector'] = np.random.default_rng(4).uniform(0.5, 2.0, a['weight_vector'].shape)
    a['sigma_sqrd'] = 1.7
    ref = rime_np.rime(a, *counts, residual=residual, use_weight_vector=use_w, sigma_sqrd=1.7)
    out = run_forward(a, counts, np.float64, residual=residual, use_weight_vector=use_w, keep_ekb=keep)
    check_outputs(out, ref, a, np.float64, residual, use_w, 1.7)
    if keep:
        ekb = out['ekb_vis']
        ekb = ekb.cpu().numpy() if hasattr(ekb, 'cpu') else np.asarray(ekb)
        assert np.abs(ekb - ref['ekb_vis']).max() <= 1e-10 * np.abs(ref['ekb_vis']).max()


@pytest.mark.parametrize('dtype,na', [(np.float32, 44), (np.float64, 48)])
def test_shaped_matrix_pass_without_point_sources(dtype, na):
    """No point sources: the point pass is skipped and the shaped matrix
    pass (k_shaped_mx in float, on its own) starts from no running sum (and
    in float applies the epilogue)."""
    counts = (0, 6, 4)
    a = synth.make_inputs(na, 2, 3, *counts, beam=(9, 8, 6), seed=na + 1, flag_fraction=0.2)
    check(a, counts, dtype)
