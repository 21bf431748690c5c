"""GPU parity of the forward path against the numpy oracle.

Tolerances (stated by the task): float64 within 1e-10 relative, float32
(with a float64 phase argument and float64 beam coordinates) within 1e-5
relative, both against the float64 oracle. Relative error is
max|gpu - ref| / max|ref| over the array (norm-wise, max norm), which is
robust to the exact zeros flags produce.
"""

import numpy as np
import pytest

from oracle import rime_np, synth

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-10, np.float32: 1e-5}


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def check_outputs(out, ref, a, dtype, residual, use_w, sigma_sqrd):
    """Outputs vs the oracle with the stated tolerance.

    float64: every output within 1e-10 relative.
    float32: the model visibilities within 1e-5 relative. Residuals and
    chi^2 are differences against the observed data, which cancel when
    the model fits (|model| ~ 25 against |residual| ~ 0.1 here), so they
    are held to that same model error propagated through the residual:
      |dR| <= tol |M|max,
      |d term| <= 2 * 4 * w_max * |R|max * tol |M|max,
      |dX2| <= 2 w_max ||R||_2 tol ||M||_2 / sigma^2   (Cauchy-Schwarz).
    """
    tol = TOL[dtype]
    if dtype == np.float64:
        assert rel(out['model_vis'], ref['model_vis']) < tol
        assert rel(out['chi_sqrd_result'], ref['chi_sqrd_result']) < tol
        assert abs(out['X2'] - ref['X2']) < tol * abs(ref['X2'])
        return
    up = synth.upcast(synth.as_dtype(a, dtype))
    unflagged = up['flag'] == 0
    M = ref['model_vis'] if not residual else (up['observed_vis'] - ref['model_vis']) * unflagged
    R = (up['observed_vis'] - M) * unflagged
    mmax, rmax = np.abs(M).max(), np.abs(R).max()
    wmax = up['weight_vector'].max() if use_w else 1.0
    if residual:
        assert np.abs(out['model_vis'] - ref['model_vis']).max() <= tol * mmax
    else:
        assert rel(out['model_vis'], ref['model_vis']) < tol
    dterm = np.abs(out['chi_sqrd_result'] - ref['chi_sqrd_result']).max()
    assert dterm <= 8 * wmax * rmax * tol * mmax
    bound = 2 * wmax * np.linalg.norm(R) * tol * np.linalg.norm(M) / (1.0 if use_w else sigma_sqrd)
    assert abs(out['X2'] - ref['X2']) <= bound


def run_forward(a, counts, dtype, **kw):
    from montblanc_amd import rime
    print('forward', np.dtype(dtype).name, a['uvw'].shape, counts, kw, flush=True)
    return rime.forward_numpy(synth.as_dtype(a, dtype), counts, **kw)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('residual,use_w', [(False, False), (True, False), (False, True)])
def test_fused_forward_cpu_parity_workload(dtype, residual, use_w):
    w = dict(synth.WORKLOADS['cpu_parity_14ant'])
    a = synth.make_inputs(**w, seed=1000, flag_fraction=0.5, observed='model')
    if use_w:
        a['weight_vector'] = np.random.default_rng(1).uniform(0.5, 2.0, a['weight_vector'].shape)
    a['sigma_sqrd'] = 1.7
    counts = (w['npsrc'], w['ngsrc'], w['nssrc'])
    ref = rime_np.rime(synth.upcast(synth.as_dtype(a, dtype)), *counts, residual=residual,
                       use_weight_vector=use_w, sigma_sqrd=1.7)
    out = run_forward(a, counts, dtype, residual=residual, use_weight_vector=use_w)
    check_outputs(out, ref, a, dtype, residual, use_w, 1.7)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('counts', [(1, 0, 0), (0, 3, 0), (0, 0, 2), (2, 2, 2), (37, 5, 3)])
def test_fused_forward_source_mixes(dtype, counts):
    a = synth.make_inputs(9, 3, 5, *counts, beam=(11, 9, 7), seed=sum(counts), flag_fraction=0.2)
    a['sersic_shape'][:2] = np.random.default_rng(2).uniform(-0.3, 0.3, (2, counts[2]))
    ref = rime_np.rime(synth.upcast(synth.as_dtype(a, dtype)), *counts)
    out = run_forward(a, counts, dtype, keep_ekb=True)
    assert rel(out['ekb_vis'], ref['ekb_vis']) < TOL[dtype]
    assert rel(out['model_vis'], ref['model_vis']) < TOL[dtype]
    assert abs(out['X2'] - ref['X2']) < TOL[dtype] * abs(ref['X2'])


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_fused_forward_autocorrelations_and_degenerate_beam(dtype):
    a = synth.make_inputs(6, 2, 3, 3, 2, 0, beam=(2, 2, 2), seed=17, auto_correlations=True)
    ref = rime_np.rime(synth.upcast(synth.as_dtype(a, dtype)), 3, 2, 0)
    out = run_forward(a, (3, 2, 0), dtype)
    assert rel(out['model_vis'], ref['model_vis']) < TOL[dtype]


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_fused_forward_many_baselines_chunked(dtype):
    """na=100 -> 4950 baselines: more than one baseline chunk per workgroup."""
    a = synth.make_inputs(100, 1, 2, 3, 1, 0, beam=(9, 9, 5), seed=23)
    ref = rime_np.rime(synth.upcast(synth.as_dtype(a, dtype)), 3, 1, 0)
    out = run_forward(a, (3, 1, 0), dtype)
    assert rel(out['model_vis'], ref['model_vis']) < TOL[dtype]
    assert abs(out['X2'] - ref['X2']) < TOL[dtype] * abs(ref['X2'])


@pytest.mark.parametrize('dtype,na', [(np.float64, 8), (np.float32, 8), (np.float64, 48), (np.float32, 40)])
def test_source_batches_match_single_launch(dtype, na):
    """Reference batch semantics: lower extent > 0 reads the running sum,
    upper extent < global writes it (RimeSumCoherencies.py:183-190, 307-316).
    From 40 (float) / 48 (float64) antennas the batches run through the
    matrix-core passes, in float with the epilogue in the shaped pass."""
    import torch
    from montblanc_amd import _lib, rime
    counts = (7, 5, 3)
    a = synth.make_inputs(na, 3, 4, *counts, beam=(9, 9, 5), seed=31, flag_fraction=0.3)
    arrays = rime.to_device(synth.as_dtype(a, dtype))
    whole = rime.forward(arrays, counts)
    torch.cuda.synchronize()
    out = rime.alloc_outputs(arrays, keep_ekb=True)
    # batches over the global source index: [0,4) [4,10) [10,15)
    from montblanc_amd import src_types
    nr = dict(npsrc=7, ngsrc=5, nssrc=3)
    for lo, hi in ((0, 4), (4, 10), (10, 15)):
        rng = src_types.source_range_tuple(lo, hi, nr)
        sub = dict(arrays)
        sub['lm'] = arrays['lm'][lo:hi].contiguous()
        sub['stokes'] = arrays['stokes'][lo:hi].contiguous()
        sub['alpha'] = arrays['alpha'][lo:hi].contiguous()
        gs, ge = rng['ngsrc']
        ss, se = rng['nssrc']
        sub['gauss_shape'] = arrays['gauss_shape'][:, gs:ge].contiguous()
        sub['sersic_shape'] = arrays['sersic_shape'][:, ss:se].contiguous()
        bc = tuple(e - s for s, e in rng.values())
        ext = {'nsrc': (lo, lo + sum(bc))}
        glob = {'nsrc': 15, 'npsrc': 7, 'ngsrc': 5, 'nssrc': 3}
        sizes = rime.make_dims(sub, bc)
        d = _lib.make_dims({n: getattr(sizes, n).local_size for n in _lib.DIM_NAMES},
                           extents=ext, global_sizes=glob)
        rime.forward(sub, bc, outputs=out, dims=d)
    torch.cuda.synchronize()
    assert rel(out['model_vis'].cpu().numpy(), whole['model_vis'].cpu().numpy()) < TOL[dtype]
    assert abs(out['X2'].item() - whole['X2'].item()) < TOL[dtype] * abs(whole['X2'].item())


def test_forward_rejects_bad_arguments():
    import ctypes
    from montblanc_amd import _lib, rime
    a = synth.make_inputs(4, 1, 2, 1, 0, 0, beam=(3, 3, 3), seed=1)
    arrays = rime.to_device(a)
    with pytest.raises(ValueError, match='lm'):
        rime.forward(arrays, (1, 1, 0))       # counts disagree with the source arrays
    bad = dict(arrays)
    bad['beam_ul'] = bad['beam_ll']
    with pytest.raises(_lib.MontblancHipError) as e:   # the library's own validation
        rime.forward(bad, (1, 0, 0))
    assert e.value.code < 0


VARIANTS = ['generic', (1, 1, -1), (2, 2, -1), (2, 4, -1), (1, 1, 0), (2, 2, 0), (2, 2, 3),
            (2, 4, 1)]


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('pair_block', VARIANTS)
@pytest.mark.parametrize('na,auto', [(9, False), (11, True), (64, False)])
def test_kernel_variants_agree_with_oracle(dtype, pair_block, na, auto):
    """Every kernel variant (generic per-baseline, and pair blocks of 1x1,
    2x2, 2x4 with ragged last antenna groups) against the oracle."""
    counts = (3, 2, 1)
    a = synth.make_inputs(na, 2, 3, *counts, beam=(9, 9, 5), seed=na + 100 * auto,
                          flag_fraction=0.3, auto_correlations=auto)
    ref = rime_np.rime(synth.upcast(synth.as_dtype(a, dtype)), *counts)
    out = run_forward(a, counts, dtype, pair_block=pair_block, keep_ekb=True)
    assert rel(out['ekb_vis'], ref['ekb_vis']) < TOL[dtype]
    assert rel(out['model_vis'], ref['model_vis']) < TOL[dtype]
    assert abs(out['X2'] - ref['X2']) < TOL[dtype] * abs(ref['X2'])


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_time_varying_layout_uses_generic_kernel(dtype):
    """Antenna pairs that change between time steps cannot be blocked: the
    plan builder declines and the generic kernel runs."""
    counts = (4, 1, 0)
    a = synth.make_inputs(6, 3, 2, *counts, beam=(7, 7, 4), seed=77)
    r = np.random.default_rng(5)
    for t in range(3):
        perm = r.permutation(a['antenna1'].shape[1])
        a['antenna1'][t] = a['antenna1'][t][perm]
        a['antenna2'][t] = a['antenna2'][t][perm]
    ref = rime_np.rime(synth.upcast(synth.as_dtype(a, dtype)), *counts)
    out = run_forward(a, counts, dtype)
    assert rel(out['model_vis'], ref['model_vis']) < TOL[dtype]
