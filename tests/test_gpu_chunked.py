"""Chunked (version 'v5') solver: the planner's chunks over time and
channels reproduce the single-chunk result, and vis_write_mode='sum'
accumulates across solves (montblanc/tests/test_rime_v5.py:73-104)."""

import numpy as np
import pytest

from oracle import rime_np, synth

pytestmark = pytest.mark.gpu

NAMES = ('uvw', 'antenna1', 'antenna2', 'frequency', 'ref_frequency', 'parallactic_angles',
         'point_errors', 'antenna_scaling', 'E_beam', 'G_term', 'lm', 'stokes', 'alpha',
         'gauss_shape', 'sersic_shape', 'flag', 'weight_vector', 'observed_vis')


def configure(version, dtype, counts, budget=None, write_mode='overwrite'):
    import montblanc_amd as mb
    kw = dict(na=9, ntime=6, nchan=8, dtype=dtype, version=version,
              sources=mb.sources(point=counts[0], gaussian=counts[1], sersic=counts[2]),
              beam_lw=9, beam_mh=8, beam_nud=6, data_source='default', vis_write_mode=write_mode)
    if budget is not None:
        kw['mem_budget'] = budget
    return mb.rime_solver(mb.rime_solver_cfg(**kw))


def load(slvr, a, ft):
    ct = np.complex64 if ft == np.float32 else np.complex128
    for n in NAMES:
        v = a[n]
        if v.dtype.kind == 'f':
            v = v.astype(ft)
        elif v.dtype.kind == 'c':
            v = v.astype(ct)
        else:
            v = v.astype(slvr.retrieve_antenna1().dtype if n.startswith('antenna') else np.uint8)
        getattr(slvr, 'transfer_' + n)(v)
    for k in rime_np.BEAM_KEYS:
        getattr(slvr, 'set_' + k)(a[k])


@pytest.mark.parametrize('dtype', ['double', 'float'])
def test_chunked_solver_matches_resident_solver_and_oracle(dtype):
    ft = np.float64 if dtype == 'double' else np.float32
    counts = (4, 2, 1)
    a = synth.make_inputs(9, 6, 8, *counts, beam=(9, 8, 6), seed=19, flag_fraction=0.3)
    ref = rime_np.rime(synth.upcast(synth.as_dtype(a, ft)), *counts)
    v4 = configure('v4', dtype, counts)
    load(v4, a, ft)
    v4.solve()
    # about the whole problem's size: forces at least two time chunks
    v5 = configure('v5', dtype, counts, budget=(400 if dtype == 'double' else 200) * 1024)
    load(v5, a, ft)
    v5.solve()
    plan = v5.last_plan
    print('plan', plan['modified'], flush=True)
    assert plan['chunk_dims']['ntime'] < 6 or plan['chunk_dims']['nchan'] < 8
    tol = 1e-10 if dtype == 'double' else 1e-5
    m4, m5 = v4.retrieve_model_vis(), v5.retrieve_model_vis()
    assert np.abs(m5 - m4).max() <= tol * np.abs(m4).max()
    assert np.abs(m5 - ref['model_vis']).max() <= tol * np.abs(ref['model_vis']).max()
    # chi^2 holds the model error propagated through the residual
    # R = O - M (Cauchy-Schwarz): |dX2| <= 2 ||R|| tol ||M||; for float64
    # this is within the plain relative bound as well
    up = synth.upcast(synth.as_dtype(a, ft))
    keep = up['flag'] == 0
    R = (up['observed_vis'] - ref['model_vis']) * keep
    bound = 2 * np.linalg.norm(R) * tol * np.linalg.norm(ref['model_vis'] * keep)
    if dtype == 'double':
        bound = min(bound, tol * ref['X2'])
    assert abs(float(v5.X2) - ref['X2']) <= bound
    assert abs(float(v5.X2) - float(v4.X2)) <= bound


def test_sum_write_mode_accumulates():
    counts = (3, 0, 0)
    a = synth.make_inputs(9, 6, 8, *counts, beam=(9, 8, 6), seed=2)
    s = configure('v5', 'double', counts, budget=400 * 1024, write_mode='sum')
    load(s, a, np.float64)
    s.solve()
    once = s.retrieve_model_vis().copy()
    for _ in range(4):
        s.solve()
    np.testing.assert_allclose(s.retrieve_model_vis(), 5 * once, rtol=1e-12, atol=1e-12)


def test_pipelined_chunks_keep_unchanged_arrays_on_device():
    """The chunk loop (inputs of the next chunk copied in while a chunk
    computes) sends the arrays without time or channel axes once; a second
    solve sends none of them and gives the same result, and after
    transfer_lm only lm is sent again, with the result following the new
    positions (the oracle on the new inputs)."""
    counts = (4, 2, 1)
    a = synth.make_inputs(9, 6, 8, *counts, beam=(9, 8, 6), seed=41, flag_fraction=0.2)
    s = configure('v5', 'double', counts, budget=300 * 1024)
    load(s, a, np.float64)
    s.solve()
    first = s.retrieve_model_vis().copy()
    print('solve 1', s.last_solve, s.last_plan['chunk_dims'], flush=True)
    assert s.last_solve['chunks'] >= 3
    assert sorted(s.last_solve['static_sent']) == ['E_beam', 'gauss_shape', 'lm', 'sersic_shape']
    s.solve()
    assert s.last_solve['static_sent'] == []
    np.testing.assert_array_equal(s.retrieve_model_vis(), first)
    b = dict(a)
    b['lm'] = a['lm'] + 1e-3
    s.transfer_lm(b['lm'].astype(np.float64))
    s.solve()
    assert s.last_solve['static_sent'] == ['lm']
    ref = rime_np.rime(b, *counts)
    assert np.abs(s.retrieve_model_vis() - ref['model_vis']).max() <= 1e-10 * np.abs(ref['model_vis']).max()
    assert abs(float(s.X2) - ref['X2']) <= 1e-10 * ref['X2']
