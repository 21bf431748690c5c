"""Chunk planner (CPU): reduction-ladder semantics, sizing for one
288 GB device, and the MemoryError path a small budget forces."""

import numpy as np
import pytest

from montblanc_amd import planner

GB = 1024 ** 3


def dims_of(na, ntime, nchan, npsrc, ngsrc, nssrc=0, beam=(65, 65, 32)):
    d = dict(ntime=ntime, na=na, nbl=na * (na - 1) // 2, nchan=nchan, npol=4,
             npsrc=npsrc, ngsrc=ngsrc, nssrc=nssrc, nsrc=npsrc + ngsrc + nssrc,
             beam_lw=beam[0], beam_mh=beam[1], beam_nud=beam[2])
    return d


def test_reduction_strings_follow_viable_dim_config():
    """montblanc/util/__init__.py:166-285 semantics on a toy cost model."""
    cost = lambda T: T['ntime'] * T['nchan'] * 10
    ok, mod, T = planner.apply_reductions(1000, dict(ntime=100, nchan=64), ['ntime=20'], cost)
    assert not ok and mod == {'ntime': 20}
    ok, mod, T = planner.apply_reductions(1000, dict(ntime=100, nchan=64),
                                          ['ntime=20', 'ntime', 'nchan=50%'], cost)
    assert ok and T['ntime'] == 1 and T['nchan'] == 64   # 1 x 64 x 10 fits: no channel step
    ok, mod, T = planner.apply_reductions(100, dict(ntime=100, nchan=64),
                                          ['ntime', 'nchan=50%'], cost)
    assert ok and T == {'ntime': 1, 'nchan': 8}
    # percentages repeat until they reach 1; '&' reduces together
    ok, mod, T = planner.apply_reductions(10, dict(nbl=100, na=14), ['nbl=1&na=2'],
                                          lambda T: T['nbl'] * T['na'])
    assert ok and T == {'nbl': 1, 'na': 2}
    # a reduction never raises a dimension
    ok, mod, T = planner.apply_reductions(0, dict(ntime=5), ['ntime=20'], lambda T: T['ntime'])
    assert not ok and T['ntime'] == 5 and mod == {}
    with pytest.raises(ValueError):
        planner.apply_reductions(0, dict(ntime=5), ['ntime==3'], lambda T: 1)


def test_large_workload_fits_one_device_without_chunking():
    """64 antennas, 500 times, 256 channels, 12000 sources, float32 with the
    gradient's buffers: resident on one 288 GB device in one chunk."""
    d = dims_of(64, 500, 256, 10000, 2000, beam=(129, 129, 64))
    mod, T, need = planner.plan(d, np.float32, 288 * GB - planner.DEVICE_RESERVE, gradients=True)
    assert mod == {} and T['ntime'] == 500
    assert 30 * GB < need < 60 * GB
    # the reference's staged pipeline would need the jones scratch (TBs)
    staged = planner.device_bytes(dict(d, npolchan=1024, nvis=500 * 2016 * 256), np.float32,
                                  pipeline='staged')
    assert staged > 1024 * GB


def test_small_budget_halves_time_then_channels_then_fails():
    d = dims_of(64, 500, 256, 10000, 2000, beam=(129, 129, 64))
    mod, T, need = planner.plan(d, np.float32, 8 * GB)
    assert T['ntime'] < 500 and need <= 8 * GB
    assert T['ntime'] in (250, 125, 62, 31, 15, 7, 3, 1)
    mod, T, need = planner.plan(d, np.float32, 120 * 1024 ** 2)
    assert T['ntime'] == 1 and T['nchan'] < 256
    with pytest.raises(MemoryError):
        planner.plan(d, np.float32, 1024)


def test_chunks_cover_exactly():
    assert planner.chunks(10, 4) == [(0, 4), (4, 8), (8, 10)]
    assert planner.chunks(3, 10) == [(0, 3)]
