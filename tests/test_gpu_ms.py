"""A MeasurementSet through rime_solver(data_source='ms') on the GPU.

The tables are injected in place of python-casacore's (absent here): the
columns come from a known geometry (test_ms_loader.make_columns), the
factory reads the dimensions and transfers the loaded arrays
(factory.py:113-148), the solver's fused forward runs on them, and the
result is compared with the oracle on the same arrays.
"""

import numpy as np
import pytest

from oracle import rime_np, synth
from test_ms_loader import make_columns

pytestmark = pytest.mark.gpu

TOL = {'double': 1e-10, 'float': 1e-5}


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize('dtype', ['double', 'float'])
def test_ms_data_source_forward_against_oracle(monkeypatch, dtype):
    import montblanc_amd as mb
    from montblanc_amd import ms_loader, parallactic as pa
    cols, _, _, (ntime, nbl, na, nbands, nchan) = make_columns(na=6, ntime=3, nbands=2, cpb=4)
    rng = np.random.default_rng(21)
    # spread the antennas over a few km so the phases are not trivial
    scale = 30.0
    cols['UVW'] = cols['UVW'] * scale
    cols['POSITION'] = pa.itrf(np.radians(-30.7 + 0.02 * np.arange(na)),
                               np.radians(21.4 + 0.02 * np.arange(na)), 1000.0)
    cols['PHASE_DIR'] = np.array([-1.1, -0.6])
    cols['TIME'] = 4.9e9 + np.repeat(np.arange(ntime) * 300.0, nbl * nbands)
    monkeypatch.setattr(ms_loader, 'open_tables',
                        lambda msfile, auto_correlations=False: ms_loader.ColumnTables(cols))
    cfg = mb.rime_solver_cfg(msfile='synthetic.ms', dtype=dtype, init_weights='weight',
                             weight_vector=True, sources=mb.sources(point=3, gaussian=2, sersic=1),
                             beam_lw=9, beam_mh=8, beam_nud=5)
    slvr = mb.rime_solver(cfg)
    assert slvr.dim_local_size('ntime', 'na', 'nbl', 'nchan') == (ntime, na, nbl, nchan)
    ft = slvr.ft
    # source parameters and the beam are not in a MeasurementSet
    slvr.transfer_lm(rng.uniform(-0.01, 0.01, (6, 2)).astype(ft))
    st = np.zeros((6, ntime, 4))
    st[..., 0] = rng.lognormal(0, 1, (6, 1))
    st[..., 1:] = st[..., :1] * rng.uniform(-0.1, 0.1, (6, 1, 3))
    slvr.transfer_stokes(st.astype(ft))
    slvr.transfer_alpha(rng.normal(-0.7, 0.2, (6, ntime)).astype(ft))
    slvr.transfer_gauss_shape(rng.uniform(-1e-4, 1e-4, (3, 2)).astype(ft))
    slvr.transfer_sersic_shape(np.array([[0.2], [-0.1], [5e-5]]).astype(ft))
    E = rng.uniform(0, 1, (9, 8, 5, 4)) + 1j * rng.uniform(0, 1, (9, 8, 5, 4))
    slvr.transfer_E_beam(E.astype(slvr.ct))
    slvr.set_beam_ll(-0.05); slvr.set_beam_lm(-0.05); slvr.set_beam_ul(0.05)
    slvr.set_beam_um(0.05); slvr.set_beam_lfreq(1.39e9); slvr.set_beam_ufreq(1.51e9)
    with slvr:
        slvr.solve()
        assert slvr.pipeline.last_error is None, slvr.pipeline.last_error
        mv, chi, x2 = slvr.retrieve_model_vis(), slvr.retrieve_chi_sqrd_result(), slvr.X2
    a = {}
    for name in ('uvw', 'antenna1', 'antenna2', 'frequency', 'ref_frequency', 'parallactic_angles',
                 'point_errors', 'antenna_scaling', 'E_beam', 'G_term', 'lm', 'stokes', 'alpha',
                 'gauss_shape', 'sersic_shape', 'flag', 'weight_vector', 'observed_vis'):
        a[name] = getattr(slvr, 'retrieve_' + name)()
    for k in rime_np.BEAM_KEYS + ('gauss_scale', 'two_pi_over_c', 'sigma_sqrd'):
        a[k] = float(getattr(slvr, k))
    # what the solver holds is what the loader read
    want = ms_loader.ms_arrays(cols, ft, init_weights='weight')
    for k, v in want.items():
        np.testing.assert_array_equal(a[k], v, err_msg=k)
    ref = rime_np.rime(synth.upcast(a), 3, 2, 1, use_weight_vector=True)
    assert rel(mv, ref['model_vis']) < TOL[dtype]
    assert abs(float(x2) - ref['X2']) <= TOL[dtype] * abs(ref['X2']) * 10
