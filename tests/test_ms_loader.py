"""MeasurementSet columns -> solver arrays (montblanc_amd/ms_loader.py),
following montblanc/impl/rime/v4/loaders/loaders.py:49-253, on columns
built here from a known geometry (python-casacore, which reads real
tables, is not installed: read_ms reports that)."""

import numpy as np
import pytest

from montblanc_amd import ms_loader


def make_columns(na=5, ntime=3, nbands=2, cpb=3, auto=False, reverse=()):
    """Rows ordered (time, antenna1, antenna2, band) as the loader's view."""
    r = np.random.default_rng(7)
    ant_uvw = r.uniform(-100, 100, (ntime, na, 3))
    pairs = [(p, q) for p in range(na) for q in range(p if auto else p + 1, na)]
    pairs = [(q, p) if (p, q) in reverse else (p, q) for p, q in pairs]
    nbl = len(pairs)
    rows = []
    for t in range(ntime):
        for p, q in pairs:
            for b in range(nbands):
                rows.append((t, p, q, b))
    T = np.array([10.0 + t for t, _, _, _ in rows])
    A1 = np.array([p for _, p, _, _ in rows])
    A2 = np.array([q for _, _, q, _ in rows])
    UVW = np.array([ant_uvw[t, q] - ant_uvw[t, p] for t, p, q, _ in rows])
    n = len(rows)
    cols = dict(TIME=T, ANTENNA1=A1, ANTENNA2=A2, UVW=UVW,
                DATA=r.standard_normal((n, cpb, 4)) + 1j * r.standard_normal((n, cpb, 4)),
                FLAG=r.uniform(size=(n, cpb, 4)) < 0.2, FLAG_ROW=r.uniform(size=n) < 0.1,
                WEIGHT=r.uniform(0.5, 2.0, (n, 4)), SIGMA=r.uniform(0.5, 2.0, (n, 4)),
                CHAN_FREQ=np.linspace(1.4e9, 1.5e9, nbands * cpb).reshape(nbands, cpb),
                REF_FREQUENCY=np.array([1.4e9 + 1e7 * b for b in range(nbands)]),
                NUM_CHAN=np.full(nbands, cpb), NA=na)
    return cols, ant_uvw, pairs, (ntime, nbl, na, nbands, nbands * cpb)


@pytest.mark.parametrize('auto,reverse', [(False, ()), (True, ((0, 3),))])
def test_columns_to_solver_arrays(auto, reverse):
    cols, ant_uvw, pairs, dims = make_columns(auto=auto, reverse=reverse)
    assert ms_loader.ms_dims(cols) == dims
    ntime, nbl, na, nbands, nchan = dims
    a = ms_loader.ms_arrays(cols, np.float64, init_weights='weight')
    np.testing.assert_array_equal(a['frequency'], cols['CHAN_FREQ'].reshape(-1))
    np.testing.assert_array_equal(a['ref_frequency'], np.repeat(cols['REF_FREQUENCY'], 3))
    # per-antenna uvw: antenna 0 at the origin, every baseline = uvw_q - uvw_p
    uvw = a['uvw']
    np.testing.assert_allclose(uvw[:, 0], 0)
    np.testing.assert_allclose(uvw, ant_uvw - ant_uvw[:, :1], atol=1e-9)
    for b, (p, q) in enumerate(pairs):
        np.testing.assert_array_equal(a['antenna1'][:, b], p)
        np.testing.assert_array_equal(a['antenna2'][:, b], q)
        np.testing.assert_allclose(uvw[:, q] - uvw[:, p],
                                   cols['UVW'][::nbands].reshape(ntime, nbl, 3)[:, b], atol=1e-9)
    # visibilities: a (time, baseline) row's bands side by side
    D = cols['DATA'].reshape(ntime, nbl, nbands, 3, 4)
    np.testing.assert_array_equal(a['observed_vis'][:, :, 3:6], D[:, :, 1])
    F = cols['FLAG'] | cols['FLAG_ROW'][:, None, None]
    np.testing.assert_array_equal(a['flag'], F.reshape(ntime, nbl, nchan, 4).astype(np.uint8))
    W = np.repeat(cols['WEIGHT'].reshape(ntime, nbl, nbands, 1, 4), 3, axis=3).reshape(ntime, nbl, nchan, 4)
    np.testing.assert_allclose(a['weight_vector'], W)
    s = ms_loader.ms_arrays(cols, np.float32, init_weights='sigma')
    assert s['weight_vector'].dtype == np.float32 and s['observed_vis'].dtype == np.complex64
    np.testing.assert_allclose(s['weight_vector'][0, 0, 0], cols['SIGMA'][0], rtol=1e-6)


def test_parallactic_angles_from_positions():
    from montblanc_amd import parallactic as pa
    cols, _, _, (ntime, nbl, na, nbands, nchan) = make_columns()
    pos = pa.itrf(np.radians(-30.7 + 0.01 * np.arange(na)), np.radians(21.4 + 0.01 * np.arange(na)), 1000.0)
    cols = dict(cols, POSITION=pos, PHASE_DIR=np.array([1.2, -0.5]), TIME=cols['TIME'] * 0 + 4.9e9 +
                np.repeat(np.arange(ntime) * 300.0, nbl * nbands))
    a = ms_loader.ms_arrays(cols, np.float64)
    want = pa.parallactic_angles(np.array([1.2, -0.5]), pos, 4.9e9 + np.arange(ntime) * 300.0)
    assert a['parallactic_angles'].shape == (ntime, na)
    np.testing.assert_allclose(a['parallactic_angles'], want, atol=1e-15)


def test_layout_checks_and_missing_casacore():
    cols, _, _, _ = make_columns()
    bad = dict(cols, NUM_CHAN=np.array([3, 4]))
    with pytest.raises(ValueError, match='not equal'):
        ms_loader.ms_dims(bad)
    short = {k: (v[:-1] if isinstance(v, np.ndarray) and v.shape[:1] == cols['TIME'].shape else v)
             for k, v in cols.items()}
    with pytest.raises(ValueError, match='rows'):
        ms_loader.ms_dims(short)
    no_weights = {k: v for k, v in cols.items() if k not in ('WEIGHT', 'SIGMA')}
    np.testing.assert_array_equal(ms_loader.ms_arrays(no_weights, init_weights='weight')['weight_vector'], 1.0)
    with pytest.raises(ValueError, match='init_weights'):
        ms_loader.ms_arrays(cols, init_weights='bogus')
    try:
        import casacore.tables  # noqa: F401
    except ImportError:
        with pytest.raises(ImportError, match='python-casacore'):
            ms_loader.read_ms('/nonexistent.ms')


def test_row_increments_match_whole_columns():
    """Reading the main table in increments of whole time steps (v5 loader,
    loaders.py:160-200) gives the arrays a whole-column read gives, for
    increments that do and do not divide the row count."""
    cols, _, _, (ntime, nbl, na, nbands, nchan) = make_columns(na=6, ntime=7)
    assert ms_loader.row_increment(nbl, nbands) % (nbl * nbands) == 0
    assert ms_loader.row_increment(nbl, nbands) >= 5000
    whole = ms_loader.ms_arrays(cols, np.float64, init_weights='weight', min_rows=10 ** 9)
    for min_rows in (1, nbl * nbands * 2, nbl * nbands * 3 + 1):
        part = ms_loader.ms_arrays(cols, np.float64, init_weights='weight', min_rows=min_rows)
        assert sorted(part) == sorted(whole)
        for k in whole:
            np.testing.assert_array_equal(part[k], whole[k], err_msg=k)


class _RecordingSolver(object):
    """Stands in for the solver class: records the configuration it was
    built with and every transfer_<name>()."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.ft = np.float32 if cfg['dtype'] == 'float' else np.float64
        self.got = {}

    def __getattr__(self, name):
        if name.startswith('transfer_'):
            return lambda ary, _n=name[9:]: self.got.__setitem__(_n, ary)
        raise AttributeError(name)


def test_factory_ms_data_source(monkeypatch):
    """data_source='ms' (factory.py:113-148): KeyError without msfile; the
    loader's dimensions go into the configuration the solver is built with;
    then the loader transfers the arrays. The tables are injected in place
    of python-casacore's."""
    import montblanc_amd
    from montblanc_amd.config import RimeSolverConfig as O
    cols, ant_uvw, pairs, (ntime, nbl, na, nbands, nchan) = make_columns(na=6, ntime=4)
    opened = []

    def fake_open(msfile, auto_correlations=False):
        opened.append((msfile, auto_correlations))
        return ms_loader.ColumnTables(cols)
    monkeypatch.setattr(ms_loader, 'open_tables', fake_open)
    cfg = montblanc_amd.rime_solver_cfg(dtype='float', init_weights='sigma')
    assert cfg[O.DATA_SOURCE] == O.DATA_SOURCE_MS
    with pytest.raises(KeyError, match='msfile'):
        montblanc_amd.create_rime_solver_from_ms(_RecordingSolver, cfg)
    cfg['msfile'] = 'obs.ms'
    s = montblanc_amd.create_rime_solver_from_ms(_RecordingSolver, cfg)
    assert opened == [('obs.ms', False)]
    assert (s.cfg['ntime'], s.cfg['na'], s.cfg['nbl'], s.cfg['nbands'], s.cfg['nchan']) == \
        (ntime, na, nbl, nbands, nchan)
    want = ms_loader.ms_arrays(cols, np.float32, init_weights='sigma')
    assert sorted(s.got) == sorted(want)
    for k in want:
        assert s.got[k].dtype == want[k].dtype
        np.testing.assert_array_equal(s.got[k], want[k], err_msg=k)
    # rime_solver dispatches 'ms' to the same path (its solver needs a GPU,
    # so only the missing-msfile error is reached here)
    with pytest.raises(KeyError, match='msfile'):
        montblanc_amd.rime_solver(montblanc_amd.rime_solver_cfg())


def test_negative_right_ascension_is_wrapped():
    from montblanc_amd import parallactic as pa
    cols, _, _, (ntime, nbl, na, nbands, nchan) = make_columns()
    pos = pa.itrf(np.radians(-30.7 + 0.01 * np.arange(na)), np.radians(21.4 + 0.01 * np.arange(na)), 1000.0)
    t = 4.9e9 + np.repeat(np.arange(ntime) * 300.0, nbl * nbands)
    neg = ms_loader.ms_arrays(dict(cols, POSITION=pos, PHASE_DIR=np.array([-1.0, -0.5]), TIME=t))
    pos_ra = ms_loader.ms_arrays(dict(cols, POSITION=pos, PHASE_DIR=np.array([2 * np.pi - 1.0, -0.5]),
                                      TIME=t))
    np.testing.assert_allclose(neg['parallactic_angles'], pos_ra['parallactic_angles'], atol=1e-14)
