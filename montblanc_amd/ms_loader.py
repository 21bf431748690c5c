"""MeasurementSet input for the v4 solver: the MS columns turned into the
solver's arrays.

Reference: montblanc/impl/common/loaders/loaders.py:35-137 (the ordered
view of the main table and the dimensions) and
montblanc/impl/rime/v4/loaders/loaders.py:49-253 (the arrays). The main
table's rows are taken ordered by (TIME, ANTENNA1, ANTENNA2, band), field
0, auto-correlations left out unless asked for; then

* frequency: the bands' CHAN_FREQ one after the other; ref_frequency:
  each band's REF_FREQUENCY repeated over its channels;
* uvw per antenna from the baselines' UVW (one row per time and
  baseline): antenna 0 at the origin and antenna k at the UVW of baseline
  (0, k), so that a baseline (p, q) is uvw_q - uvw_p. The reference takes
  the first na - 1 baselines, which are these in the triangular order
  without auto-correlations; here baseline (0, k) is found wherever it is
  (negated when stored as (k, 0));
* antenna1 / antenna2 per time and baseline; observed_vis from DATA;
  flag = FLAG or FLAG_ROW; weight_vector from WEIGHT_SPECTRUM, else WEIGHT
  broadcast over the band's channels, else ones (init_weights 'weight'),
  or the same from SIGMA_SPECTRUM / SIGMA ('sigma').

* parallactic_angles from the ANTENNA POSITIONs, the FIELD's PHASE_DIR
  and the unique TIMEs (montblanc_amd/parallactic.py; the reference uses
  casacore measures), when those columns are given.

Reading the tables needs python-casacore, which ``read_ms`` imports and
reports clearly when absent; ``ms_dims`` and ``ms_arrays`` work on column
arrays however they were read.
"""

import numpy as np

MAIN_COLUMNS = ('TIME', 'UVW', 'ANTENNA1', 'ANTENNA2', 'DATA', 'FLAG', 'FLAG_ROW', 'WEIGHT',
                'WEIGHT_SPECTRUM', 'SIGMA', 'SIGMA_SPECTRUM')


def _tables():
    try:
        import casacore.tables as pt
    except ImportError:
        try:
            import pyrap.tables as pt
        except ImportError:
            raise ImportError('reading a MeasurementSet needs python-casacore (casacore.tables), '
                              'which is not installed; ms_arrays() takes the columns read '
                              'by other means')
    return pt


def read_ms(msfile, auto_correlations=False):
    """The columns ms_dims / ms_arrays use, read with python-casacore:
    main-table columns (rows ordered by time, baselines, band), the
    SPECTRAL_WINDOW's CHAN_FREQ, REF_FREQUENCY, NUM_CHAN and the antenna
    count."""
    pt = _tables()
    ms = pt.table(msfile, ack=False)
    query = ' '.join(['SELECT FROM $ms WHERE FIELD_ID=0',
                      '' if auto_correlations else 'AND ANTENNA1 != ANTENNA2',
                      'ORDERBY TIME, ANTENNA1, ANTENNA2, '
                      '[SELECT SPECTRAL_WINDOW_ID FROM ::DATA_DESCRIPTION][DATA_DESC_ID]'])
    main = pt.taql(query)
    cols = {c: main.getcol(c) for c in MAIN_COLUMNS if c in main.colnames()}
    spw = pt.table('::'.join((msfile, 'SPECTRAL_WINDOW')), ack=False, readonly=True)
    ant = pt.table('::'.join((msfile, 'ANTENNA')), ack=False, readonly=True)
    field = pt.table('::'.join((msfile, 'FIELD')), ack=False, readonly=True)
    cols.update(CHAN_FREQ=spw.getcol('CHAN_FREQ'), REF_FREQUENCY=spw.getcol('REF_FREQUENCY'),
                NUM_CHAN=spw.getcol('NUM_CHAN'), NA=ant.nrows(), POSITION=ant.getcol('POSITION'),
                PHASE_DIR=field.getcol('PHASE_DIR')[0][0])
    for t in (main, spw, ant, field, ms):
        t.close()
    return cols


def ms_dims(cols):
    """(ntime, nbl, na, nbands, nchan) of the ordered rows; the bands must
    have equal channel counts and the rows fill ntime x nbl x nbands."""
    num_chan = np.asarray(cols['NUM_CHAN'])
    if not np.all(num_chan == num_chan[0]):
        raise ValueError('channels per band %s are not equal' % list(num_chan))
    ntime = len(np.unique(cols['TIME']))
    pairs = np.stack([np.asarray(cols['ANTENNA1']), np.asarray(cols['ANTENNA2'])], 1)
    nbl = len(np.unique(pairs, axis=0))
    nbands, nchan = len(num_chan), int(num_chan.sum())
    rows = len(cols['TIME'])
    if rows != ntime * nbl * nbands:
        raise ValueError('%d rows are not ntime x nbl x nbands = %d x %d x %d' % (rows, ntime, nbl, nbands))
    return ntime, nbl, int(cols['NA']), nbands, nchan


def ms_arrays(cols, ft=np.float64, init_weights=None):
    """Solver arrays (v4 names) from MS columns (see the module docstring)."""
    ntime, nbl, na, nbands, nchan = ms_dims(cols)
    ct = np.complex64 if np.dtype(ft) == np.float32 else np.complex128
    cpb = nchan // nbands
    out = {}
    out['frequency'] = np.asarray(cols['CHAN_FREQ'], ft).reshape(nchan)
    out['ref_frequency'] = np.repeat(np.asarray(cols['REF_FREQUENCY'], ft), np.asarray(cols['NUM_CHAN']))
    # one row per (time, baseline): the first band's
    uvw_bl = np.asarray(cols['UVW'])[::nbands].reshape(ntime, nbl, 3)
    a1 = np.asarray(cols['ANTENNA1'])[::nbands].reshape(ntime, nbl).astype(np.int32)
    a2 = np.asarray(cols['ANTENNA2'])[::nbands].reshape(ntime, nbl).astype(np.int32)
    uvw = np.zeros((ntime, na, 3), ft)
    for k in range(1, na):
        fwd = np.flatnonzero((a1[0] == 0) & (a2[0] == k))
        rev = np.flatnonzero((a1[0] == k) & (a2[0] == 0))
        if len(fwd):
            uvw[:, k] = uvw_bl[:, fwd[0]]
        elif len(rev):
            uvw[:, k] = -uvw_bl[:, rev[0]]
        else:
            raise ValueError('no baseline between antenna 0 and antenna %d to place it' % k)
    out['uvw'] = uvw
    out['antenna1'] = a1
    out['antenna2'] = a2
    if 'POSITION' in cols and 'PHASE_DIR' in cols:
        from .parallactic import parallactic_angles
        times = np.unique(np.asarray(cols['TIME'], np.float64))
        out['parallactic_angles'] = parallactic_angles(np.asarray(cols['PHASE_DIR'], np.float64),
                                                       cols['POSITION'], times).astype(ft)
    if 'DATA' in cols:
        out['observed_vis'] = np.asarray(cols['DATA']).reshape(ntime, nbl, nchan, 4).astype(ct)
    if 'FLAG' in cols:
        flag = np.asarray(cols['FLAG'], bool)
        if 'FLAG_ROW' in cols:
            flag = flag | np.asarray(cols['FLAG_ROW'], bool)[:, None, None]
        out['flag'] = flag.reshape(ntime, nbl, nchan, 4).astype(np.uint8)
    if init_weights is not None:
        if init_weights not in ('weight', 'sigma'):
            raise ValueError("init_weights must be None, 'weight' or 'sigma' (got %r)" % (init_weights,))
        spectrum, row = (('WEIGHT_SPECTRUM', 'WEIGHT') if init_weights == 'weight'
                         else ('SIGMA_SPECTRUM', 'SIGMA'))
        if spectrum in cols:
            w = np.asarray(cols[spectrum])
        elif row in cols:
            w = np.asarray(cols[row])[:, None, :] * np.ones((cpb, 1))
        else:
            w = np.ones((ntime * nbl * nbands, cpb, 4))
        out['weight_vector'] = w.reshape(ntime, nbl, nchan, 4).astype(ft)
    return out


class MeasurementSetLoader(object):
    """The reference's loader interface (montblanc/impl/rime/v4/loaders/
    loaders.py:49-253): ``get_dims()`` and ``load(solver, slvr_cfg)``,
    which transfers the arrays into a v4 solver."""

    def __init__(self, msfile, auto_correlations=False):
        self.msfile = msfile
        self.cols = read_ms(msfile, auto_correlations)
        self.ntime, self.nbl, self.na, self.nbands, self.nchan = ms_dims(self.cols)

    def get_dims(self):
        return self.ntime, self.nbl, self.na, self.nbands, self.nchan

    def load(self, solver, slvr_cfg):
        ft = getattr(solver, 'ft', np.float64)
        arrays = ms_arrays(self.cols, ft, slvr_cfg.get('init_weights'))
        for name, value in arrays.items():
            getattr(solver, 'transfer_' + name)(np.ascontiguousarray(value))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False
