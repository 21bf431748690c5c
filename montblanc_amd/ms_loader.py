"""MeasurementSet input for the v4 solver: the MS columns turned into the
solver's arrays.

Reference: montblanc/impl/common/loaders/loaders.py:35-137 (the ordered
view of the main table and the dimensions),
montblanc/impl/rime/v4/loaders/loaders.py:49-253 (the arrays) and
montblanc/impl/rime/v5/loaders/loaders.py:160-320 (the main table read in
row increments of whole time steps). The main table's rows are taken
ordered by (TIME, ANTENNA1, ANTENNA2, band), field 0, auto-correlations
left out unless asked for; then

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
  or the same from SIGMA_SPECTRUM / SIGMA ('sigma');
* parallactic_angles from the ANTENNA POSITIONs, the FIELD's PHASE_DIR
  (a negative right ascension taken modulo 2 pi) and the unique TIMEs
  (montblanc_amd/parallactic.py), when those are present.

The main table is read ``row_increment`` rows at a time (whole time steps,
at least 5000 rows, as the reference's v5 loader does) straight into the
host arrays, so no column is ever held twice.

The tables come from ``open_tables(msfile, auto_correlations)``: the
ordered python-casacore view (``CasacoreTables``), which reports clearly
when python-casacore is absent. ``ColumnTables`` holds the same columns in
memory (read by other means); ``ms_arrays(cols)`` and
``MeasurementSetLoader(msfile, tables=ColumnTables(cols))`` take those.
"""

import logging

import numpy as np

log = logging.getLogger('montblanc')

MAIN_COLUMNS = ('TIME', 'UVW', 'ANTENNA1', 'ANTENNA2', 'DATA', 'FLAG', 'FLAG_ROW', 'WEIGHT',
                'WEIGHT_SPECTRUM', 'SIGMA', 'SIGMA_SPECTRUM')
SUB_VALUES = ('CHAN_FREQ', 'REF_FREQUENCY', 'NUM_CHAN', 'NA', 'POSITION', 'PHASE_DIR')
LOG_PREFIX = 'LOADER:'


def _tables():
    try:
        import casacore.tables as pt
    except ImportError:
        try:
            import pyrap.tables as pt
        except ImportError:
            raise ImportError('reading a MeasurementSet needs python-casacore (casacore.tables), '
                              'which is not installed; ms_arrays() and ColumnTables take the '
                              'columns read by other means')
    return pt


class ColumnTables(object):
    """The loader's tables held as in-memory columns: the ordered main table
    (MAIN_COLUMNS, one row per time, baseline and band) plus the sub-table
    values CHAN_FREQ (nbands, cpb), REF_FREQUENCY, NUM_CHAN (nbands,), NA
    and optionally POSITION (na, 3) and PHASE_DIR (2,)."""

    def __init__(self, cols):
        self.cols = cols

    def nrows(self):
        return len(self.cols['TIME'])

    def colnames(self):
        return [c for c in MAIN_COLUMNS if c in self.cols]

    def getcol(self, name, startrow=0, nrow=-1):
        c = np.asarray(self.cols[name])
        return c[startrow:] if nrow < 0 else c[startrow:startrow + nrow]

    def sub(self, name):
        return self.cols.get(name)

    def close(self):
        pass


class CasacoreTables(object):
    """The ordered view of a MeasurementSet's main table and its ANTENNA,
    SPECTRAL_WINDOW and FIELD sub-tables, through python-casacore
    (montblanc/impl/common/loaders/loaders.py:47-79)."""

    def __init__(self, msfile, auto_correlations=False):
        pt = _tables()
        self._ms = pt.table(msfile, ack=False)
        query = ' '.join(['SELECT FROM $ms WHERE FIELD_ID=0',
                          '' if auto_correlations else 'AND ANTENNA1 != ANTENNA2',
                          'ORDERBY TIME, ANTENNA1, ANTENNA2, '
                          '[SELECT SPECTRAL_WINDOW_ID FROM ::DATA_DESCRIPTION][DATA_DESC_ID]'])
        ms = self._ms  # noqa: F841  (named in the query)
        self._main = pt.taql(query)
        spw = pt.table('::'.join((msfile, 'SPECTRAL_WINDOW')), ack=False, readonly=True)
        ant = pt.table('::'.join((msfile, 'ANTENNA')), ack=False, readonly=True)
        field = pt.table('::'.join((msfile, 'FIELD')), ack=False, readonly=True)
        self._sub = dict(CHAN_FREQ=spw.getcol('CHAN_FREQ'),
                         REF_FREQUENCY=spw.getcol('REF_FREQUENCY'),
                         NUM_CHAN=spw.getcol('NUM_CHAN'), NA=ant.nrows(),
                         POSITION=ant.getcol('POSITION'),
                         PHASE_DIR=field.getcol('PHASE_DIR')[0][0])
        for t in (spw, ant, field):
            t.close()

    def nrows(self):
        return self._main.nrows()

    def colnames(self):
        return [c for c in MAIN_COLUMNS if c in self._main.colnames()]

    def getcol(self, name, startrow=0, nrow=-1):
        return self._main.getcol(name, startrow=startrow, nrow=nrow)

    def sub(self, name):
        return self._sub.get(name)

    def close(self):
        self._main.close()
        self._ms.close()


def open_tables(msfile, auto_correlations=False):
    """The tables a MeasurementSetLoader reads (python-casacore's view)."""
    return CasacoreTables(msfile, auto_correlations)


def read_ms(msfile, auto_correlations=False):
    """Every column ms_arrays uses, read whole with python-casacore."""
    t = open_tables(msfile, auto_correlations)
    try:
        cols = {c: t.getcol(c) for c in t.colnames()}
        cols.update({k: t.sub(k) for k in SUB_VALUES})
    finally:
        t.close()
    return cols


def table_dims(tables):
    """(ntime, nbl, na, nbands, nchan) of the ordered rows; the bands must
    have equal channel counts and the rows fill ntime x nbl x nbands."""
    num_chan = np.asarray(tables.sub('NUM_CHAN'))
    if not np.all(num_chan == num_chan[0]):
        raise ValueError('channels per band %s are not equal' % list(num_chan))
    ntime = len(np.unique(tables.getcol('TIME')))
    pairs = np.stack([np.asarray(tables.getcol('ANTENNA1')),
                      np.asarray(tables.getcol('ANTENNA2'))], 1)
    nbl = len(np.unique(pairs, axis=0))
    nbands, nchan = len(num_chan), int(num_chan.sum())
    rows = tables.nrows()
    if rows != ntime * nbl * nbands:
        raise ValueError('%d rows are not ntime x nbl x nbands = %d x %d x %d'
                         % (rows, ntime, nbl, nbands))
    return ntime, nbl, int(tables.sub('NA')), nbands, nchan


def ms_dims(cols):
    """table_dims of in-memory columns."""
    return table_dims(ColumnTables(cols))


def row_increment(nbl, nbands, min_rows=5000):
    """Rows per read: whole time steps, doubled until at least min_rows
    (montblanc/impl/rime/v5/loaders/loaders.py:163-169)."""
    time_inc = 1
    while time_inc * nbl * nbands < min_rows:
        time_inc *= 2
    return time_inc * nbl * nbands


def load_arrays(tables, ft=np.float64, init_weights=None, min_rows=5000):
    """Solver arrays (v4 names) read from ``tables`` in row increments."""
    ntime, nbl, na, nbands, nchan = table_dims(tables)
    ct = np.complex64 if np.dtype(ft) == np.float32 else np.complex128
    cpb = nchan // nbands
    names = tables.colnames()
    nrows = tables.nrows()
    out = {}
    out['frequency'] = np.asarray(tables.sub('CHAN_FREQ'), ft).reshape(nchan)
    out['ref_frequency'] = np.repeat(np.asarray(tables.sub('REF_FREQUENCY'), ft),
                                     np.asarray(tables.sub('NUM_CHAN')))
    if init_weights not in (None, 'weight', 'sigma'):
        raise ValueError("init_weights must be None, 'weight' or 'sigma' (got %r)" % (init_weights,))
    wcol = None
    if init_weights is not None:
        spectrum, row = (('WEIGHT_SPECTRUM', 'WEIGHT') if init_weights == 'weight'
                         else ('SIGMA_SPECTRUM', 'SIGMA'))
        wcol = spectrum if spectrum in names else row if row in names else 'ones'
        log.info("%s '%s' will be loaded into 'weight_vector'." % (LOG_PREFIX, wcol))
    # row views (time * baseline * band, channel of the band, correlation)
    vis = np.zeros((ntime, nbl, nchan, 4), ct) if 'DATA' in names else None
    flag = np.zeros((ntime, nbl, nchan, 4), np.uint8) if 'FLAG' in names else None
    wv = np.ones((ntime, nbl, nchan, 4), ft) if wcol is not None else None
    a1 = np.empty(ntime * nbl, np.int32)
    a2 = np.empty(ntime * nbl, np.int32)
    uvw_bl = np.empty((ntime * nbl, 3), np.float64)
    inc = row_increment(nbl, nbands, min_rows)
    for start in range(0, nrows, inc):
        n = min(inc, nrows - start)
        sl = slice(start, start + n)
        if vis is not None:
            vis.reshape(-1, cpb, 4)[sl] = tables.getcol('DATA', start, n)
        if flag is not None:
            f = np.asarray(tables.getcol('FLAG', start, n), bool)
            if 'FLAG_ROW' in names:
                f = f | np.asarray(tables.getcol('FLAG_ROW', start, n), bool)[:, None, None]
            flag.reshape(-1, cpb, 4)[sl] = f
        if wcol in ('WEIGHT_SPECTRUM', 'SIGMA_SPECTRUM'):
            wv.reshape(-1, cpb, 4)[sl] = tables.getcol(wcol, start, n)
        elif wcol in ('WEIGHT', 'SIGMA'):
            wv.reshape(-1, cpb, 4)[sl] = np.asarray(tables.getcol(wcol, start, n))[:, None, :]
        # one row per (time, baseline): the first band's (rows are band-minor)
        bl = slice(start // nbands, (start + n) // nbands)
        a1[bl] = np.asarray(tables.getcol('ANTENNA1', start, n))[::nbands]
        a2[bl] = np.asarray(tables.getcol('ANTENNA2', start, n))[::nbands]
        uvw_bl[bl] = np.asarray(tables.getcol('UVW', start, n))[::nbands]
    a1 = a1.reshape(ntime, nbl)
    a2 = a2.reshape(ntime, nbl)
    uvw_bl = uvw_bl.reshape(ntime, nbl, 3)
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
    pos, pdir = tables.sub('POSITION'), tables.sub('PHASE_DIR')
    if pos is not None and pdir is not None:
        from .parallactic import parallactic_angles
        pdir = np.array(pdir, np.float64)
        if pdir[0] < 0:
            pdir[0] += 2 * np.pi
        times = np.unique(np.asarray(tables.getcol('TIME'), np.float64))
        out['parallactic_angles'] = parallactic_angles(pdir, pos, times).astype(ft)
    if vis is not None:
        out['observed_vis'] = vis
    if flag is not None:
        out['flag'] = flag
    if wv is not None:
        out['weight_vector'] = wv
    return out


def ms_arrays(cols, ft=np.float64, init_weights=None, min_rows=5000):
    """Solver arrays (v4 names) from in-memory MS columns."""
    return load_arrays(ColumnTables(cols), ft, init_weights, min_rows)


class MeasurementSetLoader(object):
    """The reference's loader interface (montblanc/impl/rime/v4/loaders/
    loaders.py:49-253): ``get_dims()`` and ``load(solver, slvr_cfg)``,
    which transfers the arrays into a v4 solver; a context manager that
    closes the tables. ``tables`` replaces the python-casacore view (a
    ColumnTables, or any object with its methods)."""

    def __init__(self, msfile, auto_correlations=False, tables=None):
        self.msfile = msfile
        self.tables = tables if tables is not None else open_tables(msfile, auto_correlations)
        self.ntime, self.nbl, self.na, self.nbands, self.nchan = table_dims(self.tables)
        log.info('%s Found %d antenna(s), %d baseline(s), %d band(s) of %d channels, %d time(s).'
                 % (LOG_PREFIX, self.na, self.nbl, self.nbands, self.nchan // self.nbands,
                    self.ntime))

    def get_dims(self):
        return self.ntime, self.nbl, self.na, self.nbands, self.nchan

    def load(self, solver, slvr_cfg):
        ft = getattr(solver, 'ft', np.float64)
        arrays = load_arrays(self.tables, ft, slvr_cfg.get('init_weights'))
        for name, value in arrays.items():
            getattr(solver, 'transfer_' + name)(np.ascontiguousarray(value))

    def close(self):
        self.tables.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False
