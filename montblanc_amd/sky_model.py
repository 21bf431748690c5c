"""Sky-model catalog files: the reference's text format, read into the
source arrays of the solver and of a batched-evaluation sample.

Format (util/sky_model_parser.py:1-67, examples/sky_model.txt:1-6): comma
separated lines. A line whose first field reads ``# format <count>:
<name>`` opens a section: its sources count towards ``<count>`` (npsrc,
ngsrc or nssrc) and its columns are named by ``<name>`` and the line's
further fields. Every other line is one source; each value is appended to
the column of its position's name, so a column collects its values over
all sections in file order (``l`` holds every source's l, ``el`` only the
gaussians').

``parse_sky_model`` returns the reference's result object (column name ->
list of the value strings, count name -> number of sources, and
``shape_arrays`` with the reference's reshape semantics). Two departures,
both for lines the reference cannot read: blank lines are skipped (the
reference indexes their first field and fails) and ``#`` lines that are
not format lines are comments (the reference would take them as a source).

``source_arrays`` turns the result into lm (n, 2), stokes (n, ntime, 4),
alpha (n, ntime), gauss_shape (3, ng) = (el, em, eR) and sersic_shape
(3, ns) = (e1, e2, eS), with the sources in the solver's order (point,
gaussian, sersic; file order within a type) and the brightness repeated
over time as examples/MS_sky_model_example.py:29-64 does. The dict also
carries ``counts`` = (npsrc, ngsrc, nssrc): it is one sample of
``batched.Batch`` as it stands, and its arrays go to a solver through
``transfer_lm`` and the like.
"""

import csv
import re

import numpy as np

_FORMAT = re.compile(r'^\s*#\s*format\s+(?P<count>\S*?)\s*:\s*(?P<name>\S*?)\s*$')

# per-source columns and their value when a file leaves them out
_BRIGHTNESS = (('I', None), ('Q', 0.0), ('U', 0.0), ('V', 0.0))
_SHAPE_COLUMNS = {'ngsrc': ('el', 'em', 'eR'), 'nssrc': ('e1', 'e2', 'eS')}
_TYPE_ORDER = ('npsrc', 'ngsrc', 'nssrc')


class ParseResults(object):
    """Columns (name -> list of value strings, file order) and source
    counts (count name -> sources) of a sky-model file; ``rows`` keeps each
    source's count name and values by column, in file order."""

    def __init__(self):
        self.arrays = {}
        self.src_counts = {}
        self.rows = []

    def shape_arrays(self, fmt_list, shape, dtype):
        """The named columns as one array of ``shape`` and ``dtype``: the
        columns stacked in ``fmt_list`` order, then reshaped (not
        transposed), as the reference does (util/__init__.py:340-342)."""
        return shape_list([self.arrays[a] for a in fmt_list], shape, dtype)


def shape_list(lists, shape, dtype):
    """A list of lists as an array of ``shape`` and ``dtype`` (values
    converted by numpy, then reshaped)."""
    return np.array(lists, dtype=dtype).reshape(shape)


def parse_sky_model(filename):
    """Read a sky-model file (see the module docstring)."""
    res = ParseResults()
    names, count = [], None
    with open(filename, 'r', newline='') as f:
        for row in csv.reader(f):
            if not row:
                continue
            first = row[0]
            if first.lstrip().startswith('#'):
                m = _FORMAT.match(first)
                if m:
                    count = m.group('count')
                    names = [m.group('name')] + [v.strip() for v in row[1:]]
                continue
            values = {}
            for name, v in zip(names, row):
                v = v.strip()
                res.arrays.setdefault(name, []).append(v)
                values[name] = v
            if count is not None:
                res.src_counts[count] = res.src_counts.get(count, 0) + 1
                res.rows.append((count, values))
    return res


def source_arrays(res, ntime, dtype=np.float64):
    """Solver / batched-sample source arrays of a parse result (see the
    module docstring). A source without I, l or m, or a gaussian or sersic
    source without its shape columns, is an error."""
    unknown = sorted({c for c, _ in res.rows} - set(_TYPE_ORDER))
    if unknown:
        raise ValueError('sky model: unknown source counts %s (expected %s)' % (unknown, _TYPE_ORDER))
    ordered = [r for t in _TYPE_ORDER for r in res.rows if r[0] == t]
    n = len(ordered)

    def value(i, values, name, default=None):
        v = values.get(name)
        if v is None or v == '':
            if default is None:
                raise ValueError('sky model: source %d has no %s value' % (i, name))
            return default
        return float(v)

    lm = np.zeros((n, 2), dtype)
    iquv = np.zeros((n, 4), dtype)
    alpha = np.zeros(n, dtype)
    shapes = {t: [] for t in _SHAPE_COLUMNS}
    for i, (t, values) in enumerate(ordered):
        lm[i] = (value(i, values, 'l'), value(i, values, 'm'))
        iquv[i] = [value(i, values, name, d) for name, d in _BRIGHTNESS]
        alpha[i] = value(i, values, 'alpha', 0.0)
        if t in _SHAPE_COLUMNS:
            shapes[t].append([value(i, values, name) for name in _SHAPE_COLUMNS[t]])
    counts = tuple(sum(1 for r in ordered if r[0] == t) for t in _TYPE_ORDER)
    return {
        'lm': lm,
        'stokes': np.repeat(iquv[:, None, :], ntime, axis=1),
        'alpha': np.repeat(alpha[:, None], ntime, axis=1),
        'gauss_shape': np.array(shapes['ngsrc'], dtype).reshape(-1, 3).T.copy(),
        'sersic_shape': np.array(shapes['nssrc'], dtype).reshape(-1, 3).T.copy(),
        'counts': counts,
    }
