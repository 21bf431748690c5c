"""Source types and source-range bookkeeping.

Same public behaviour as montblanc/src_types.py:40-189: a fixed type order
point -> gaussian -> sersic inside the global source index, and helpers that
map a global source range [start, end) to per-type ranges. The chunk planner
uses these to split source batches.
"""

from collections import OrderedDict
import logging

log = logging.getLogger('montblanc')

POINT_TYPE, POINT_NR_VAR = 'point', 'npsrc'
GAUSSIAN_TYPE, GAUSSIAN_NR_VAR = 'gaussian', 'ngsrc'
SERSIC_TYPE, SERSIC_NR_VAR = 'sersic', 'nssrc'

# src_types.py:40-44
SOURCE_VAR_TYPES = OrderedDict([
    (POINT_TYPE, POINT_NR_VAR),
    (GAUSSIAN_TYPE, GAUSSIAN_NR_VAR),
    (SERSIC_TYPE, SERSIC_NR_VAR),
])


def source_types():
    return list(SOURCE_VAR_TYPES.keys())


def source_nr_vars():
    return list(SOURCE_VAR_TYPES.values())


def default_sources(**kwargs):
    """Counts for every registered type; one point source if all are zero.

    src_types.py:54-99 (unknown types are logged and ignored, values that do
    not convert to int raise TypeError).
    """
    for t in kwargs:
        if t not in SOURCE_VAR_TYPES:
            log.warning('Source type %s is not yet implemented in montblanc. '
                        'Valid source types are %s', t, source_types())
    out = OrderedDict()
    for name in SOURCE_VAR_TYPES:
        value = kwargs.get(name, 0)
        try:
            out[name] = int(value)
        except (TypeError, ValueError):
            raise TypeError('Supplied value %s for source %s cannot be '
                            'converted to an integer' % (value, name))
    if sum(out.values()) == 0:
        out[POINT_TYPE] = 1
    return out


def sources_to_nr_vars(sources):
    """{'point': 10} -> OrderedDict(npsrc=10, ngsrc=0, nssrc=0). src_types.py:101-125."""
    counts = default_sources(**sources)
    return OrderedDict((SOURCE_VAR_TYPES[k], v) for k, v in counts.items())


def source_range_tuple(start, end, nr_var_dict):
    """Per-type (start, end) local index pairs covering global [start, end).

    Same answers as src_types.py:127-167 (pinned by the known-answer cases of
    montblanc/tests/test_source_utils.py:46-106); written as an interval
    intersection of [start, end) with each type's global interval.
    """
    out = OrderedDict()
    offset = 0
    for nr_var in SOURCE_VAR_TYPES.values():
        n = int(nr_var_dict.get(nr_var, 0))
        lo = max(start, offset)
        hi = min(end, offset + n)
        if hi > lo:
            out[nr_var] = (lo - offset, hi - offset)
        else:
            out[nr_var] = (0, 0)
        offset += n
    return out


def source_range(start, end, nr_var_dict):
    """Per-type counts inside global [start, end). src_types.py:169-179."""
    return OrderedDict((k, e - s) for k, (s, e)
                       in source_range_tuple(start, end, nr_var_dict).items())


def source_range_slices(start, end, nr_var_dict):
    """Per-type slices inside global [start, end). src_types.py:181-189."""
    return OrderedDict((k, slice(s, e, 1)) for k, (s, e)
                       in source_range_tuple(start, end, nr_var_dict).items())
