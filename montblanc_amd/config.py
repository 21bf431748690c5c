"""Solver configuration: the reference's option keys, defaults and rules.

Mirrors montblanc/slvr_config.py:61-366 and
montblanc/impl/rime/slvr_config.py:35-276: ``gen_cfg(**kwargs)`` merges an
optional ``cfg_file`` ([montblanc] ini section), then the keyword arguments,
over the defaults; string values are converted with the option's type (as
argparse does for string defaults); keys that are not options pass through
unchanged (e.g. ``pipeline``, ``mem_budget``, ``context``).
"""

import argparse
import configparser
import json
import logging
import re

from . import src_types

log = logging.getLogger('montblanc')


class ArgumentParserError(Exception):
    pass


class _Parser(argparse.ArgumentParser):
    _ARG_RE = re.compile(r'^(?:argument\s*?(\S*?)):\s*?(.*?)$')

    def error(self, message):
        m = self._ARG_RE.match(message)
        if m and m.group(1).strip('-'):
            raise ArgumentParserError("argument '%s': %s" % (m.group(1).strip('-'), m.group(2)))
        raise ArgumentParserError(message)


def _bool(v):
    if isinstance(v, bool):
        return v
    s = str(v).strip().lower()
    if s in ('true', '1', 'yes'):
        return True
    if s in ('false', '0', 'no', ''):
        return False
    raise argparse.ArgumentTypeError('%r is not a boolean' % v)


def _sources(v):
    return json.loads(v) if isinstance(v, str) else v


def _none_or_str(v):
    return None if v in (None, 'None') else str(v)


class SolverConfig(object):
    CFG_FILE = 'cfg_file'

    MODE = 'mode'
    MODE_CHI_SQUARED = 'chi-squared'
    MODE_SIMULATOR = 'simulator'
    SOURCES = 'sources'
    NTIME = 'ntime'
    NA = 'na'
    NBL = 'nbl'
    NCHAN = 'nchan'
    NBANDS = 'nbands'
    NPOL = 'npol'
    NSRC = 'nsrc'
    DTYPE = 'dtype'
    DTYPE_FLOAT = 'float'
    DTYPE_DOUBLE = 'double'
    AUTO_CORRELATIONS = 'auto_correlations'
    DATA_SOURCE = 'data_source'
    DATA_SOURCE_DEFAULT = 'default'
    DATA_SOURCE_MS = 'ms'
    DATA_SOURCE_TEST = 'test'
    DATA_SOURCE_EMPTY = 'empty'
    MS_FILE = 'msfile'
    DATA_ORDER = 'data_order'
    CONTEXT = 'context'

    # name: (type, default, choices, description)
    OPTIONS = {
        MODE: (str, MODE_CHI_SQUARED, None, "Execution mode: 'chi-squared' or 'simulator'"),
        SOURCES: (_sources, json.dumps(dict(src_types.default_sources())), None,
                  'Dictionary of source type numbers e.g. {"point": 1}'),
        NTIME: (int, 10, None, 'Timesteps'),
        NA: (int, 7, None, 'Antenna'),
        NPOL: (int, 4, None, 'Polarisations'),
        NCHAN: (int, 16, None, 'Channels'),
        NBANDS: (int, 1, None, 'Bands'),
        DTYPE: (str, DTYPE_DOUBLE, [DTYPE_FLOAT, DTYPE_DOUBLE], 'Floating point precision'),
        AUTO_CORRELATIONS: (_bool, False, [True, False], 'Include auto-correlations'),
        DATA_SOURCE: (str, DATA_SOURCE_MS,
                      [DATA_SOURCE_DEFAULT, DATA_SOURCE_MS, DATA_SOURCE_TEST, DATA_SOURCE_EMPTY],
                      'Data source for initialising arrays'),
        MS_FILE: (_none_or_str, None, None, 'MeasurementSet file'),
        DATA_ORDER: (str, 'casa', ['casa', 'other'], 'MeasurementSet data ordering'),
        CONTEXT: (None, None, None, 'Device context (a torch device or index)'),
    }

    def parser(self):
        p = _Parser(add_help=False)
        for name, (typ, default, choices, desc) in self.all_options().items():
            kw = dict(required=False, help=desc, default=default)
            if typ is not None:
                kw['type'] = typ
            if choices is not None:
                kw['choices'] = choices
            p.add_argument('--' + name, **kw)
        return p

    @classmethod
    def all_options(cls):
        opts = {}
        for klass in reversed(cls.__mro__):
            opts.update(getattr(klass, 'OPTIONS', {}))
        return opts

    def gen_cfg(self, **kwargs):
        """Defaults <- cfg_file [montblanc] section <- kwargs (slvr_config.py:345-366)."""
        p = self.parser()
        if self.CFG_FILE in kwargs:
            log.info("Loading defaults from '%s'.", kwargs[self.CFG_FILE])
            cp = configparser.ConfigParser()
            cp.read([kwargs[self.CFG_FILE]])
            if cp.has_section('montblanc'):
                p.set_defaults(**dict(cp.items('montblanc')))
        p.set_defaults(**kwargs)
        return vars(p.parse_args([]))


class RimeSolverConfig(SolverConfig):
    E_BEAM_WIDTH = 'beam_lw'
    E_BEAM_HEIGHT = 'beam_mh'
    E_BEAM_DEPTH = 'beam_nud'
    WEIGHT_VECTOR = 'weight_vector'
    INIT_WEIGHTS = 'init_weights'
    SOURCE_BATCH_SIZE = 'source_batch_size'
    NSOLVERS = 'nsolvers'
    VISIBILITY_THROTTLE_FACTOR = 'visibility_throttle_factor'
    VISIBILITY_OUTPUT = 'vis_output'
    VISIBILITY_OUTPUT_MODEL = 'model'
    VISIBILITY_OUTPUT_RESIDUALS = 'residuals'
    VISIBILITY_WRITE_MODE = 'vis_write_mode'
    VISIBILITY_WRITE_MODE_OVERWRITE = 'overwrite'
    VISIBILITY_WRITE_MODE_SUM = 'sum'
    VERSION = 'version'
    VERSION_TWO = 'v2'
    VERSION_FOUR = 'v4'
    VERSION_FIVE = 'v5'
    DEFAULT_VERSION = VERSION_FOUR
    MEM_BUDGET = 'mem_budget'

    OPTIONS = {
        E_BEAM_WIDTH: (int, 50, None, 'Width of the E beam cube (l)'),
        E_BEAM_HEIGHT: (int, 50, None, 'Height of the E beam cube (m)'),
        E_BEAM_DEPTH: (int, 50, None, 'Depth of the E beam cube (frequency)'),
        WEIGHT_VECTOR: (_bool, False, [True, False], 'Weight chi-squared terms per visibility'),
        INIT_WEIGHTS: (_none_or_str, None, [None, 'sigma', 'weight'],
                       'How a MeasurementSet initialises the weight vector'),
        SOURCE_BATCH_SIZE: (int, 500, None, 'Minimum source batch size'),
        NSOLVERS: (int, 2, None, 'Concurrent solvers per device'),
        VISIBILITY_THROTTLE_FACTOR: (int, 6, None, 'Visibility chunks enqueued before throttling'),
        VISIBILITY_OUTPUT: (str, VISIBILITY_OUTPUT_MODEL,
                            [VISIBILITY_OUTPUT_MODEL, VISIBILITY_OUTPUT_RESIDUALS],
                            'Model visibilities or residuals'),
        VISIBILITY_WRITE_MODE: (str, VISIBILITY_WRITE_MODE_OVERWRITE,
                                [VISIBILITY_WRITE_MODE_SUM, VISIBILITY_WRITE_MODE_OVERWRITE],
                                'Overwrite or accumulate model visibilities'),
        VERSION: (str, DEFAULT_VERSION, [VERSION_TWO, VERSION_FOUR, VERSION_FIVE], 'RIME version'),
    }


def describe_options():
    """One line per option: name, default, description (config.py:87-95)."""
    lines = []
    for name, (_, default, choices, desc) in sorted(RimeSolverConfig.all_options().items()):
        extra = '' if choices is None else ' Valid: %s.' % (choices,)
        lines.append("'%s': %s Default: %r.%s" % (name, desc, default, extra))
    return '\n'.join(lines)
