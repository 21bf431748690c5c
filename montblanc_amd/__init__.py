"""montblanc_amd: the v4 RIME hot path of Montblanc on MI355X (gfx950).

Public surface mirrors the reference package (montblanc/__init__.py:52-147):
``rime_solver_cfg(**kw)``, ``rime_solver(cfg)``, ``sources(**kw)``,
``source_types()``, ``constants.C`` and the ``montblanc`` logger, plus the
kernel-node classes and ``Pipeline``. The arithmetic runs in the HIP library
``libmontblanc_amd.so`` built from ``montblanc_amd/csrc``; nothing falls back
to the CPU.
"""

import logging

from . import src_types
from .config import RimeSolverConfig, describe_options
from .pipeline import Node, NullNode, Pipeline, PipeLineError

log = logging.getLogger('montblanc')
if not log.handlers:
    log.addHandler(logging.NullHandler())


class _Constants(object):
    @property
    def C(self):
        """Speed of light in m/s (montblanc/__init__.py:52-59)."""
        return 299792458

    def __setattr__(self, name, value):
        raise SyntaxError('constants are read-only')


constants = _Constants()


def source_types():
    return src_types.source_types()


def sources(**kwargs):
    """Per-type source counts, e.g. sources(point=10, gaussian=20)."""
    return src_types.default_sources(**kwargs)


def rime_solver_cfg(**kwargs):
    """Solver configuration dict (montblanc/__init__.py:96-128).

    Valid options and their defaults are listed by describe_options().
    """
    cfg = RimeSolverConfig().gen_cfg(**kwargs)
    cfg.setdefault(RimeSolverConfig.DATA_SOURCE, RimeSolverConfig.DATA_SOURCE_MS)
    return cfg


def rime_solver(slvr_cfg):
    """A RIME solver for ``slvr_cfg`` (montblanc/factory.py:150-183).

    version 'v4' -> the HIP v4 solver; 'v5' -> the same solver driven by the
    chunk planner over time and source batches; 'v2' is not provided.
    """
    from .solver import RimeSolver
    cfg = dict(slvr_cfg)
    version = cfg.get(RimeSolverConfig.VERSION, RimeSolverConfig.DEFAULT_VERSION)
    if version == RimeSolverConfig.VERSION_FOUR:
        cls = RimeSolver
    elif version == RimeSolverConfig.VERSION_FIVE:
        from .planner import ChunkedRimeSolver
        cls = ChunkedRimeSolver
    elif version == RimeSolverConfig.VERSION_TWO:
        raise ValueError('RIME version v2 is not provided by montblanc_amd (v4 only)')
    else:
        raise ValueError('Invalid version %s' % version)
    data_source = cfg.get(RimeSolverConfig.DATA_SOURCE, RimeSolverConfig.DATA_SOURCE_MS)
    if data_source == RimeSolverConfig.DATA_SOURCE_MS:
        return create_rime_solver_from_ms(cls, cfg)
    if data_source not in (RimeSolverConfig.DATA_SOURCE_TEST, RimeSolverConfig.DATA_SOURCE_DEFAULT,
                           RimeSolverConfig.DATA_SOURCE_EMPTY):
        raise ValueError('Invalid data source %s' % data_source)
    return cls(cfg)


def create_rime_solver_from_ms(slvr_class_type, slvr_cfg):
    """A solver of ``slvr_class_type`` filled from the MeasurementSet named by
    slvr_cfg['msfile'] (montblanc/factory.py:113-148): KeyError without one;
    the dimensions (ntime, na, nbl, nbands, nchan) come from the loader
    into the configuration, the solver is built on them and the loader
    transfers the arrays into it. The tables are opened by
    ``ms_loader.open_tables`` (python-casacore)."""
    from . import ms_loader
    cfg = dict(slvr_cfg)
    O = RimeSolverConfig
    if cfg.get(O.MS_FILE) is None:
        raise KeyError('%s key is set to %s in the Solver Configuration, but no MeasurementSet '
                       'file has been specified in the %s key'
                       % (O.DATA_SOURCE, O.DATA_SOURCE_MS, O.MS_FILE))
    with ms_loader.MeasurementSetLoader(cfg[O.MS_FILE],
                                        auto_correlations=cfg.get(O.AUTO_CORRELATIONS, False)) as ld:
        ntime, nbl, na, nbands, nchan = ld.get_dims()
        cfg[O.NTIME], cfg[O.NA], cfg[O.NBL] = ntime, na, nbl
        cfg[O.NBANDS], cfg[O.NCHAN] = nbands, nchan
        slvr = slvr_class_type(cfg)
        ld.load(slvr, cfg)
        return slvr


def __getattr__(name):
    # Node classes import torch; load them on first use.
    if name in ('RimeBSqrt', 'RimeEBeam', 'RimeEKBSqrt', 'RimeSumCoherencies',
                'RimeReduction', 'RimeFusedForward', 'default_pipeline', 'staged_pipeline'):
        from . import nodes
        return getattr(nodes, name)
    raise AttributeError(name)
