"""The v4 RIME solver on a HIP device.

Drop-in for montblanc/impl/rime/v4/RimeSolver.py:42-88 on top of
montblanc/solvers/rime_solver.py:35-369 and mb_cuda_solver.py:35-191:

* dimensions with global / local sizes and lower / upper extents (the
  semantics of docs/dims.rst), ``dim_local_size`` & co.;
* the v4 arrays, allocated on the device as torch tensors, initialised
  from the 'default' / 'test' / 'empty' data sources, with
  ``transfer_<name>(ndarray)`` and ``retrieve_<name>()`` per array;
* the v4 properties with ``set_<name>(value)``; ``X2`` is filled by the
  chi-squared reduction of the last solve;
* ``solve()`` / ``initialise()`` / ``shutdown()``, context-manager use, and
  a ``pipeline`` of kernel nodes taken from ``slvr_cfg['pipeline']``.

Device memory, streams and the tensor boundary are torch's; the arithmetic
is the library's (montblanc_amd/_lib.py).
"""

import logging
from collections import OrderedDict

import numpy as np
import torch

from . import _lib
from . import schema
from . import src_types
from .config import RimeSolverConfig as Options

log = logging.getLogger('montblanc')

_TORCH = {np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64,
          np.dtype(np.complex64): torch.complex64, np.dtype(np.complex128): torch.complex128,
          np.dtype(np.int32): torch.int32, np.dtype(np.uint8): torch.uint8}


def nr_of_baselines(na, auto_correlations=False):
    """montblanc/util/__init__.py:42-50."""
    m = (na - 1) if auto_correlations is False else (na + 1)
    return (na * m) // 2


def nr_of_antenna(nbl, auto_correlations=False):
    """montblanc/util/__init__.py:52-60."""
    t = 1 if auto_correlations is False else -1
    return int(t + np.sqrt(1 + 8 * nbl)) // 2


class Dimension(object):
    __slots__ = ('name', 'global_size', 'local_size', 'lower_extent', 'upper_extent',
                 'description', 'zero_valid')

    def __init__(self, name, size, description='', zero_valid=False):
        self.name = name
        self.global_size = self.local_size = self.upper_extent = int(size)
        self.lower_extent = 0
        self.description = description
        self.zero_valid = zero_valid


def resolve_device(context):
    if context is None:
        return torch.device('cuda', 0)
    if isinstance(context, torch.device):
        return context
    if isinstance(context, int):
        return torch.device('cuda', context)
    if isinstance(context, str):
        return torch.device(context)
    if isinstance(context, (list, tuple)) and context:
        return resolve_device(context[0])
    raise TypeError('context must be a torch.device, device index or device string')


class RimeSolver(object):
    """v4 RIME solver whose kernels are the gfx950 library's."""

    def __init__(self, slvr_cfg):
        self._slvr_cfg = slvr_cfg
        if slvr_cfg[Options.DTYPE] == Options.DTYPE_FLOAT:
            self.ft, self.ct = np.float32, np.complex64
        elif slvr_cfg[Options.DTYPE] == Options.DTYPE_DOUBLE:
            self.ft, self.ct = np.float64, np.complex128
        else:
            raise TypeError('Invalid dtype %s ' % slvr_cfg[Options.DTYPE])
        self._use_weight_vector = slvr_cfg.get(Options.WEIGHT_VECTOR)
        self._is_auto_correlated = slvr_cfg.get(Options.AUTO_CORRELATIONS)
        self._visibility_output = slvr_cfg.get(Options.VISIBILITY_OUTPUT)
        self.device = resolve_device(slvr_cfg.get(Options.CONTEXT))
        if self.device.type != 'cuda':
            raise ValueError('RimeSolver needs a HIP device, got %s' % self.device)
        _lib.load()

        from .nodes import default_pipeline
        self.pipeline = slvr_cfg.get('pipeline') or default_pipeline()

        self._dims = OrderedDict()
        self._arrays = OrderedDict()
        self._props = OrderedDict()
        self._tensors = {}
        self._workspaces = {}
        self._x2_dev = None
        self._x2_host = None
        self._pair_plan = None
        self._host_antenna = [None, None]

        self.register_default_dimensions()
        self.register_dimension('beam_lw', slvr_cfg[Options.E_BEAM_WIDTH], 'E cube l width')
        self.register_dimension('beam_mh', slvr_cfg[Options.E_BEAM_HEIGHT], 'E cube m height')
        self.register_dimension('beam_nud', slvr_cfg[Options.E_BEAM_DEPTH], 'E cube nu depth')
        for name, dtype, default in schema.PROPERTIES:
            self.register_property(name, dtype, default)
        for entry in schema.ARRAYS:
            self.register_array(*entry)
        self.create_arrays(supplied=slvr_cfg.get('array_cfg', {}).get('supplied', None)
                           if isinstance(slvr_cfg.get('array_cfg'), dict) else None)

    # ---- flags (rime_solver.py:73-91) --------------------------------
    def is_float(self):
        return self.ft == np.float32

    def is_double(self):
        return self.ft == np.float64

    def use_weight_vector(self):
        return bool(self._use_weight_vector)

    def outputs_model_visibilities(self):
        return self._visibility_output == Options.VISIBILITY_OUTPUT_MODEL

    def outputs_residuals(self):
        return self._visibility_output == Options.VISIBILITY_OUTPUT_RESIDUALS

    def is_autocorrelated(self):
        return self._is_auto_correlated is True

    def config(self):
        return self._slvr_cfg

    def type_dict(self):
        return {'ft': self.ft, 'ct': self.ct, 'int': int}

    # ---- dimensions ----------------------------------------------------
    def register_dimension(self, name, size, description='', zero_valid=False):
        size = int(size)
        if size < 0 or (size == 0 and not zero_valid):
            raise ValueError("Dimension '%s' has invalid size %d" % (name, size))
        self._dims[name] = Dimension(name, size, description, zero_valid)

    def register_default_dimensions(self):
        """rime_solver.py:101-161."""
        cfg = self._slvr_cfg
        autocor = cfg.get(Options.AUTO_CORRELATIONS, False)
        self.register_dimension('ntime', cfg[Options.NTIME], 'Timesteps')
        self.register_dimension('na', cfg[Options.NA], 'Antenna')
        self.register_dimension('nbands', cfg[Options.NBANDS], 'Bands')
        self.register_dimension('nchan', cfg[Options.NCHAN], 'Channels')
        self.register_dimension('npol', cfg[Options.NPOL], 'Polarisations')
        ntime, na, nchan, npol = self.dim_local_size('ntime', 'na', 'nchan', 'npol')
        nbl = cfg.get(Options.NBL) or nr_of_baselines(na, autocor)
        self.register_dimension('nbl', nbl, 'Baselines')
        self.register_dimension('npolchan', nchan * npol, 'Polarised channels')
        self.register_dimension('nvis', ntime * nbl * nchan, 'Visibilities')
        nr_vars = src_types.sources_to_nr_vars(cfg[Options.SOURCES])
        self.register_dimension('nsrc', sum(nr_vars.values()), 'Sources (total)')
        for nr_var, n in nr_vars.items():
            self.register_dimension(nr_var, n, nr_var, zero_valid=True)

    def _dim_attr(self, attr, names):
        vals = tuple(getattr(self._dims[n], attr) for n in names)
        return vals[0] if len(vals) == 1 else vals

    def dim_local_size(self, *names):
        return self._dim_attr('local_size', names)

    def dim_global_size(self, *names):
        return self._dim_attr('global_size', names)

    def dim_lower_extent(self, *names):
        return self._dim_attr('lower_extent', names)

    def dim_upper_extent(self, *names):
        return self._dim_attr('upper_extent', names)

    def dim_extents(self, *names):
        vals = tuple((self._dims[n].lower_extent, self._dims[n].upper_extent) for n in names)
        return vals[0] if len(vals) == 1 else vals

    def dim_local_size_dict(self):
        return {n: d.local_size for n, d in self._dims.items()}

    def dimensions(self):
        return self._dims

    def update_dimension(self, name, global_size=None, local_size=None,
                         lower_extent=None, upper_extent=None):
        """Change a dimension's sizes or extents (docs/dims.rst semantics).
        Array shapes are not changed; extents select what the kernels touch."""
        d = self._dims[name]
        if global_size is not None:
            d.global_size = int(global_size)
        if local_size is not None:
            d.local_size = int(local_size)
        if lower_extent is not None:
            d.lower_extent = int(lower_extent)
        if upper_extent is not None:
            d.upper_extent = int(upper_extent)
        if not (0 <= d.lower_extent <= d.upper_extent <= d.global_size):
            raise ValueError("Dimension '%s': extents [%d, %d) outside global size %d"
                             % (name, d.lower_extent, d.upper_extent, d.global_size))

    def const_data(self):
        """The dims struct the library takes (const_data.py:118-169)."""
        d = _lib.mbx_dims()
        for name in _lib.DIM_NAMES:
            x = self._dims[name]
            setattr(d, name, _lib.mbx_dim(x.global_size, x.local_size,
                                          x.lower_extent, x.upper_extent))
        # nsrc extents are the sums of the per-type extents (const_data.py:135-145)
        lo = sum(self._dims[n].lower_extent for n in src_types.source_nr_vars())
        hi = sum(self._dims[n].upper_extent for n in src_types.source_nr_vars())
        ext = sum(self._dims[n].upper_extent - self._dims[n].lower_extent
                  for n in src_types.source_nr_vars())
        d.nsrc.lower_extent = lo
        d.nsrc.upper_extent = lo + ext if hi >= lo else lo
        npc = self._dims['npolchan']
        d.npolchan = _lib.mbx_dim(npc.global_size, npc.local_size, npc.lower_extent,
                                  npc.upper_extent)
        return d

    # ---- antenna pairs (v4/ant_pairs.py:23-37) --------------------------
    def default_base_ant_pairs(self):
        na = self.dim_local_size('na')
        return np.triu_indices(na, 0 if self.is_autocorrelated() else 1)

    def default_ant_pairs(self):
        ntime, nbl = self.dim_local_size('ntime', 'nbl')
        a1, a2 = (a[:nbl] for a in self.default_base_ant_pairs())
        return (np.tile(a1, ntime).reshape(ntime, nbl), np.tile(a2, ntime).reshape(ntime, nbl))

    # ---- properties ------------------------------------------------------
    def register_property(self, name, dtype, default):
        """Property read back in the solver's type (slvr.two_pi_over_c is a
        float32 in a float solver, as in the reference); the value as given
        is kept in float64 and that is what the kernels receive, so the
        float64 phase argument is not built on a float32-rounded 2 pi / C."""
        ft = self.type_dict().get(dtype, dtype) if isinstance(dtype, str) else dtype
        self._props[name] = ft
        setattr(self, '_prop_' + name, np.float64(default))

        def setter(value, _n=name):
            setattr(self, '_prop_' + _n, np.float64(value))
            if _n == 'X2':
                self._x2_dev = None
        setattr(self, 'set_' + name, setter)

    def exact_property(self, name):
        """The float64 value of a property as it was given."""
        if name == 'X2':
            self._read_x2()
        return float(getattr(self, '_prop_' + name))

    def __getattr__(self, name):
        # Array and property access as attributes (slvr.uvw, slvr.sigma_sqrd)
        d = self.__dict__
        if name == 'X2' and '_props' in d:
            return self._read_x2()
        if '_props' in d and name in d['_props']:
            return d['_props'][name](d['_prop_' + name])
        if '_arrays' in d and name in d['_arrays']:
            return self.tensor(name)
        raise AttributeError(name)

    def _read_x2(self):
        if self._x2_dev is not None:
            self._prop_X2 = np.float64(self._x2_dev.item())
            self._x2_dev = None
        return self.ft(self._prop_X2)

    def x2_device_scalar(self):
        """Device float64 scalar the reduction writes into; X2 reads it."""
        t = self._tensors.get('_X2')
        if t is None:
            t = torch.zeros((), dtype=torch.float64, device=self.device)
            self._tensors['_X2'] = t
        self._x2_dev = t
        return t

    def properties(self):
        return OrderedDict((n, getattr(self, n)) for n in self._props)

    # ---- arrays ----------------------------------------------------------
    def register_array(self, name, shape, dtype, default=None, test=None, scratch=False):
        dt = np.dtype(self.type_dict().get(dtype, dtype) if isinstance(dtype, str) else dtype)
        self._arrays[name] = dict(name=name, shape=shape, dtype=dt, default=default,
                                  test=test, scratch=scratch)

    def array_shape(self, name):
        return tuple(self._dims[s].local_size if isinstance(s, str) else int(s)
                     for s in self._arrays[name]['shape'])

    def arrays(self):
        return self._arrays

    def tensor(self, name):
        """The device tensor of array ``name`` (scratch arrays on first use)."""
        t = self._tensors.get(name)
        if t is None:
            a = self._arrays[name]
            t = torch.zeros(self.array_shape(name), dtype=_TORCH[a['dtype']], device=self.device)
            self._tensors[name] = t
        return t

    def init_array(self, name, ary, value):
        """montblanc/solvers/rime_solver.py:295-350."""
        if value is None:
            ary.fill(0)
        elif callable(value):
            out = value(self, ary)
            if out is not None:
                ary[...] = out
        elif isinstance(value, np.ndarray):
            ary[...] = value
        else:
            ary.fill(value)

    def create_arrays(self, ignore=None, supplied=None):
        ignore = set(ignore or [])
        supplied = supplied or {}
        # 'ms' (the factory transfers the MeasurementSet's arrays after
        # construction) and 'default' start from the schema defaults
        # (mb_cuda_solver.py:43-61)
        source = self._slvr_cfg.get(Options.DATA_SOURCE, Options.DATA_SOURCE_DEFAULT)
        for name, a in self._arrays.items():
            if name in ignore:
                continue
            shape = self.array_shape(name)
            if name in supplied:
                s = np.asarray(supplied[name])
                if s.shape != shape:
                    raise ValueError("Supplied array '%s's' shape '%s' does not match the "
                                     "expected shape of '%s'" % (name, s.shape, shape))
                self.transfer(name, s.astype(a['dtype'], copy=False))
                continue
            if a['scratch']:
                continue    # allocated on first use (jones is very large)
            if source == Options.DATA_SOURCE_EMPTY:
                self.tensor(name)
                continue
            init = a['test'] if source == Options.DATA_SOURCE_TEST else a['default']
            host = np.empty(shape, dtype=a['dtype'])
            self.init_array(name, host, init)
            self.transfer(name, host)
        for name in self._arrays:
            setattr(self, 'transfer_' + name, (lambda ary, _n=name: self.transfer(_n, ary)))
            setattr(self, 'retrieve_' + name, (lambda _n=name: self.retrieve(_n)))

    def pair_plan(self):
        """Device launch plan of the current antenna layout (rebuilt after
        antenna1/antenna2 transfers); None for the generic kernel."""
        if self._pair_plan is None:
            from .rime import build_plan, default_pair_block
            form = default_pair_block('f32' if self.is_float() else 'f64', self.dim_local_size('na'))
            for i, name in enumerate(('antenna1', 'antenna2')):
                if self._host_antenna[i] is None:
                    self._host_antenna[i] = self.retrieve(name)
            plan = build_plan(self._host_antenna[0], self._host_antenna[1],
                              self.dim_local_size('na'), self.device, form)
            self._pair_plan = plan if plan is not None else False
        return self._pair_plan or None

    def transfer(self, name, ary):
        """Host ndarray -> device array; shape and dtype must match (H2D set)."""
        a = self._arrays[name]
        ary = np.asarray(ary)
        shape = self.array_shape(name)
        if ary.shape != shape:
            raise ValueError("transfer_%s: shape %s does not match %s" % (name, ary.shape, shape))
        if ary.dtype != a['dtype']:
            raise ValueError("transfer_%s: dtype %s does not match %s" % (name, ary.dtype, a['dtype']))
        t = self.tensor(name)
        if t.numel():
            t.copy_(torch.from_numpy(np.ascontiguousarray(ary)))
        if name in ('antenna1', 'antenna2'):
            i = 0 if name == 'antenna1' else 1
            self._host_antenna[i] = np.array(ary, dtype=np.int32, copy=True)
            self._pair_plan = None

    def retrieve(self, name):
        """Device array -> host ndarray (D2H get)."""
        t = self.tensor(name)
        return t.detach().cpu().numpy().copy()

    def ekb_vis_tensor(self):
        """(ntime, nbl, nchan, 4) complex pre-gain source sum (adjoint input,
        and the running sum of partial source batches)."""
        t = self._tensors.get('_ekb_vis')
        if t is None:
            ntime, nbl, nchan = self.dim_local_size('ntime', 'nbl', 'nchan')
            t = torch.zeros((ntime, nbl, nchan, 4), dtype=_TORCH[np.dtype(self.ct)],
                            device=self.device)
            self._tensors['_ekb_vis'] = t
        return t

    def workspace(self, key, nbytes):
        """A device byte buffer of at least nbytes, kept between solves."""
        nbytes = max(int(nbytes), 1)
        t = self._workspaces.get(key)
        if t is None or t.numel() < nbytes:
            t = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
            self._workspaces[key] = t
        return t

    # ---- lifecycle (mb_cuda_solver.py:171-191) ---------------------------
    def solve(self):
        with torch.cuda.device(self.device):
            return self.pipeline.execute(self)

    def initialise(self):
        with torch.cuda.device(self.device):
            return self.pipeline.initialise(self)

    def shutdown(self):
        with torch.cuda.device(self.device):
            return self.pipeline.shutdown(self)

    def __enter__(self):
        self.initialise()
        return self

    def __exit__(self, etype, value, tb):
        self.shutdown()
