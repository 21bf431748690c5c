"""Memory planner and chunked solver for problems larger than one device.

The planner keeps the reference's reduction-ladder semantics
(montblanc/util/__init__.py:166-285 viable_dim_config, driven by
montblanc/impl/rime/v5/CompositeRimeSolver.py:644-718 _thread_budget):
reductions are applied in order until the device arrays fit the budget;
'ntime=20' lowers a dimension to 20 (never raises it), 'nchan=50%' halves
it repeatedly, 'nbl=1&na=2' lowers several at once, and MemoryError is
raised when the ladder runs out.

What changes for one MI355X:
* the budget is the device's free memory (hipMemGetInfo through torch)
  less a reserve, or cfg['mem_budget'];
* the arrays budgeted are those the fused path actually keeps on the
  device -- the v4 inputs and outputs, the pre-gain sums when gradients
  are wanted and the prepared beam cube -- not the reference's
  (nsrc, ntime, na, nchan, 4) jones scratch, which the fused kernel never
  materialises (the staged pipeline still budgets it);
* one solver per device (no per-device sub-solver streams), so the ladder
  first halves the time chunk, then batches sources (staged pipeline
  only: the fused path's source arrays are small), then halves channels.
"""

import logging
import math
import os
import re
from collections import OrderedDict

import numpy as np

from . import schema, src_types
from .config import RimeSolverConfig as Options

log = logging.getLogger('montblanc')

ONE_MB = 1024 ** 2
DEVICE_RESERVE = 512 * ONE_MB

_RED_RE = re.compile(r'^\s*(?P<name>[A-Za-z_0-9]+)\s*(=\s*(?P<value>\d+)\s*(?P<percent>%)?)?\s*$')

FUSED_ARRAYS = ('uvw', 'antenna1', 'antenna2', 'frequency', 'ref_frequency', 'parallactic_angles',
                'point_errors', 'antenna_scaling', 'E_beam', 'G_term', 'lm', 'stokes', 'alpha',
                'gauss_shape', 'sersic_shape', 'flag', 'weight_vector', 'observed_vis', 'model_vis',
                'chi_sqrd_result')


def _itemsize(dtype, ft):
    if isinstance(dtype, str):
        return {'ft': np.dtype(ft).itemsize, 'ct': 2 * np.dtype(ft).itemsize}[dtype]
    return np.dtype(dtype).itemsize


def array_bytes(shape, dtype, dims, ft):
    n = 1
    for s in shape:
        n *= dims[s] if isinstance(s, str) else int(s)
    return n * _itemsize(dtype, ft)


def device_bytes(dims, ft, pipeline='fused', gradients=False):
    """Device bytes the chosen pipeline needs for problem dims."""
    total = 0
    for name, shape, dtype, _d, _t, scratch in schema.ARRAYS:
        if pipeline == 'fused' and (scratch or name not in FUSED_ARRAYS):
            continue
        total += array_bytes(shape, dtype, dims, ft)
    vis = dims['ntime'] * dims['nbl'] * dims['nchan'] * 4 * 2 * np.dtype(ft).itemsize
    beam = dims['beam_lw'] * dims['beam_mh'] * dims['beam_nud'] * 4 * 4 * np.dtype(ft).itemsize
    total += beam
    if gradients:
        total += 2 * vis          # pre-gain sums + the adjoint's Vb workspace
    return total


def apply_reductions(budget, dims, ladder, bytes_of):
    """viable_dim_config semantics: returns (viable, reduced dims dict)."""
    ladder = list(ladder)
    T = dict(dims)
    modified = OrderedDict()
    if budget < 0:
        budget = 0
    used = bytes_of(T)
    while used > budget:
        if not ladder:
            return False, modified, T
        step = ladder.pop(0)
        for part in step.strip().split('&'):
            m = _RED_RE.match(part)
            if not m:
                raise ValueError("%s is an invalid dimension reduction string; valid strings are "
                                 "e.g. 'ntime', 'ntime=20' or 'ntime=20%%'" % part)
            name, value, pct = m.group('name'), m.group('value'), m.group('percent')
            value = 1 if value is None else int(value)
            if pct == '%':
                value = int(T[name] * value / 100.0)
                if value < 1:
                    value = 1
                else:
                    ladder.insert(0, part)
            if T[name] > value:
                T[name] = value
                modified[name] = value
            else:
                log.info('Ignored reduction of %s of size %s to %s.', name, T[name], value)
        used = bytes_of(T)
    return True, modified, T


def default_ladder(source_batch_size, pipeline='fused'):
    ladder = ['ntime=50%']
    if pipeline != 'fused':
        srcs = '&'.join('%s=%d' % (n, source_batch_size)
                        for n in ['nsrc'] + src_types.source_nr_vars())
        ladder.append(srcs)
    ladder.append('nchan=50%')
    return ladder


def device_budget(cfg=None, device=None):
    if cfg is not None and cfg.get('mem_budget') is not None:
        return int(cfg['mem_budget'])
    import torch
    free, _total = torch.cuda.mem_get_info(device)
    return int(free) - DEVICE_RESERVE


def plan(dims, ft, budget, pipeline='fused', gradients=False, source_batch_size=500, ladder=None):
    """Reduced per-chunk dims that fit ``budget`` bytes.

    Raises MemoryError, listing the largest arrays, when no reduction fits
    (CompositeRimeSolver.py:703-716).
    """
    ladder = default_ladder(source_batch_size, pipeline) if ladder is None else ladder

    def bytes_of(T):
        T2 = dict(T)
        T2['nvis'] = T2['ntime'] * T2['nbl'] * T2['nchan']
        T2['npolchan'] = T2['nchan'] * T2.get('npol', 4)
        return device_bytes(T2, ft, pipeline, gradients)
    viable, modified, T = apply_reductions(budget, dims, ladder, bytes_of)
    if not viable:
        worst = sorted(((array_bytes(s, dt, dict(T, npolchan=4 * T['nchan']), ft), n)
                        for n, s, dt, *_ in schema.ARRAYS), reverse=True)[:5]
        raise MemoryError(
            "Tried reducing the problem size by setting '%s' on all arrays, but the required "
            "memory of %d bytes does not fit the budget of %d bytes. Largest arrays: %s"
            % (', '.join('%s=%s' % kv for kv in modified.items()), bytes_of(T), budget,
               ', '.join('%s %d' % (n, b) for b, n in worst)))
    return modified, T, bytes_of(T)


def chunks(n, size):
    """[start, end) ranges of at most ``size`` covering [0, n)."""
    size = max(1, int(size))
    return [(s, min(n, s + size)) for s in range(0, n, size)]


# Arrays with neither a time nor a channel axis (source parameters, the
# beam cube): sent to the device once and kept there across chunks and
# solves, sent again only after a transfer_<name> (the reference's v5
# caches unchanged array slices the same way, CompositeRimeSolver.py:
# 433-593).
STATIC_ARRAYS = tuple(n for n, shape, _dt, _d, _t, scratch in schema.ARRAYS
                      if not scratch and 'ntime' not in shape and 'nchan' not in shape)
INT_ARRAYS = ('antenna1', 'antenna2')


def pin_limit():
    """Host bytes the chunked solver may keep page-locked for its full-size
    arrays: a quarter of the physical memory, at most 16 GiB; 0 without a
    device (page-locked memory belongs to the device runtime)."""
    import torch
    if not torch.cuda.is_available():
        return 0
    try:
        phys = os.sysconf('SC_PHYS_PAGES') * os.sysconf('SC_PAGE_SIZE')
    except (ValueError, OSError, AttributeError):
        phys = 0
    return int(min(16 << 30, phys // 4))


def _device_ready(name, v):
    """A host array in the dtype the device path takes (as rime.to_device)."""
    if name in INT_ARRAYS:
        return np.ascontiguousarray(v, dtype=np.int32)
    if name == 'flag':
        return np.ascontiguousarray(v, dtype=np.uint8)
    return np.ascontiguousarray(v)


class ChunkedRimeSolver(object):
    """The v4 solver interface over problems that do not fit one device.

    Full-size arrays live on the host (transfer_<name>/retrieve_<name>
    work on them); solve() runs the fused forward chunk by chunk over time
    and channels (the fused forward's ladder reduces those two: its
    resident arrays do not grow with the source count the way the staged
    pipeline's per-source scratch does), writing model visibilities back to the host array
    (overwritten, or accumulated with vis_write_mode='sum') and summing
    X2 over chunks -- the host side of the reference's v5 solve loop
    (CompositeRimeSolver.py:793-1012, 1027-1215), on one device.

    The loop is pipelined: chunk k runs on the current stream while a copy
    stream moves chunk k + 1's inputs in and chunk k - 1's outputs out, so
    the host-device copies overlap the kernels. Two chunks are then
    resident at once: the planner is given half of the device budget.
    Arrays without time or channel axes stay on the device between chunks
    and solves (STATIC_ARRAYS). Where the full-size host arrays fit
    pin_limit() they are page-locked, and a chunk's slices that are
    contiguous in them (time-major arrays, chunks over time only) go to and
    from the device directly; other slices pass through page-locked staging
    buffers.
    """

    def __init__(self, slvr_cfg):
        import torch
        from . import rime
        self._cfg = slvr_cfg
        self.ft = np.float32 if slvr_cfg[Options.DTYPE] == Options.DTYPE_FLOAT else np.float64
        self.ct = np.complex64 if self.ft == np.float32 else np.complex128
        self._use_w = bool(slvr_cfg.get(Options.WEIGHT_VECTOR))
        self._residual = slvr_cfg.get(Options.VISIBILITY_OUTPUT) == Options.VISIBILITY_OUTPUT_RESIDUALS
        self._sum = slvr_cfg.get(Options.VISIBILITY_WRITE_MODE) == Options.VISIBILITY_WRITE_MODE_SUM
        from .solver import RimeSolver, resolve_device
        self.device = resolve_device(slvr_cfg.get(Options.CONTEXT))
        # a host-side twin holds dims, properties and the full arrays
        cfg = dict(slvr_cfg)
        self._host = _HostArrays(cfg, self.ft, self.ct, pin_limit())
        self.X2 = self.ft(0.0)
        self.last_plan = None
        self.last_solve = None
        self._rime = rime
        self._torch = torch
        self._copy = None          # copy stream, created on first solve
        self._static = {}          # device copies of STATIC_ARRAYS

    def __getattr__(self, name):
        h = self.__dict__.get('_host')
        if h is None:
            raise AttributeError(name)
        if name.startswith('transfer_'):
            return lambda ary, _n=name[9:]: h.transfer(_n, ary)
        if name.startswith('retrieve_'):
            return lambda _n=name[9:]: h.arrays[_n].copy()
        if name.startswith('set_'):
            return lambda v, _n=name[4:]: h.props.__setitem__(_n, np.float64(v))
        if name in h.props:
            # read back in the solver's type; kept (and used) in float64
            return self.ft(h.props[name])
        if name in h.arrays:
            return h.arrays[name]
        if name.startswith('dim_'):
            return getattr(h, name)
        raise AttributeError(name)

    def plan(self):
        h = self._host
        dims = dict(h.dims)
        # two chunks in flight (see the class docstring)
        budget = device_budget(self._cfg, self.device) // 2
        modified, T, need = plan(dims, self.ft, budget,
                                 source_batch_size=self._cfg.get(Options.SOURCE_BATCH_SIZE, 500))
        self.last_plan = dict(modified=dict(modified), chunk_dims=T, bytes=need, budget=budget)
        log.info('montblanc planner: %s per chunk (%d bytes of %d)', dict(modified), need, budget)
        return T

    def _static_on_device(self):
        """Device copies of STATIC_ARRAYS, re-sent where transfer_* changed
        them since the last solve. Returns the names sent."""
        torch, h = self._torch, self._host
        sent = []
        for n in STATIC_ARRAYS:
            if n in self._static and n not in h.dirty:
                continue
            self._static[n] = torch.from_numpy(_device_ready(n, h.arrays[n])).to(self.device)
            sent.append(n)
        h.dirty.difference_update(STATIC_ARRAYS)
        return sent

    def solve(self):
        rime, torch = self._rime, self._torch
        h = self._host
        T = self.plan()
        nt, nc = h.dims['ntime'], h.dims['nchan']
        counts = (h.dims['npsrc'], h.dims['ngsrc'], h.dims['nssrc'])
        jobs = [(t0, t1, c0, c1) for t0, t1 in chunks(nt, T['ntime']) for c0, c1 in chunks(nc, T['nchan'])]
        model = h.arrays['model_vis']      # every chunk overwrites (or adds to) its slice
        chi = h.arrays['chi_sqrd_result']
        dev = self.device
        comp = torch.cuda.current_stream(dev)
        if self._copy is None:
            self._copy = torch.cuda.Stream(device=dev)   # default priority
        copy = self._copy
        sent = self._static_on_device()
        h.dirty.clear()
        x2 = torch.zeros((), dtype=torch.float64, device=dev)
        sigma = float(h.props['sigma_sqrd'])

        sfx = 'f32' if self.ft == np.float32 else 'f64'
        na = h.dims['na']
        plans = {}     # launch plans by antenna layout, built from the host arrays

        def stage(job):
            # chunk inputs: host slice -> pinned buffer -> device, on the
            # copy stream; an event marks them ready
            t0, t1, c0, c1 = job
            a = h.slice(t0, t1, c0, c1, skip=STATIC_ARRAYS)
            arrays = dict(self._static)
            with torch.cuda.stream(copy):
                # the launch plan from the host layout (the forward would
                # otherwise copy the antennas back from the device, blocking
                # the pipeline), shared by chunks of the same layout
                a1 = _device_ready('antenna1', a['antenna1'])
                a2 = _device_ready('antenna2', a['antenna2'])
                lk = (a1.shape, a1.tobytes(), a2.tobytes())
                if lk not in plans:
                    plans[lk] = rime.plan_entry(a1, a2, na, dev, sfx)
                arrays['_pair_plan'] = plans[lk]
                for k, v in a.items():
                    if not isinstance(v, np.ndarray):
                        arrays[k] = v
                        continue
                    src = torch.from_numpy(_device_ready(k, v))
                    d = torch.empty(src.shape, dtype=src.dtype, device=dev)
                    if h.pinned and src.is_pinned():
                        d.copy_(src, non_blocking=True)      # a view of the page-locked array
                    else:
                        pin = torch.empty(src.shape, dtype=src.dtype, pin_memory=True)
                        pin.copy_(src)
                        d.copy_(pin, non_blocking=True)
                    d.record_stream(comp)
                    arrays[k] = d
                ready = torch.cuda.Event()
                ready.record(copy)
            return arrays, ready

        # outputs straight into the page-locked host arrays where the chunk
        # is contiguous in them (and the model is overwritten, not summed)
        direct = h.pinned and not self._sum and T['nchan'] == nc

        def drain(p):
            # a finished chunk's outputs into the host arrays
            (t0, t1, c0, c1), mv, cs, done = p
            done.synchronize()
            if direct:
                return
            if self._sum:
                model[t0:t1, :, c0:c1] += mv.numpy()
            else:
                model[t0:t1, :, c0:c1] = mv.numpy()
            chi[t0:t1, :, c0:c1] = cs.numpy()

        pending = None
        nxt = stage(jobs[0]) if jobs else None
        for k, job in enumerate(jobs):
            arrays, ready = nxt
            comp.wait_event(ready)
            out = rime.forward(arrays, counts, residual=self._residual, use_weight_vector=self._use_w,
                               sigma_sqrd=sigma)
            x2 += out['X2']
            computed = torch.cuda.Event()
            computed.record(comp)
            if k + 1 < len(jobs):
                nxt = stage(jobs[k + 1])      # overlaps this chunk's kernels
            with torch.cuda.stream(copy):
                copy.wait_event(computed)
                if direct:
                    t0, t1 = job[0], job[1]
                    mv = torch.from_numpy(model[t0:t1])
                    cs = torch.from_numpy(chi[t0:t1])
                else:
                    mv = torch.empty(out['model_vis'].shape, dtype=out['model_vis'].dtype,
                                     pin_memory=True)
                    cs = torch.empty(out['chi_sqrd_result'].shape, dtype=out['chi_sqrd_result'].dtype,
                                     pin_memory=True)
                mv.copy_(out['model_vis'], non_blocking=True)
                cs.copy_(out['chi_sqrd_result'], non_blocking=True)
                out['model_vis'].record_stream(copy)
                out['chi_sqrd_result'].record_stream(copy)
                done = torch.cuda.Event()
                done.record(copy)
            if pending is not None:
                drain(pending)
            pending = (job, mv, cs, done)
            del arrays, out
        if pending is not None:
            drain(pending)
        total = float(x2.item())
        self.X2 = self.ft(total)
        h.props['X2'] = np.float64(total)
        self.last_solve = dict(chunks=len(jobs), static_sent=sent)
        return False

    def initialise(self):
        return True

    def shutdown(self):
        return True

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False


class _HostArrays(object):
    """Full-size numpy arrays, dims and properties of the v4 schema;
    page-locked (views of pinned torch tensors, kept in _pins) when their
    total is at most pin_limit bytes."""

    def __init__(self, cfg, ft, ct, pin_limit=0):
        from .solver import nr_of_baselines
        autocor = cfg.get(Options.AUTO_CORRELATIONS, False)
        nr = src_types.sources_to_nr_vars(cfg[Options.SOURCES])
        na = cfg[Options.NA]
        self.dims = dict(ntime=cfg[Options.NTIME], na=na, nchan=cfg[Options.NCHAN], npol=4,
                         nbl=cfg.get(Options.NBL) or nr_of_baselines(na, autocor),
                         beam_lw=cfg[Options.E_BEAM_WIDTH], beam_mh=cfg[Options.E_BEAM_HEIGHT],
                         beam_nud=cfg[Options.E_BEAM_DEPTH], **nr)
        self.dims['nsrc'] = sum(nr.values())
        self.dims['npolchan'] = 4 * self.dims['nchan']
        self.dims['nvis'] = self.dims['ntime'] * self.dims['nbl'] * self.dims['nchan']
        self.ft, self.ct = ft, ct
        self.props = {n: np.float64(v) for n, _t, v in schema.PROPERTIES}
        self.arrays = {}
        self.dirty = set()          # arrays changed since the last solve
        a1, a2 = np.triu_indices(na, 0 if autocor else 1)
        nbl = self.dims['nbl']
        specs = []
        for name, shape, dtype, default, _test, scratch in schema.ARRAYS:
            if scratch:
                continue
            dt = np.dtype({'ft': ft, 'ct': ct}.get(dtype, dtype) if isinstance(dtype, str) else dtype)
            shp = tuple(self.dims[s] if isinstance(s, str) else s for s in shape)
            specs.append((name, shp, dt, default))
        total = sum(int(np.prod(shp)) * dt.itemsize for _n, shp, dt, _d in specs)
        self.pinned = 0 < total <= pin_limit
        self._pins = {}
        for name, shp, dt, default in specs:
            if self.pinned:
                import torch
                t = torch.empty(shp, dtype=torch.from_numpy(np.zeros(0, dt)).dtype, pin_memory=True)
                self._pins[name] = t
                arr = t.numpy()
                arr[...] = 0
            else:
                arr = np.zeros(shp, dtype=dt)
            if name == 'antenna1':
                arr[...] = a1[:nbl]
            elif name == 'antenna2':
                arr[...] = a2[:nbl]
            elif name == 'frequency':
                arr[...] = np.linspace(1e9, 2e9, shp[0])
            elif isinstance(default, np.ndarray) or np.isscalar(default):
                arr[...] = default
            self.arrays[name] = arr

    def transfer(self, name, ary):
        ary = np.asarray(ary)
        if ary.shape != self.arrays[name].shape:
            raise ValueError('transfer_%s: shape %s does not match %s'
                             % (name, ary.shape, self.arrays[name].shape))
        if ary.dtype != self.arrays[name].dtype:
            raise ValueError('transfer_%s: dtype %s does not match %s'
                             % (name, ary.dtype, self.arrays[name].dtype))
        self.arrays[name][...] = ary
        self.dirty.add(name)

    def dim_local_size(self, *names):
        v = tuple(self.dims[n] for n in names)
        return v[0] if len(v) == 1 else v

    dim_global_size = dim_local_size

    def slice(self, t0, t1, c0, c1, skip=()):
        """The chunk's arrays (time slice t0:t1, channel slice c0:c1),
        leaving out the names in ``skip``."""
        A = self.arrays
        out = {}
        for name, arr in A.items():
            if name in ('model_vis', 'chi_sqrd_result') or name in skip:
                continue
            shape = dict((n, s) for n, s, *_ in ((x[0], x[1]) for x in schema.ARRAYS))[name]
            idx = []
            for s in shape:
                if s == 'ntime':
                    idx.append(slice(t0, t1))
                elif s == 'nchan':
                    idx.append(slice(c0, c1))
                else:
                    idx.append(slice(None))
            out[name] = np.ascontiguousarray(arr[tuple(idx)])
        for k in ('beam_ll', 'beam_lm', 'beam_lfreq', 'beam_ul', 'beam_um', 'beam_ufreq',
                  'gauss_scale', 'two_pi_over_c', 'sigma_sqrd'):
            out[k] = float(self.props[k])
        return out
