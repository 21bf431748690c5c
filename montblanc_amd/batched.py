"""Batched evaluation of many sky-model samples against shared data.

``Batch`` holds one set of visibility-space inputs (uvw, antennas,
frequencies, beam, gains, flags, weights, observed visibilities) on the
device and the concatenated source lists of K samples; ``evaluate()`` returns the
per-sample chi^2 and, on request, the gradients of every sample's chi^2
with respect to its own source parameters, by one of three strategies:
the library's matrix-core passes each launched once over a group of
samples, its single persistent launch on the vector units, or every
sample through the single-model entry points in a captured graph. No
reference counterpart.
"""

import ctypes

import numpy as np
import torch

from . import _lib, rime

SOURCE_ROWS = ('lm', 'stokes', 'alpha')


class Batch(object):
    """K sky-model samples over one shared observation.

    shared: dict (v4 names) of the visibility-space inputs and beam
        extents/constants (numpy or torch); its source entries are ignored.
    samples: list of dicts with 'lm' (n, 2), 'stokes' (n, ntime, 4),
        'alpha' (n, ntime), 'gauss_shape' (3, ng), 'sersic_shape' (3, ns)
        and 'counts' = (npsrc, ngsrc, nssrc) with n = npsrc + ngsrc + nssrc.
    """

    def __init__(self, shared, samples, device='cuda:0', dtype=np.float64):
        self.device = torch.device(device)
        self.dtype = np.dtype(dtype)
        base = {k: v for k, v in shared.items() if k not in SOURCE_ROWS + ('gauss_shape', 'sersic_shape')}
        self.counts = [tuple(int(x) for x in s['counts']) for s in samples]
        ntime = np.asarray(shared['uvw']).shape[0]
        ft = self.dtype
        cat = {
            'lm': np.concatenate([np.asarray(s['lm'], ft).reshape(-1, 2) for s in samples]),
            'stokes': np.concatenate([np.asarray(s['stokes'], ft).reshape(-1, ntime, 4) for s in samples]),
            'alpha': np.concatenate([np.asarray(s['alpha'], ft).reshape(-1, ntime) for s in samples]),
            'gauss_shape': np.concatenate([np.asarray(s['gauss_shape'], ft).reshape(3, -1) for s in samples], 1),
            'sersic_shape': np.concatenate([np.asarray(s['sersic_shape'], ft).reshape(3, -1) for s in samples], 1),
        }
        host = dict(base)
        host.update(cat)
        self.arrays = rime.to_device(host, device, dtype=ft)
        n = np.array([sum(c) for c in self.counts], np.int32)
        cnt = np.array(self.counts, np.int32).reshape(-1, 3)
        src_off = np.concatenate([[0], np.cumsum(n)]).astype(np.int32)
        g_off = np.concatenate([[0], np.cumsum(cnt[:, 1])[:-1]]).astype(np.int32)
        s_off = np.concatenate([[0], np.cumsum(cnt[:, 2])[:-1]]).astype(np.int32)
        order = np.argsort(-n, kind='stable').astype(np.int32)      # largest first
        self.src_offset = src_off
        self._dev = {k: torch.from_numpy(np.ascontiguousarray(v)).to(device)
                     for k, v in dict(src_offset=src_off, npsrc=cnt[:, 0].copy(),
                                      ngsrc=cnt[:, 1].copy(), nssrc=cnt[:, 2].copy(),
                                      gauss_offset=g_off, sersic_offset=s_off, order=order).items()}
        self.batch = _lib.mbx_batch(len(samples), *(self._dev[k].data_ptr() for k in (
            'src_offset', 'npsrc', 'ngsrc', 'nssrc', 'gauss_offset', 'sersic_offset', 'order')))
        self.totals = tuple(int(x) for x in cnt.sum(0))
        self.X2 = torch.zeros(len(samples), dtype=torch.float64, device=device)
        self._grads = None
        self._ws = None
        self._per_sample = {}
        self._matrix_group = None

    @property
    def nsamples(self):
        return len(self.counts)

    def dims(self):
        return rime.make_dims(self.arrays, self.totals)

    def grads_buffers(self):
        if self._grads is None:
            a = self.arrays
            f64 = dict(dtype=torch.float64, device=self.device)
            self._grads = dict(lm=torch.zeros(a['lm'].shape, **f64),
                               stokes=torch.zeros(a['stokes'].shape, **f64),
                               alpha=torch.zeros(a['alpha'].shape, **f64),
                               gauss_shape=torch.zeros(a['gauss_shape'].shape, **f64),
                               sersic_shape=torch.zeros(a['sersic_shape'].shape, **f64))
        return self._grads

    def split(self, t):
        """Split a concatenated per-source tensor into per-sample pieces."""
        o = self.src_offset
        return [t[o[k]:o[k + 1]] for k in range(self.nsamples)]

    def split_shapes(self, t, kind):
        col = 1 if kind == 'gauss' else 2
        out, start = [], 0
        for c in self.counts:
            out.append(t[:, start:start + c[col]])
            start += c[col]
        return out


class _PerSample(object):
    """The per-sample strategy: every sample's chi^2 (and gradients) by the
    single-model entry points -- point sources on the matrix cores, the
    adjoint's product too -- the whole sequence captured as one HIP graph
    and replayed. The samples are spread over a few streams (each with its
    own scratch outputs and workspaces, balanced by source count), so that
    small samples run side by side instead of one small grid at a time.

    ntime_probe limits the time extent (autotuner probes)."""

    STREAMS = 4

    def __init__(self, batch, gradients, use_weight_vector, sigma_sqrd, ntime_probe=None):
        from . import graph
        self.batch = batch
        self.gradients = gradients
        self.kw = dict(use_weight_vector=use_weight_vector, sigma_sqrd=sigma_sqrd)
        a = batch.arrays
        shared = {k: v for k, v in a.items() if k not in SOURCE_ROWS + ('gauss_shape', 'sersic_shape')}
        # one launch plan and baseline map for every sample (same layout)
        rime.pair_plan(a)
        if gradients:
            rime.baseline_map(a)
        shared['_pair_plan'] = a['_pair_plan']
        if '_baseline_map' in a:
            shared['_baseline_map'] = a['_baseline_map']
        ft = a['uvw'].dtype
        ntime, nbl = a['antenna1'].shape
        nchan = a['frequency'].shape[0]
        dev = batch.device
        ct = rime.complex_dtype(a)
        # each stream holds its own scratch (model and pre-gain sums, chi^2
        # terms, and a workspace of about one more visibility array): at
        # most a quarter of the free device memory goes to them
        per_group = ntime * nbl * nchan * (4 * ct.itemsize * (3 if gradients else 2) + ft.itemsize)
        free = torch.cuda.mem_get_info(dev)[0]
        ngroups = max(1, min(self.STREAMS, len(batch.counts), int(0.25 * free // max(per_group, 1))))
        self.streams = [torch.cuda.Stream(dev) for _ in range(ngroups)]
        self.workspaces = [rime.Workspace(dev) for _ in range(ngroups)]
        outs = []
        for _ in range(ngroups):
            o = dict(model_vis=torch.empty((ntime, nbl, nchan, 4), dtype=ct, device=dev),
                     chi_sqrd_result=torch.empty((ntime, nbl, nchan), dtype=ft, device=dev))
            if gradients:
                o['ekb_vis'] = torch.empty((ntime, nbl, nchan, 4), dtype=ct, device=dev)
            outs.append(o)
        if gradients:
            g = batch.grads_buffers()
        # largest samples first, each to the least loaded stream
        order = sorted(range(len(batch.counts)), key=lambda k: -sum(batch.counts[k]))
        load = [0] * ngroups
        self.groups = [[] for _ in range(ngroups)]
        g0s = np.concatenate([[0], np.cumsum([c[1] for c in batch.counts])])
        s0s = np.concatenate([[0], np.cumsum([c[2] for c in batch.counts])])
        o = batch.src_offset
        for k in order:
            c = batch.counts[k]
            grp = int(np.argmin(load))
            load[grp] += sum(c)
            sa = dict(shared)
            for n in SOURCE_ROWS:
                sa[n] = a[n][o[k]:o[k + 1]]
            g0, s0 = int(g0s[k]), int(s0s[k])
            sa['gauss_shape'] = a['gauss_shape'][:, g0:g0 + c[1]].contiguous()
            sa['sersic_shape'] = a['sersic_shape'][:, s0:s0 + c[2]].contiguous()
            dims = rime.make_dims(sa, c)
            if ntime_probe is not None:
                dims.ntime.upper_extent = min(dims.ntime.local_size, ntime_probe)
            out = dict(outs[grp])
            out['X2'] = batch.X2[k]
            entry = dict(arrays=sa, counts=c, dims=dims, out=out, g0=g0, s0=s0)
            if gradients:
                entry['grads'] = dict(lm=g['lm'][o[k]:o[k + 1]], stokes=g['stokes'][o[k]:o[k + 1]],
                                      alpha=g['alpha'][o[k]:o[k + 1]],
                                      gauss_shape=torch.empty((3, c[1]), dtype=torch.float64, device=dev),
                                      sersic_shape=torch.empty((3, c[2]), dtype=torch.float64, device=dev))
            self.groups[grp].append(entry)
        self.graph = graph.Graphed(self._run, device=dev)

    def _run_group(self, entries, ws):
        b = self.batch
        for e in entries:
            rime.forward(e['arrays'], e['counts'], outputs=e['out'], keep_ekb=self.gradients,
                         dims=e['dims'], workspace=ws, **self.kw)
            if self.gradients:
                rime.adjoint(e['arrays'], e['counts'], e['out']['ekb_vis'], grads=e['grads'],
                             dims=e['dims'], workspace=ws, **self.kw)
                g = b.grads_buffers()
                c = e['counts']
                if c[1]:
                    g['gauss_shape'][:, e['g0']:e['g0'] + c[1]].copy_(e['grads']['gauss_shape'])
                if c[2]:
                    g['sersic_shape'][:, e['s0']:e['s0'] + c[2]].copy_(e['grads']['sersic_shape'])

    def _run(self):
        # fork onto the group streams and join back (also inside capture)
        main = torch.cuda.current_stream(self.batch.device)
        for s in self.streams:
            s.wait_stream(main)
        for s, entries, ws in zip(self.streams, self.groups, self.workspaces):
            with torch.cuda.stream(s):
                self._run_group(entries, ws)
        for s in self.streams:
            main.wait_stream(s)
        return self.batch.X2

    def __call__(self):
        return self.graph.replay()


def drop_probes(batch):
    """Release the per-sample graphs built for shortened (probe) dims."""
    for k in [k for k in batch._per_sample if k[-1] is not None]:
        del batch._per_sample[k]


# fraction of the free device memory the matrix strategy's per-sample
# visibility buffers may take
MATRIX_MEMORY_FRACTION = 0.4


def matrix_group(batch):
    """Samples per group of launches for the matrix strategy: as many as
    fit MATRIX_MEMORY_FRACTION of the free device memory with two
    visibility arrays (running / pre-gain sums and Vb) per sample. Decided
    once per batch (so later calls, graph capture included, query no
    device memory)."""
    if batch._matrix_group is None:
        a = batch.arrays
        ntime, nbl = a['antenna1'].shape
        nchan = a['frequency'].shape[0]
        per = 2 * ntime * nbl * nchan * 4 * rime.complex_dtype(a).itemsize
        free = torch.cuda.mem_get_info(batch.device)[0]
        batch._matrix_group = max(1, min(batch.nsamples, int(MATRIX_MEMORY_FRACTION * free // per)))
    return batch._matrix_group


def tuning_struct(tune, batch=None):
    if tune is None or tune.get('strategy') == 'per_sample':
        return None
    if tune.get('strategy') == 'matrix':
        g = int(tune.get('samples_per_group', 0)) or matrix_group(batch)
        return _lib.mbx_batch_tuning(0, 0, 0, 1, g)
    return _lib.mbx_batch_tuning(int(tune.get('workgroups', 0)), int(tune.get('channels_per_item', 0)),
                                 int(tune.get('sources_per_tile', 0)), 0, 0)


def evaluate(batch, gradients=False, use_weight_vector=False, sigma_sqrd=None, tune='auto',
             stream=None, dims=None):
    """Per-sample chi^2 (float64 device tensor [K]) and, with gradients=True,
    a dict of float64 gradient tensors over the concatenated sources.

    tune: a dict (workgroups, channels_per_item, sources_per_tile) for the
    persistent launch; {'strategy': 'matrix'} (float, at most 64
    antennas; optional 'samples_per_group', default: what fits
    MATRIX_MEMORY_FRACTION of the free memory) for the matrix-core passes,
    each one launch over a group of samples; {'strategy': 'per_sample'} for
    the per-sample graph (which runs on the current stream; ``stream``
    applies to the library launches); None for the library defaults, or
    'auto' for the cached autotuner choice.
    """
    if tune == 'auto':
        from . import autotune
        tune = autotune.batched_tuning(batch, gradients=gradients)
    if isinstance(tune, dict) and tune.get('strategy') == 'per_sample':
        key = ('per_sample', bool(gradients), bool(use_weight_vector), sigma_sqrd,
               None if dims is None else int(dims.ntime.upper_extent))
        ps = batch._per_sample.get(key)
        if ps is None:
            ps = _PerSample(batch, gradients, use_weight_vector, sigma_sqrd,
                            None if dims is None else int(dims.ntime.upper_extent))
            batch._per_sample[key] = ps
        ps()
        if gradients:
            return batch.X2, batch.grads_buffers()
        return batch.X2
    a = batch.arrays
    d = batch.dims() if dims is None else dims
    ts = tuning_struct(tune, batch)
    tref = ctypes.byref(ts) if ts is not None else None
    matrix = ts is not None and ts.strategy == 1
    lib = _lib.load()
    nb = lib.mbx_batched_workspace_bytes(ctypes.byref(d), int(rime.is_double(a)), tref)
    if nb == 0:
        raise _lib.MontblancHipError(-5, 'mbx_batched_workspace_bytes', lib.mbx_last_error().decode())
    if batch._ws is None or batch._ws.numel() < nb:
        batch._ws = torch.empty(nb, dtype=torch.uint8, device=batch.device)
    ins = _lib.mbx_inputs(**{n: rime._raw(a.get(n)) for n in _lib.INPUT_NAMES})
    g = None
    bmap = rime.baseline_map(a) if (gradients or matrix) else None
    if gradients:
        gb = batch.grads_buffers()
        g = _lib.mbx_grads(**{n: rime._raw(gb.get(n)) for n in _lib.GRAD_NAMES})
    b = rime.beam_extents(a)
    k = rime.consts(a, sigma_sqrd)
    h = rime.stream_of(batch.device) if stream is None else ctypes.c_void_p(stream)
    _lib.call('mbx_batched_' + rime.suffix(a), ctypes.byref(d), ctypes.byref(ins),
              ctypes.byref(batch.batch), rime._ptr(batch.X2),
              ctypes.byref(g) if g is not None else None, rime._ptr(bmap), ctypes.byref(b),
              ctypes.byref(k), tref, rime.options(False, use_weight_vector),
              rime._ptr(batch._ws), nb, h)
    if gradients:
        return batch.X2, batch.grads_buffers()
    return batch.X2


def sample_counts(rng, nsamples, lo=10, hi=2000, gaussian_fraction=0.3):
    """Log-uniform source counts in [lo, hi], a fraction of them gaussian."""
    n = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), nsamples))).astype(int)
    ng = np.rint(n * gaussian_fraction).astype(int)
    return [(int(a - b), int(b), 0) for a, b in zip(n, ng)]
