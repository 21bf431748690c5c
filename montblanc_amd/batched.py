"""Batched evaluation of many sky-model samples against shared data.

``Batch`` holds one set of visibility-space inputs (uvw, antennas,
frequencies, beam, gains, flags, weights, observed visibilities) on the
device and the concatenated source lists of K samples; ``evaluate()`` runs
the library's single persistent launch over all samples and returns the
per-sample chi^2 and, on request, the gradients of every sample's chi^2
with respect to its own source parameters. No reference counterpart.
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


def tuning_struct(tune):
    if tune is None:
        return None
    t = _lib.mbx_batch_tuning(int(tune.get('workgroups', 0)), int(tune.get('channels_per_item', 0)),
                              int(tune.get('sources_per_tile', 0)), 0)
    return t


def evaluate(batch, gradients=False, use_weight_vector=False, sigma_sqrd=None, tune='auto',
             stream=None, dims=None):
    """Per-sample chi^2 (float64 device tensor [K]) and, with gradients=True,
    a dict of float64 gradient tensors over the concatenated sources.

    tune: a dict (workgroups, channels_per_item, sources_per_tile), None for
    the library defaults, or 'auto' for the cached autotuner choice.
    """
    if tune == 'auto':
        from . import autotune
        tune = autotune.batched_tuning(batch, gradients=gradients)
    a = batch.arrays
    d = batch.dims() if dims is None else dims
    ts = tuning_struct(tune)
    tref = ctypes.byref(ts) if ts is not None else None
    lib = _lib.load()
    nb = lib.mbx_batched_workspace_bytes(ctypes.byref(d), int(rime.is_double(a)), tref)
    if nb == 0:
        raise _lib.MontblancHipError(-5, 'mbx_batched_workspace_bytes', lib.mbx_last_error().decode())
    if batch._ws is None or batch._ws.numel() < nb:
        batch._ws = torch.empty(nb, dtype=torch.uint8, device=batch.device)
    ins = _lib.mbx_inputs(**{n: rime._raw(a.get(n)) for n in _lib.INPUT_NAMES})
    g = None
    bmap = None
    if gradients:
        gb = batch.grads_buffers()
        g = _lib.mbx_grads(**{n: rime._raw(gb.get(n)) for n in _lib.GRAD_NAMES})
        bmap = rime.baseline_map(a)
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
