"""Project-written autotuner for launch shapes.

Candidates are timed on the device (median of a few runs, synchronised)
on a probe of the real problem -- the same arrays with the time extent cut
to a few steps through the dims struct, so the per-item work is the real
one -- and the fastest is cached per (kernel, dtype, shape, device) in
memory and in a JSON file ($MBX_AUTOTUNE_CACHE, default
~/.cache/montblanc_amd/autotune.json). Cached choices are reused without
timing, so tuned calls are safe inside HIP-graph capture once warmed up.
"""

import ctypes
import json
import os
import time

import torch

from . import _lib, rime

_MEM = {}


def cache_path():
    return os.environ.get('MBX_AUTOTUNE_CACHE',
                          os.path.join(os.path.expanduser('~'), '.cache', 'montblanc_amd',
                                       'autotune.json'))


def _load_disk():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_disk(key, value):
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        data = _load_disk()
        data[key] = value
        tmp = path + '.tmp%d' % os.getpid()
        with open(tmp, 'w') as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass      # a read-only home only loses the persistence


def _device_name(device):
    return torch.cuda.get_device_name(device).replace(' ', '_')


def time_call(fn, repeats=3):
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[len(ts) // 2]


def tune(key, candidates, run, use_cache=True, measure=None, final=None, finalists=2):
    """Return the candidate with the smallest run(candidate) time; failing
    candidates (MontblancHipError) are skipped. measure(candidate), when
    given, replaces the timing of run(candidate) by time_call. final(c),
    when given, times the `finalists` best candidates again (on the full
    problem, where measure only estimates it) and decides among them."""
    if use_cache:
        if key in _MEM:
            return _MEM[key]
        disk = _load_disk()
        if key in disk:
            _MEM[key] = disk[key]['choice']
            return _MEM[key]
    times, ran = {}, []
    for c in candidates:
        try:
            t = measure(c) if measure is not None else time_call(lambda: run(c))
        except _lib.MontblancHipError:
            continue
        times[json.dumps(c)] = t
        ran.append((t, c))
    if not ran:
        raise RuntimeError('autotune %s: no candidate ran' % key)
    ran.sort(key=lambda tc: tc[0])
    best = ran[0][1]
    record = {'choice': best, 'seconds': times}
    if final is not None and len(ran) > 1:
        finals = {}
        for _, c in ran[:finalists]:
            finals[json.dumps(c)] = final(c)
        best = min(ran[:finalists], key=lambda tc: finals[json.dumps(tc[1])])[1]
        record = {'choice': best, 'seconds': times, 'final_seconds': finals}
    _MEM[key] = best
    _store_disk(key, record)
    return best


def _probe_dims(arrays, counts, ntime_probe):
    d = rime.make_dims(arrays, counts)
    d.ntime.upper_extent = min(d.ntime.local_size, ntime_probe)
    return d


# threads per workgroup of the persistent batched kernel (k_batched)
BATCH_THREADS = 512


def batched_tuning(batch, gradients=False, ntime_probe=2):
    """Launch shape (or strategy) of the batched evaluation for this batch.

    Candidates are timed at two probe lengths of the time extent
    (ntime_probe and 3*ntime_probe) and the time at the full extent is
    extrapolated linearly: the per-sample strategy pays a fixed cost per
    sample that a single short probe would weigh far above its share of
    the full run. The two best estimates are then timed on the full
    extent, which decides between them (the matrix-core strategies can
    come out within the probes' noise of each other)."""
    from . import batched as bm
    a = batch.arrays
    na = a['uvw'].shape[1]
    ntime = a['antenna1'].shape[0]
    key = 'batched|%s|na%d|nbl%d|nchan%d|nt%d|K%d|grad%d|%s' % (
        rime.suffix(a), na, a['antenna1'].shape[1], a['frequency'].shape[0], ntime,
        batch.nsamples, int(gradients), _device_name(batch.device))
    props = torch.cuda.get_device_properties(batch.device)
    cus = props.multi_processor_count
    cands = []
    nbl = a['antenna1'].shape[1]
    for wg in (cus, 2 * cus, 4 * cus):
        for cpi in (1, 2):
            # more than four (baseline, channel) items per thread of the
            # 512-thread persistent kernel spill its accumulators to scratch
            if nbl * cpi > 4 * BATCH_THREADS:
                continue
            for st in (0, 8):
                cands.append({'workgroups': wg, 'channels_per_item': cpi, 'sources_per_tile': st})
    # the matrix-core passes, each one launch over groups of samples (float,
    # one baseline per antenna pair at every time; the library rejects
    # more than 64 antennas), or every sample through the single-model
    # matrix-core kernels, as one captured graph
    if not rime.is_double(a):
        try:
            rime.baseline_map(a)
            cands.append({'strategy': 'matrix'})
        except ValueError:
            pass
    cands.append({'strategy': 'per_sample'})

    p1 = min(ntime, ntime_probe)
    p2 = min(ntime, 3 * ntime_probe)
    probes = [_probe_dims(a, batch.totals, p) for p in (p1, p2)]

    def run(c, d=probes[0]):
        bm.evaluate(batch, gradients=gradients, tune=c, dims=d)

    def measure(c):
        t1 = time_call(lambda: run(c, probes[0]))
        if p2 == p1:
            return t1 * ntime / p1
        t2 = time_call(lambda: run(c, probes[1]))
        return t1 + (t2 - t1) * (ntime - p1) / (p2 - p1)

    def final(c):
        # the two best estimates again, on the full time extent
        return time_call(lambda: bm.evaluate(batch, gradients=gradients, tune=c))
    try:
        return tune(key, cands, run, measure=measure, final=final)
    finally:
        bm.drop_probes(batch)


def forward_tuning(arrays, counts, ntime_probe=4):
    """Kernel form (pair blocks / producer waves / generic) of the fused
    forward for this shape."""
    a = arrays
    key = 'forward|%s|na%d|nbl%d|nchan%d|src%d|%s' % (
        rime.suffix(a), a['uvw'].shape[1], a['antenna1'].shape[1], a['frequency'].shape[0],
        sum(counts), _device_name(a['uvw'].device))
    cands = ['generic', [2, 2, -1], [2, 2, 0], [2, 2, 3], [2, 2, 5], [1, 1, 0]]
    d = _probe_dims(a, counts, ntime_probe)

    def run(c):
        a.pop('_pair_plan', None)
        rime.forward(a, counts, dims=d, pair_block=c)
    choice = tune(key, cands, run)
    a.pop('_pair_plan', None)
    return choice if choice == 'generic' else tuple(choice)
