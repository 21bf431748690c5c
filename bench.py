#!/usr/bin/env python
"""Throughput of the fused RIME forward on one MI355X.

Workload (default): 64 antennas (2016 baselines), 100 times, 64 channels at
1.40-1.42 GHz, 1000 point + 500 gaussian + 100 sersic sources, beam cube
65x65x32, synthetic inputs (oracle/synth.py, seed 1001). A step is one
fused forward (model visibilities, chi^2 terms and X2) over the whole
problem with every input resident in HBM. Metric: source-visibility terms
per second (one term = one source, time, baseline and channel with all four
correlations).

Both precisions are timed; ``value`` is the float32 run (float64 phase
argument and beam coordinates), the float64 run is reported beside it.
The roofline object is for the forward launch (its matrix-core and
vector kernels): algorithmic FLOPs per launch over its HIP-event-timed
duration, against the dtype's dense MI355X peak. cpu_baseline times the
numpy restatement of the reference's CPU chain on a bounded slice of the
same workload in one process; cpu_baseline_parallel the same slice with
the coherency sum split over source ranges in one process per core (up to
16).

Default legs: the forward workload in float32 and float64, chi^2 + gradient
(27 antennas in float64; the forward workload in float32), the 256-sample
batched evaluation, the large single-model workload (64 antennas, 500
times, 256 channels, 12000 sources: forward and forward + gradient, with
the error against the float64 oracle on a slice), the attainable library
GEMM rates and the CPU baselines.

Usage: python bench.py [--gpus 1] [--steps 20] [--warmup 3]
"""

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument('--gpus', type=int, default=1)
    p.add_argument('--steps', type=int, default=20)
    p.add_argument('--warmup', type=int, default=3)
    p.add_argument('--workload', default='forward_64ant')
    p.add_argument('--dtypes', default='f32,f64')
    p.add_argument('--no-cpu-baseline', action='store_true')
    p.add_argument('--no-gradient', action='store_true')
    p.add_argument('--no-large', action='store_true',
                   help='skip the large single-model workload (about 40 s)')
    p.add_argument('--no-batched', action='store_true')
    p.add_argument('--calibrate-copy', type=int, default=0, metavar='MIB',
                   help='first run a plain copy kernel over MIB MiB (reads and writes '
                        'MIB MiB), the reference point for HBM byte counters')
    return p.parse_args()


def cpu_model():
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or 'unknown'


def measured_traffic(workload, dtype):
    """HBM bytes per forward from the committed rocprofv3 PMC summary of this
    workload (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel), or None."""
    import re
    path = os.path.join(ROOT, 'profiles', '%s_%s_hbm_traffic.md' % (workload, dtype))
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r'traffic_bytes_per_forward:\s*([0-9.eE+]+)', text)
    return float(m.group(1)) if m else None


def copy_calibration(mib):
    """One elementwise copy (dst = src * 1) of ``mib`` MiB of float32: a
    kernel of known traffic (mib MiB read, mib MiB written) whose byte
    counters scale the fused kernel's."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device='cuda')
    dst = torch.empty_like(src)
    torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print('# copy calibration: %d MiB read, %d MiB written (torch.mul)' % (mib, mib),
          file=sys.stderr, flush=True)
    del src, dst
    torch.cuda.empty_cache()


def time_forward(w, dtype, steps, warmup):
    import numpy as np
    import torch
    from montblanc_amd import rime, roofline
    from oracle import synth

    counts = (w['npsrc'], w['ngsrc'], w['nssrc'])
    a = synth.make_inputs(**w, seed=1001)
    arrays = rime.to_device(synth.as_dtype(a, np.float32 if dtype == 'f32' else np.float64))
    del a
    out = rime.alloc_outputs(arrays)
    # observed = model + CN(0, 0.1), the model from this GPU path (untimed)
    rime.forward(arrays, counts, outputs=out)
    g = torch.Generator(device='cuda').manual_seed(7)
    noise = torch.randn(out['model_vis'].shape + (2,), generator=g, device='cuda',
                        dtype=arrays['uvw'].dtype) * 0.1
    arrays['observed_vis'] = (out['model_vis'] + torch.view_as_complex(noise)).contiguous()
    del noise
    for _ in range(warmup):
        rime.forward(arrays, counts, outputs=out)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        starts[i].record(stream)
        rime.forward(arrays, counts, outputs=out)
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev = sorted(s.elapsed_time(e) for s, e in zip(starts, ends))
    x2 = float(out['X2'].item())
    nbl = w['na'] * (w['na'] - 1) // 2
    n_terms = roofline.terms(w['na'], w['ntime'], w['nchan'], sum(counts), nbl)
    flops = roofline.forward_flops(w['na'], w['ntime'], w['nchan'], *counts, nbl=nbl)
    ms_mean = sum(ev) / len(ev)
    ms_median = ev[len(ev) // 2]
    del arrays, out
    torch.cuda.empty_cache()
    return dict(ms_per_step=wall * 1e3 / steps, ms_event_mean=ms_mean, ms_event_median=ms_median,
                terms=n_terms, flops=flops, X2=x2,
                terms_per_s=n_terms / (wall / steps),
                terms_per_s_median=n_terms / (ms_median * 1e-3),
                tflops=flops / (ms_mean * 1e-3) / 1e12)


def time_gradient(w, dtype, steps, warmup, name='gradient_27ant'):
    """chi^2 alone vs chi^2 + full gradient (forward keeping the pre-gain
    sums, then the adjoint) on workload w."""
    import numpy as np
    import torch
    from montblanc_amd import rime
    from oracle import synth

    counts = (w['npsrc'], w['ngsrc'], w['nssrc'])
    a = synth.make_inputs(**w, seed=1002)
    arrays = rime.to_device(synth.as_dtype(a, np.float32 if dtype == 'f32' else np.float64))
    del a
    out = rime.alloc_outputs(arrays, keep_ekb=True)
    rime.forward(arrays, counts, outputs=out)
    g = torch.Generator(device='cuda').manual_seed(8)
    noise = torch.randn(out['model_vis'].shape + (2,), generator=g, device='cuda',
                        dtype=arrays['uvw'].dtype) * 0.1
    arrays['observed_vis'] = (out['model_vis'] + torch.view_as_complex(noise)).contiguous()
    del noise
    grads = rime.alloc_grads(arrays, counts)

    def fwd():
        rime.forward(arrays, counts, outputs=out)

    def fwd_grad():
        rime.forward(arrays, counts, outputs=out)
        rime.adjoint(arrays, counts, out['ekb_vis'], grads=grads)

    res, med = {}, {}
    stream = torch.cuda.current_stream()
    for leg, fn in (('forward', fwd), ('forward_gradient', fwd_grad)):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(steps)]
        t0 = time.perf_counter()
        for e0, e1 in ev:
            e0.record(stream)
            fn()
            e1.record(stream)
        torch.cuda.synchronize()
        res[leg] = (time.perf_counter() - t0) * 1e3 / steps
        med[leg] = sorted(e0.elapsed_time(e1) for e0, e1 in ev)[steps // 2]
    nbl = w['na'] * (w['na'] - 1) // 2
    n_terms = sum(counts) * w['ntime'] * nbl * w['nchan']
    del arrays, out, grads
    torch.cuda.empty_cache()
    return {'workload': name, 'dtype': dtype, 'terms': n_terms, 'steps': steps,
            'forward_ms': res['forward'], 'forward_gradient_ms': res['forward_gradient'],
            'forward_ms_median': med['forward'],
            'forward_gradient_ms_median': med['forward_gradient'],
            'forward_terms_per_s': n_terms / (res['forward'] * 1e-3),
            'gradient_terms_per_s': n_terms / (res['forward_gradient'] * 1e-3),
            'gradient_over_forward': res['forward'] / res['forward_gradient'],
            'gradient_over_forward_median': med['forward'] / med['forward_gradient']}


def time_large(steps=1):
    """Large single model: 64 antennas, 500 times, 256 channels, 10000
    point + 2000 gaussian sources, beam 129x129x64, float32 with float64
    phase. Forward and forward + full gradient, resident in HBM (the
    planner's budget for one device needs no chunking); the float32 model
    visibilities are compared with the float64 numpy oracle on a slice
    (time 0, two channels, every baseline and source)."""
    import numpy as np
    import torch
    from montblanc_amd import planner, rime
    from oracle import rime_np, synth

    w = dict(synth.WORKLOADS['large_64ant'])
    counts = (w['npsrc'], w['ngsrc'], w['nssrc'])
    a = synth.make_inputs(**w, seed=1004, vis_arrays=False)
    arrays = rime.to_device(synth.as_dtype(a, np.float32))
    nt, nbl, nc = w['ntime'], w['na'] * (w['na'] - 1) // 2, w['nchan']
    g = torch.Generator(device='cuda').manual_seed(4)
    arrays['observed_vis'] = torch.view_as_complex(
        0.1 * torch.randn((nt, nbl, nc, 4, 2), generator=g, device='cuda')).contiguous()
    arrays['weight_vector'] = torch.ones((nt, nbl, nc, 4), device='cuda')
    arrays['flag'] = torch.zeros((nt, nbl, nc, 4), dtype=torch.uint8, device='cuda')
    budget = torch.cuda.mem_get_info()[0]
    dims = dict(ntime=nt, na=w['na'], nbl=nbl, nchan=nc, npol=4, npsrc=counts[0],
                ngsrc=counts[1], nssrc=0, nsrc=sum(counts), beam_lw=w['beam'][0],
                beam_mh=w['beam'][1], beam_nud=w['beam'][2])
    mod, _, need = planner.plan(dims, np.float32, budget - planner.DEVICE_RESERVE, gradients=True)
    out = rime.alloc_outputs(arrays, keep_ekb=True)
    grads = rime.alloc_grads(arrays, counts)
    # warm-up on one time step: launch plans, baseline map, workspaces
    probe = rime.make_dims(arrays, counts)
    probe.ntime.upper_extent = 1
    rime.forward(arrays, counts, outputs=out, dims=probe)
    rime.adjoint(arrays, counts, out['ekb_vis'], grads=grads, dims=probe)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        rime.forward(arrays, counts, outputs=out)
    torch.cuda.synchronize()
    t_fwd = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        rime.forward(arrays, counts, outputs=out)
        rime.adjoint(arrays, counts, out['ekb_vis'], grads=grads)
    torch.cuda.synchronize()
    t_grad = (time.perf_counter() - t0) / steps
    # error against the float64 oracle on a slice
    up = synth.upcast(synth.as_dtype(a, np.float32))
    sl = dict(up)
    sl['uvw'] = up['uvw'][:1]
    sl['parallactic_angles'] = up['parallactic_angles'][:1]
    sl['point_errors'] = up['point_errors'][:1, :, :2]
    sl['antenna_scaling'] = up['antenna_scaling'][:, :2]
    sl['frequency'], sl['ref_frequency'] = up['frequency'][:2], up['ref_frequency'][:2]
    sl['stokes'], sl['alpha'] = up['stokes'][:, :1], up['alpha'][:, :1]
    sl['antenna1'], sl['antenna2'] = up['antenna1'][:1], up['antenna2'][:1]
    sl['G_term'] = up['G_term'][:1, :, :2]
    V = rime_np.coherencies(sl, *counts, src_chunk=256)
    M = rime_np.gekb_vis(V, sl['G_term'], sl['antenna1'], sl['antenna2'],
                         np.zeros(V.shape, np.uint8), np.zeros(V.shape, np.complex128), False)
    got = out['model_vis'][:1, :, :2].cpu().numpy()
    err = float(np.abs(got - M).max() / np.abs(M).max())
    n_terms = sum(counts) * nt * nbl * nc
    del arrays, out, grads
    torch.cuda.empty_cache()
    return {'workload': 'large_64ant', 'dtype': 'f32', 'terms': n_terms,
            'planner_reductions': dict(mod), 'device_bytes_planned': need,
            'forward_s': t_fwd, 'forward_gradient_s': t_grad,
            'forward_terms_per_s': n_terms / t_fwd, 'gradient_terms_per_s': n_terms / t_grad,
            'gradient_over_forward': t_fwd / t_grad,
            'max_rel_error_vs_f64_oracle': err,
            'error_slice': 'time 0, channels 0-1, all 2016 baselines and 12000 sources'}


def time_batched(steps=2):
    """Batched workload: 256 sky-model samples, source counts log-uniform
    in [10, 2000] (30% gaussian), against one shared observation of 64
    antennas, 50 times, 32 channels; float32 with float64 phase. chi^2 of
    every sample, then chi^2 + gradient of every sample. The autotuner
    (run in the untimed warm-up) picks either one persistent launch over
    all samples with tuned tile shapes, or the per-sample strategy (each
    sample through the single-model matrix-core kernels on four streams,
    the sequence replayed as one HIP graph); the choice is reported."""
    import numpy as np
    import torch
    from montblanc_amd import autotune, batched
    from oracle import synth

    w = dict(synth.WORKLOADS['batched_64ant'])
    rng = np.random.default_rng(1003)
    counts = batched.sample_counts(rng, 256)
    shared = synth.make_inputs(**w, seed=1003)
    nt = w['ntime']
    samples = []
    for c in counts:
        n = sum(c)
        I = rng.lognormal(0.0, 1.0, n)
        st = np.empty((n, nt, 4))
        st[..., 0] = I[:, None]
        st[..., 1:] = (I[:, None] * rng.uniform(-0.1, 0.1, (n, 3)))[:, None, :]
        samples.append(dict(lm=rng.uniform(-0.01, 0.01, (n, 2)), stokes=st,
                            alpha=np.repeat(rng.normal(-0.7, 0.2, n)[:, None], nt, axis=1),
                            gauss_shape=rng.uniform(-1e-4, 1e-4, (3, c[1])),
                            sersic_shape=np.zeros((3, 0)), counts=c))
    b = batched.Batch(shared, samples, dtype=np.float32)
    nbl = w['na'] * (w['na'] - 1) // 2
    n_terms = sum(sum(c) for c in counts) * nt * nbl * w['nchan']
    out = {}
    for label, grads in (('chi2', False), ('chi2_gradient', True)):
        choice = autotune.batched_tuning(b, gradients=grads)
        batched.evaluate(b, gradients=grads)            # warms up (graph capture)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            batched.evaluate(b, gradients=grads)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        out[label] = {'s_per_step': dt, 'terms_per_s': n_terms / dt, 'tuning': choice}
        print('# batched %s: %.1f ms, %.4g terms/s, %s' % (label, dt * 1e3, n_terms / dt, choice),
              file=sys.stderr, flush=True)
    out.update(workload='batched_64ant', samples=len(counts), sources=sum(sum(c) for c in counts),
               terms=n_terms, dtype='f32',
               gradient_over_chi2=out['chi2']['s_per_step'] / out['chi2_gradient']['s_per_step'])
    del b
    torch.cuda.empty_cache()
    return out


def attainable_peaks(reps=5):
    """Dense matrix-core rates a vendor library reaches on this device:
    torch.mm (hipBLASLt / rocBLAS) of square float32 (8192) and float64
    (4096) matrices, best of ``reps`` HIP-event-timed calls. A reference
    point beside the datasheet peaks the roofline is priced against."""
    import torch
    out = {}
    for name, dt, n in (('f32', torch.float32, 8192), ('f64', torch.float64, 4096)):
        a = torch.randn(n, n, dtype=dt, device='cuda')
        b = torch.randn(n, n, dtype=dt, device='cuda')
        c = torch.mm(a, b)
        best = float('inf')
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            torch.mm(a, b, out=c)
            e1.record()
            e1.synchronize()
            best = min(best, e0.elapsed_time(e1) * 1e-3)
        out[name + '_gemm_tflops'] = 2.0 * n ** 3 / best / 1e12
        del a, b, c
    out['note'] = 'torch.mm of square matrices (f32 8192, f64 4096), best of %d' % reps
    return out


def cpu_baseline(w):
    """numpy restatement of the reference CPU chain on four time steps and
    16 channels, all sources and baselines (~10-30 s on one core), then the
    same slice over one worker process per core (up to 16). Returns the
    single-process and the multi-process baseline."""
    import numpy as np
    from oracle import parallel, rime_np, synth
    sample = {k: v for k, v in w.items() if k != 'name'}
    sample.update(ntime=4, nchan=16)
    a = synth.make_inputs(**sample, seed=1001)
    counts = (w['npsrc'], w['ngsrc'], w['nssrc'])
    t0 = time.perf_counter()
    ref = rime_np.rime(a, *counts, src_chunk=64)
    dt = time.perf_counter() - t0
    nbl = w['na'] * (w['na'] - 1) // 2
    n = sum(counts) * 4 * nbl * 16
    what = '%s with ntime=4, nchan=16 (all %d sources, %d baselines): %.3g terms' % (
        w.get('name', 'workload'), sum(counts), nbl, n)
    single = dict(value=n / dt, unit='source-vis terms/s', cores=1, kind='port',
                  sample='%s in %.1f s, numpy float64, one process; host %s, os.cpu_count()=%d'
                         % (what, dt, cpu_model(), os.cpu_count() or 0))
    procs = max(1, min(16, os.cpu_count() or 1))
    pr = parallel.ParallelRime(a, *counts, processes=procs)
    try:
        t0 = time.perf_counter()
        out = pr()
        dtp = time.perf_counter() - t0
    finally:
        pr.close()
    diff = float(np.abs(out['model_vis'] - ref['model_vis']).max() / np.abs(ref['model_vis']).max())
    multi = dict(value=n / dtp, unit='source-vis terms/s', cores=procs, kind='port',
                 sample='%s in %.2f s, numpy float64, %d worker processes (source ranges; pool '
                        'started before timing); model vs one process: max rel diff %.1e'
                        % (what, dtp, procs, diff))
    return single, multi


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py: this build measures one GPU only (the hot path has no '
              'multi-GPU collectives); run with --gpus 1', file=sys.stderr)
        sys.exit(2)
    import torch
    from montblanc_amd import roofline
    from oracle import synth

    if args.calibrate_copy:
        copy_calibration(args.calibrate_copy)
    w = dict(synth.WORKLOADS[args.workload])
    w['name'] = args.workload
    wl = {k: v for k, v in w.items() if k != 'name'}
    results = {}
    for dt in args.dtypes.split(','):
        results[dt] = time_forward(wl, dt, args.steps, args.warmup)
        print('# %s: %.3f ms/step (events mean %.3f), %.4g terms/s, %.2f TFLOP/s'
              % (dt, results[dt]['ms_per_step'], results[dt]['ms_event_mean'],
                 results[dt]['terms_per_s'], results[dt]['tflops']), file=sys.stderr, flush=True)
    primary = 'f32' if 'f32' in results else next(iter(results))
    r = results[primary]
    peak = roofline.PEAK_TFLOPS[primary]
    line = {
        'metric': 'source-vis terms/s, fused RIME forward (model vis + chi^2), 1xMI355X',
        'value': r['terms_per_s'],
        'unit': 'source-vis terms/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': r['ms_per_step'],
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': primary,
        'data': 'synthetic (oracle/synth.py seed 1001: uvw U(+-3000 m), lm U(+-0.01), '
                'lognormal Stokes, random beam cube)',
        'config': {'workload': args.workload, 'na': w['na'], 'nbl': w['na'] * (w['na'] - 1) // 2,
                   'ntime': w['ntime'], 'nchan': w['nchan'], 'npsrc': w['npsrc'],
                   'ngsrc': w['ngsrc'], 'nssrc': w['nssrc'], 'beam': list(w['beam']),
                   'phase_argument': 'f64'},
        'roofline': {'bound': 'mfma', 'achieved': r['tflops'], 'peak': peak,
                     'unit': 'TFLOP/s', 'frac': r['tflops'] / peak,
                     'traffic': measured_traffic(args.workload, primary),
                     'kernel': 'mbx_forward_%s (k_beam_planes; k_forward_mx: point and gaussian/sersic '
                               'passes and the epilogue in one kernel; k_chi2_final)' % primary,
                     'flops_per_launch': r['flops'],
                     'peak_note': 'compute-bound: point sources on the matrix cores (16x16x4 '
                                  'products), gaussian and sersic sources on the matrix cores '
                                  '(4x4x1 blocks, shape factor on the B operand); priced against '
                                  'the dense peak of the dtype (FP32 157.3 TFLOP/s on the matrix '
                                  'cores, the same figure as the packed-FP32 vector peak; this '
                                  'build compiles packed FP32 out, leaving a 78.6 TFLOP/s vector '
                                  'rate)',
                     'limiters': 'k_forward_mx: matrix cores busy 46% of SIMD cycles, other vector '
                                 'instructions 43% (rocprofv3 PMC, forward_64ant f32): 89% together, '
                                 'matrix and vector issue taking turns on each SIMD; texture path '
                                 '37% (profiles/forward_64ant_f32_counters.md)'},
        'timing_note': 'value and ms_per_step: wall clock over the timed steps with the device '
                       'synchronised on both sides; ms_event_median / terms_per_s_median: the '
                       'median of per-step HIP events (the reference protocol: warm-up, then the '
                       'median of the timed steps)',
        'per_dtype': {k: {'terms_per_s': v['terms_per_s'], 'ms_per_step': v['ms_per_step'],
                          'ms_event_mean': v['ms_event_mean'],
                          'ms_event_median': v['ms_event_median'],
                          'terms_per_s_median': v['terms_per_s_median'], 'tflops': v['tflops'],
                          'frac_of_peak': v['tflops'] / roofline.PEAK_TFLOPS[k], 'X2': v['X2']}
                      for k, v in results.items()},
    }
    if not args.no_gradient:
        gw = {k: v for k, v in synth.WORKLOADS['gradient_27ant'].items()}
        grad = time_gradient(gw, 'f64', max(3, args.steps // 2), 1)
        print('# gradient f64: forward %.2f ms, forward+gradient %.2f ms, ratio %.3f'
              % (grad['forward_ms'], grad['forward_gradient_ms'], grad['gradient_over_forward']),
              file=sys.stderr, flush=True)
        line['gradient'] = grad
        # float32 (float64 phase) on the forward workload itself
        g32 = time_gradient(wl, 'f32', max(3, args.steps // 2), 1, name=args.workload)
        print('# gradient f32 (%s): forward %.2f ms, forward+gradient %.2f ms, ratio %.3f'
              % (args.workload, g32['forward_ms'], g32['forward_gradient_ms'], g32['gradient_over_forward']),
              file=sys.stderr, flush=True)
        line['gradient_f32'] = g32
    if not args.no_batched:
        line['batched'] = time_batched()
    if not args.no_large:
        line['large'] = time_large()
        print('# large: %s' % line['large'], file=sys.stderr, flush=True)
    att = attainable_peaks()
    if primary + '_gemm_tflops' in att:
        att['frac_of_' + primary + '_gemm'] = r['tflops'] / att[primary + '_gemm_tflops']
    line['attainable'] = att
    if not args.no_cpu_baseline:
        line['cpu_baseline'], line['cpu_baseline_parallel'] = cpu_baseline(w)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
