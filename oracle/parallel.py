"""The oracle's forward with the coherency sum split over source ranges in
worker processes: the multi-core CPU baseline (test and measurement
infrastructure, like the rest of the oracle; nothing on the product path
imports it).

Each worker sums its source range with rime_np.coherencies (the
restatement of CPUSolver.py:562-634); the parent adds the partial sums and
applies the epilogue once (rime_np.gekb_vis, chi_sqrd_sum_terms,
chi_sqrd). Workers are spawned interpreters: they share nothing with the
parent but the inputs.
"""

import multiprocessing as mp

import numpy as np

from . import rime_np

SOURCE_ROWS = ('lm', 'stokes', 'alpha')

_INPUTS = {}


def _init(a, counts):
    _INPUTS['a'] = a
    _INPUTS['counts'] = counts


def _partial(rng):
    """Coherency sum over the global source range [s0, s1)."""
    s0, s1 = rng
    a = _INPUTS['a']
    npsrc, ngsrc, nssrc = _INPUTS['counts']
    bounds = rime_np.source_type_slices(npsrc, ngsrc, nssrc)
    sub = dict(a)
    for k in SOURCE_ROWS:
        sub[k] = a[k][s0:s1]
    sub_counts = []
    for kind, (b, e) in enumerate(bounds):
        lo, hi = max(s0, b), min(s1, e)
        sub_counts.append(max(0, hi - lo))
        if kind == 1:
            sub['gauss_shape'] = a['gauss_shape'][:, max(lo - b, 0):max(hi - b, 0)]
        if kind == 2:
            sub['sersic_shape'] = a['sersic_shape'][:, max(lo - b, 0):max(hi - b, 0)]
    return rime_np.coherencies(sub, *sub_counts)


def source_ranges(nsrc, parts):
    """[0, nsrc) in ``parts`` contiguous ranges of near-equal size."""
    edges = np.linspace(0, nsrc, parts + 1).round().astype(int)
    return [(int(b), int(e)) for b, e in zip(edges[:-1], edges[1:]) if e > b]


class ParallelRime(object):
    """A pool of ``processes`` workers holding one set of inputs; call()
    runs the forward (rime_np.rime's outputs)."""

    def __init__(self, a, npsrc, ngsrc, nssrc, processes):
        self.a, self.counts = a, (npsrc, ngsrc, nssrc)
        self.processes = processes
        self.pool = mp.get_context('spawn').Pool(processes, initializer=_init, initargs=(a, self.counts))
        self.pool.map(_partial, [(0, 0)] * processes)     # workers up and holding the inputs

    def __call__(self, residual=False, use_weight_vector=False, sigma_sqrd=1.0):
        a = self.a
        ranges = source_ranges(sum(self.counts), self.processes)
        V = np.sum(self.pool.map(_partial, ranges), axis=0)
        vis = rime_np.gekb_vis(V, a['G_term'], a['antenna1'], a['antenna2'], a['flag'],
                               a['observed_vis'], residual)
        terms = rime_np.chi_sqrd_sum_terms(vis, a['observed_vis'], a['flag'], a['weight_vector'],
                                           residual, use_weight_vector)
        return dict(ekb_vis=V, model_vis=vis, chi_sqrd_result=terms,
                    X2=rime_np.chi_sqrd(terms, sigma_sqrd, use_weight_vector))

    def close(self):
        self.pool.close()
        self.pool.join()
