"""chi^2 as a differentiable torch function with a hand-written adjoint.

``chi2(arrays, counts, ...)`` returns X2 as a 0-d float64 tensor; its
backward runs the library's adjoint (recomputing the Jones terms; only the
forward's pre-gain source sums are kept between the passes) and gives
gradients for whichever of lm, stokes, alpha, gauss_shape, sersic_shape and
G_term require them. No reference counterpart: the reference computes
chi^2 only (RimeSumCoherencies.py:531-536).
"""

import torch

from . import rime

PARAMS = ('lm', 'stokes', 'alpha', 'gauss_shape', 'sersic_shape', 'G_term')


class RimeChi2(torch.autograd.Function):
    @staticmethod
    def forward(ctx, lm, stokes, alpha, gauss_shape, sersic_shape, G_term, arrays, counts,
                use_weight_vector, sigma_sqrd):
        arr = dict(arrays)
        for name, t in zip(PARAMS, (lm, stokes, alpha, gauss_shape, sersic_shape, G_term)):
            arr[name] = t.detach().contiguous()
        out = rime.forward(arr, counts, use_weight_vector=use_weight_vector,
                           sigma_sqrd=sigma_sqrd, keep_ekb=True)
        ctx.arrays = arr
        ctx.counts = counts
        ctx.opts = (use_weight_vector, sigma_sqrd)
        ctx.ekb = out['ekb_vis']
        ctx.dtypes = [t.dtype for t in (lm, stokes, alpha, gauss_shape, sersic_shape, G_term)]
        ctx.model_vis = out['model_vis']
        return out['X2'].clone()

    @staticmethod
    def backward(ctx, grad_x2):
        use_w, s2 = ctx.opts
        g = rime.adjoint(ctx.arrays, ctx.counts, ctx.ekb, grad_scale=float(grad_x2.item()),
                         use_weight_vector=use_w, sigma_sqrd=s2)
        outs = []
        for name, dt, need in zip(PARAMS, ctx.dtypes, ctx.needs_input_grad[:6]):
            outs.append(g[name].to(dt) if need else None)
        return tuple(outs) + (None, None, None, None)


def chi2(arrays, counts, use_weight_vector=False, sigma_sqrd=None):
    """X2 (0-d float64 tensor) of the arrays dict; differentiable in the
    parameter tensors of the dict that require grad.

    Raises ValueError before the forward when gradients are wanted for an
    antenna layout the adjoint does not take (pairs changing with time, or
    one pair on two baselines)."""
    if torch.is_grad_enabled() and any(
            isinstance(arrays.get(n), torch.Tensor) and arrays[n].requires_grad for n in PARAMS):
        rime.baseline_map(arrays)
    return RimeChi2.apply(*(arrays[n] for n in PARAMS), arrays, tuple(counts),
                          use_weight_vector, sigma_sqrd)
