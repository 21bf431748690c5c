"""Derived error bound for float32 gradients of chi^2 (test helper).

chi^2 = sum_v w_v |R_v|^2 / s with R_v = O_v - M_v over unflagged
visibilities v (s = sigma^2 without weights, 1 with them), so for a real
parameter theta

    g_theta = -(2 / s) sum_v w_v Re< R_v, dM_v/dtheta >.

The float32 path (float64 phase and beam coordinates) has two error
sources, each bounded by Cauchy-Schwarz over the visibilities:

* its model M~ = M + dM enters the residual: measured here as ||dM||_2
  from the float32 forward's own output (or bounded by tol ||M||_2 where
  a path does not return its model), plus 8 u ||M||_2 (u = 2^-24) for
  the adjoint recomputing M from the pre-gain sums in its own rounding;
* it forms each dM_v/dtheta by the float32 Jones recomputation, taken to
  carry a relative error tol = 1e-5 (the bound the float32 forward is held
  to), which multiplies the residual.

First order in both:

    |g~_theta - g_theta| <= (2 w_max / s) (||dM||_2 + 8 u ||M||_2
                                             + tol ||R||_2) ||dM/dtheta||_2

plus a floor of 2 u max|g| per array. The same bound with u = 2^-53 and
tol = 1e-12 holds the float64 path where the residual is much smaller
than the model (near a fit the model's own float64 rounding, dominated by
the phase's |phi| 2^-53, is amplified by |M| / |R| in the gradient). ||dM/dtheta||_2 (the column of the
model's Jacobian) comes from the torch float64 oracle by forward-mode
derivatives: one source at a time for the source parameters (a source
parameter moves only that source's contribution), and through the
oracle's pre-gain sums for the gains (M = G_p V G_q^H). Complex gains
combine the bounds of their real and imaginary directions.

Each array is checked on the elements with the largest errors and a
random sample of the rest (the Jacobian columns cost one derivative pass
per element).
"""

import numpy as np
import torch

from oracle import rime_np, rime_torch

SOURCE_PARAMS = ('lm', 'stokes', 'alpha')
TOL = 1e-5
U32 = 2.0 ** -24


def _column_norms(fn, x, flat, chunk=32):
    """||d fn / d x_i||_2 for the flat indices i of x; for a complex x,
    (real direction, imaginary direction)."""
    n = x.numel()
    out = []
    for imag in ((False, True) if x.is_complex() else (False,)):
        res = np.empty(len(flat))
        for s in range(0, len(flat), chunk):
            idx = torch.as_tensor(flat[s:s + chunk])
            k = len(idx)
            e = torch.zeros((k, n), dtype=torch.float64)
            e[torch.arange(k), idx] = 1.0
            e = e.reshape((k,) + tuple(x.shape))
            if x.is_complex():
                z = torch.zeros_like(e)
                e = torch.complex(z, e) if imag else torch.complex(e, z)
            jv = torch.func.vmap(lambda v: torch.func.jvp(fn, (x,), (v,))[1])(e)
            res[s:s + k] = torch.linalg.vector_norm(jv.reshape(k, -1), dim=1).numpy()
        out.append(res)
    return out


def column_norms(a, counts, select):
    """||dM/dtheta||_2 for the selected flat elements of each parameter
    array: select maps a name to flat indices; returns name -> norms (for
    G_term, real and imaginary directions combined by hypot)."""
    t = rime_torch.to_torch(a, requires_grad=False)
    npsrc, ngsrc, nssrc = counts
    out = {}
    for n in SOURCE_PARAMS + ('gauss_shape', 'sersic_shape'):
        if n not in select:
            continue
        flat = np.asarray(select[n])
        shape = tuple(t[n].shape)
        res = np.empty(len(flat))
        if n in SOURCE_PARAMS:
            src = np.unravel_index(flat, shape)[0]
        else:
            src = np.unravel_index(flat, shape)[1] + (npsrc if n == 'gauss_shape' else npsrc + ngsrc)
        for s in np.unique(src):
            kind = 0 if s < npsrc else 1 if s < npsrc + ngsrc else 2
            one = dict(t)
            for p in SOURCE_PARAMS:
                one[p] = t[p][s:s + 1]
            c = [0, 0, 0]
            c[kind] = 1
            shape_name = (None, 'gauss_shape', 'sersic_shape')[kind]
            col = None
            if shape_name:
                col = s - npsrc - (ngsrc if kind == 2 else 0)
                one[shape_name] = t[shape_name][:, col:col + 1]
            mine = np.flatnonzero(src == s)
            if n in SOURCE_PARAMS:
                local = flat[mine] - s * int(np.prod(shape[1:]))
            else:
                local = np.unravel_index(flat[mine], shape)[0]

            def fn(x, _n=n):
                d = dict(one)
                d[_n] = x
                return rime_torch.chi2(d, *c)[1]
            res[mine] = _column_norms(fn, one[n], local)[0]
        out[n] = res
    if 'G_term' in select:
        V = torch.from_numpy(rime_np.coherencies(a, npsrc, ngsrc, nssrc))
        a1 = t['antenna1'].long()
        a2 = t['antenna2'].long()
        tt = torch.arange(a1.shape[0])[:, None]
        keep = t['flag'] == 0

        def gains(G):
            return rime_torch._jmh(rime_torch._jm(G[tt, a1], V), G[tt, a2]) * keep
        re, im = _column_norms(gains, t['G_term'], np.asarray(select['G_term']), chunk=64)
        out['G_term'] = np.hypot(re, im)
    return out


U64 = 2.0 ** -53
TOL64 = 1e-12


def check_f32_gradients(a, counts, got, ref, model_f32, use_weight_vector=False,
                        sigma_sqrd=1.0, names=None, top=12, sample=12, seed=0, tol=TOL, u=U32):
    """Assert the float32 gradients ``got`` are within the derived bound of
    the float64 reference ``ref`` (rime_torch.gradients on the same
    float32-rounded inputs ``a``). model_f32: the float32 forward's model
    visibilities, or None where the path does not expose them (then
    ||dM||_2 is taken as tol ||M||_2, the forward's own bound). Returns
    the largest error / bound ratio per array."""
    keep = a['flag'] == 0
    M = rime_np.rime(a, *counts)['model_vis'] * keep
    R = (a['observed_vis'] - M) * keep
    if model_f32 is None:
        dM_norm = tol * np.linalg.norm(M)
    else:
        dM_norm = np.linalg.norm((np.asarray(model_f32) - M) * keep)
    wmax = float(a['weight_vector'][keep].max()) if use_weight_vector else 1.0
    s = 1.0 if use_weight_vector else sigma_sqrd
    scale = 2.0 * wmax / s * (dM_norm + 8 * u * np.linalg.norm(M) + tol * np.linalg.norm(R))
    rng = np.random.default_rng(seed)
    select, errs = {}, {}
    for n in SOURCE_PARAMS + ('gauss_shape', 'sersic_shape', 'G_term'):
        if (names is not None and n not in names) or ref.get(n) is None or ref[n].size == 0:
            continue
        err = np.abs(np.asarray(got[n]) - ref[n]).reshape(-1)
        order = np.argsort(err)[::-1]
        pick = list(order[:top])
        rest = order[top:]
        if len(rest):
            pick += list(rng.choice(rest, min(sample, len(rest)), replace=False))
        select[n] = np.asarray(pick)
        errs[n] = err
    norms = column_norms(a, counts, select)
    ratios = {}
    for n, idx in select.items():
        bound = scale * norms[n] + 2 * u * np.abs(ref[n]).max()
        e = errs[n][idx]
        r = np.full(e.shape, np.inf)
        np.divide(e, bound, out=r, where=bound > 0)
        r[e == 0] = 0.0
        ratios[n] = float(r.max())
        assert ratios[n] <= 1.0, (n, ratios[n], float(errs[n].max()), float(bound.min()))
    return ratios
