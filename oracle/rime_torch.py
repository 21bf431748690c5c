"""torch float64 twin of the v4 CPU RIME, for reference gradients.

TEST INFRASTRUCTURE ONLY (tests/ and smoke() use it as the checker).

The same chain as oracle/rime_np.py (montblanc/impl/rime/v4/cpu/
CPUSolver.py:59-737), written in torch complex128 operations so that
autograd gives d chi^2 / d(lm, stokes, alpha, gauss_shape, sersic_shape,
G_term). floor() of the beam grid positions carries no gradient and
clamp() passes none outside its range, so the beam term contributes the
derivative of the trilinear weights inside each cell.

Pinned in tests/test_oracle_grad.py: its forward equals rime_np's, and its
gradient equals central finite differences of rime_np's chi^2.
"""

import math

import torch

C = 299792458.0
GAUSS_SCALE = 1.0 / math.sqrt(math.log(256.0)) * math.sqrt(2.0) * math.pi / C
TWO_PI_OVER_C = 2.0 * math.pi / C

PARAMS = ('lm', 'stokes', 'alpha', 'gauss_shape', 'sersic_shape', 'G_term')


def _jm(A, B):
    a0, a1, a2, a3 = A.unbind(-1)
    b0, b1, b2, b3 = B.unbind(-1)
    return torch.stack([a0 * b0 + a1 * b2, a0 * b1 + a1 * b3,
                        a2 * b0 + a3 * b2, a2 * b1 + a3 * b3], -1)


def _jmh(A, B):
    a0, a1, a2, a3 = A.unbind(-1)
    b0, b1, b2, b3 = B.conj().unbind(-1)
    return torch.stack([a0 * b0 + a1 * b1, a0 * b2 + a1 * b3,
                        a2 * b0 + a3 * b1, a2 * b2 + a3 * b3], -1)


def chi2(t, npsrc, ngsrc, nssrc, residual=False, use_weight_vector=False, sigma_sqrd=1.0):
    """chi^2 (and the model visibilities) from a dict of float64/complex128
    tensors in the v4 layout; differentiable in the PARAMS entries."""
    uvw, lm, f = t['uvw'], t['lm'], t['frequency']
    a1, a2 = t['antenna1'].long(), t['antenna2'].long()
    ntime, na = uvw.shape[:2]
    nbl = a1.shape[1]
    nchan = f.shape[0]
    # brightness and its square root
    S = t['stokes']
    I, Q, U, V = S.unbind(-1)
    B0 = torch.stack([I + Q, torch.complex(U, V), torch.complex(U, -V), I - Q], -1)
    power = torch.exp(t['alpha'][..., None] * torch.log(f / t['ref_frequency']))  # (s,t,c)
    B = B0[:, :, None, :] * power[..., None]
    tr = (B[..., 0] + B[..., 3]).real
    det = (B[..., 0] * B[..., 3] - B[..., 1] * B[..., 2]).real
    s_ = torch.sqrt(det)
    t_ = torch.sqrt(tr + 2 * s_)
    t_ = torch.where((s_ == 0) & (t_ == 0), torch.ones_like(t_), t_)
    eye = torch.tensor([1, 0, 0, 1], dtype=torch.complex128)
    sqB = (B + s_[..., None] * eye) / t_[..., None]
    # phase
    l, m = lm[:, 0], lm[:, 1]
    n1 = torch.sqrt(1 - l * l - m * m) - 1
    p = (l[:, None, None] * uvw[None, :, :, 0] + m[:, None, None] * uvw[None, :, :, 1]
         + n1[:, None, None] * uvw[None, :, :, 2])
    K = torch.exp(torch.complex(torch.zeros_like(p[..., None] * f),
                                -2 * math.pi / C * p[..., None] * f))      # (s,t,a,c)
    # beam
    E = t['E_beam']
    lw, mh, nud = E.shape[:3]
    pa = t['parallactic_angles']
    sn, cs = torch.sin(pa), torch.cos(pa)
    lr = l[:, None, None] * cs - m[:, None, None] * sn
    mr = l[:, None, None] * sn + m[:, None, None] * cs
    pe, sc = t['point_errors'], t['antenna_scaling']
    lr = (lr[..., None] + pe[None, ..., 0]) * sc[None, None, :, :, 0]
    mr = (mr[..., None] + pe[None, ..., 1]) * sc[None, None, :, :, 1]
    vl = torch.clamp((lw - 1) * (lr - t['beam_ll']) / (t['beam_ul'] - t['beam_ll']), 0, lw - 1)
    vm = torch.clamp((mh - 1) * (mr - t['beam_lm']) / (t['beam_um'] - t['beam_lm']), 0, mh - 1)
    vc = torch.clamp((nud - 1) * (f - t['beam_lfreq']) / (t['beam_ufreq'] - t['beam_lfreq']),
                     0, nud - 1)
    gl0, gm0, gc0 = vl.detach().floor(), vm.detach().floor(), vc.detach().floor()
    gl1 = torch.clamp(gl0 + 1, max=lw - 1)
    gm1 = torch.clamp(gm0 + 1, max=mh - 1)
    gc1 = torch.clamp(gc0 + 1, max=nud - 1)
    ld, md, cd = vl - gl0, vm - gm0, (vc - gc0)[None, None, None, :]
    absE = E.abs()
    acc = 0
    amp = 0
    for gl, wl in ((gl0, 1 - ld), (gl1, ld)):
        for gm, wm in ((gm0, 1 - md), (gm1, md)):
            for gc, wc in ((gc0, 1 - cd), (gc1, cd)):
                ic = gc.long()[None, None, None, :].expand_as(gl)
                idx = (gl.long(), gm.long(), ic)
                w = (wl * wm * wc)[..., None]
                acc = acc + w * E[idx]
                amp = amp + w * absE[idx]
    mag = acc.abs()
    ok = (mag > 0) & torch.isfinite(mag)
    norm = torch.where(ok, 1 / torch.where(ok, mag, torch.ones_like(mag)), torch.ones_like(mag))
    Eb = acc * (norm * amp)
    J = _jm(Eb, K[..., None] * sqB[:, :, None, :, :])        # (s,t,a,c,4)
    # per-baseline products and shapes
    tt = torch.arange(ntime)[:, None]
    per = _jmh(J[:, tt, a1], J[:, tt, a2])                   # (s,t,b,c,4)
    uvb = uvw[tt, a2] - uvw[tt, a1]
    u, v = uvb[..., 0], uvb[..., 1]
    fac = [torch.ones((npsrc, ntime, nbl, nchan), dtype=torch.float64)]
    if ngsrc:
        el, em, R = t['gauss_shape'][:, :, None, None]
        s = t.get('gauss_scale', GAUSS_SCALE) * f
        x = (u * em - v * el)[..., None] * s * R[..., None]
        y = (u * el + v * em)[..., None] * s
        fac.append(torch.exp(-(x * x + y * y)))
    if nssrc:
        e1, e2, Rs = t['sersic_shape'][:, :, None, None]
        k = t.get('two_pi_over_c', TWO_PI_OVER_C) * f
        Rp = (Rs / (1 - e1 * e1 - e2 * e2))[..., None]
        x = (u * (1 + e1) + v * e2)[..., None] * k * Rp
        y = (u * e2 + v * (1 - e1))[..., None] * k * Rp
        den = 1 + x * x + y * y
        fac.append(1 / (den * torch.sqrt(den)))
    fac = torch.cat(fac, 0)
    Vs = (per * fac[..., None]).sum(0)                       # (t,b,c,4)
    G = t['G_term']
    M = _jmh(_jm(G[tt, a1], Vs), G[tt, a2])
    keep = (t['flag'] == 0)
    O = t['observed_vis'] * keep
    M = M * keep
    R = O - M
    terms = (R.real ** 2 + R.imag ** 2)
    if use_weight_vector:
        terms = terms * t['weight_vector']
    x2 = terms.sum()
    if not use_weight_vector:
        x2 = x2 / sigma_sqrd
    return x2, (R if residual else M)


def to_torch(a, requires_grad=True):
    """numpy dict (v4 names) -> dict of float64 / complex128 CPU tensors;
    the PARAMS entries require grad."""
    import numpy as np
    out = {}
    for k, v in a.items():
        if isinstance(v, np.ndarray):
            if v.dtype.kind == 'f':
                v = torch.from_numpy(v.astype(np.float64))
            elif v.dtype.kind == 'c':
                v = torch.from_numpy(v.astype(np.complex128))
            else:
                v = torch.from_numpy(v)
            if requires_grad and k in PARAMS:
                v = v.clone().requires_grad_(True)
            out[k] = v
        else:
            out[k] = v
    return out


def gradients(a, npsrc, ngsrc, nssrc, **kw):
    """dict of numpy gradients of chi^2 w.r.t. PARAMS (G_term as
    d/dRe + i d/dIm, torch's convention), plus 'X2'."""
    t = to_torch(a)
    x2, _ = chi2(t, npsrc, ngsrc, nssrc, **kw)
    params = [t[k] for k in PARAMS]
    grads = torch.autograd.grad(x2, params, allow_unused=True)
    out = {k: (g.detach().numpy() if g is not None else None) for k, g in zip(PARAMS, grads)}
    out['X2'] = float(x2.detach())
    return out
