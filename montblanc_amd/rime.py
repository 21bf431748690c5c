"""Functional tensor-level API over the C-ABI (no solver object).

``forward(arrays, counts, ...)`` runs the fused forward on device tensors
held in a dict keyed by the v4 array names; ``forward_numpy`` is the same
from/to numpy. The stage functions (``b_sqrt``, ``e_beam``, ``ekb_sqrt``,
``sum_coherencies``, ``chi2_reduce``) run one v4 kernel node each.
Every call goes through libmontblanc_amd.so; there is no CPU path.
"""

import ctypes

import numpy as np
import torch

from . import _lib

C = 299792458.0
GAUSS_SCALE = 1.0 / np.sqrt(np.log(256.0)) * np.sqrt(2.0) * np.pi / C
TWO_PI_OVER_C = 2.0 * np.pi / C

BEAM_KEYS = ('beam_ll', 'beam_lm', 'beam_lfreq', 'beam_ul', 'beam_um', 'beam_ufreq')
FLOAT_ARRAYS = ('uvw', 'lm', 'stokes', 'alpha', 'frequency', 'ref_frequency',
                'parallactic_angles', 'point_errors', 'antenna_scaling', 'gauss_shape',
                'sersic_shape', 'weight_vector')
COMPLEX_ARRAYS = ('E_beam', 'observed_vis', 'G_term')
INT_ARRAYS = ('antenna1', 'antenna2')


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() else None


def _raw(t):
    return t.data_ptr() if t is not None and t.numel() else None


def is_double(arrays):
    return arrays['uvw'].dtype == torch.float64


def suffix(arrays):
    return 'f64' if is_double(arrays) else 'f32'


def complex_dtype(arrays):
    return torch.complex128 if is_double(arrays) else torch.complex64


def to_device(a, device='cuda:0', dtype=None):
    """numpy dict (v4 names) -> dict of contiguous device tensors.

    dtype: np.float32 / np.float64 for the float arrays (complex follow);
    None keeps each array's own dtype. Scalars (beam extents, constants)
    are kept as Python floats.
    """
    out = {}
    ft = None if dtype is None else np.dtype(dtype)
    ct = None if ft is None else (np.complex64 if ft == np.float32 else np.complex128)
    for k, v in a.items():
        if isinstance(v, np.ndarray):
            if v.dtype.kind == 'f' and ft is not None:
                v = v.astype(ft)
            elif v.dtype.kind == 'c' and ct is not None:
                v = v.astype(ct)
            elif k in INT_ARRAYS:
                v = v.astype(np.int32)
            elif k == 'flag':
                v = v.astype(np.uint8)
            out[k] = torch.from_numpy(np.ascontiguousarray(v)).to(device)
        elif isinstance(v, torch.Tensor):
            out[k] = v.to(device).contiguous()
        else:
            out[k] = v
    return out


def make_dims(arrays, counts, extents=None, global_sizes=None):
    """mbx_dims from array shapes; counts = (npsrc, ngsrc, nssrc)."""
    ntime, na = arrays['uvw'].shape[:2]
    nbl = arrays['antenna1'].shape[1]
    nchan = arrays['frequency'].shape[0]
    lw, mh, nud = arrays['E_beam'].shape[:3]
    npsrc, ngsrc, nssrc = counts
    sizes = dict(ntime=ntime, na=na, nbl=nbl, nchan=nchan, npol=4, npolchan=4 * nchan,
                 nsrc=npsrc + ngsrc + nssrc, npsrc=npsrc, ngsrc=ngsrc, nssrc=nssrc,
                 beam_lw=lw, beam_mh=mh, beam_nud=nud)
    return _lib.make_dims(sizes, extents, global_sizes)


def beam_extents(arrays):
    return _lib.mbx_beam_extents(*(float(arrays[k]) for k in BEAM_KEYS))


def consts(arrays, sigma_sqrd=None):
    return _lib.mbx_consts(float(arrays.get('gauss_scale', GAUSS_SCALE)),
                           float(arrays.get('two_pi_over_c', TWO_PI_OVER_C)),
                           float(arrays.get('sigma_sqrd', 1.0) if sigma_sqrd is None else sigma_sqrd))


def options(residual=False, use_weight_vector=False):
    return (_lib.RESIDUAL_OUTPUT if residual else 0) | (_lib.APPLY_WEIGHTS if use_weight_vector else 0)


def stream_of(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


class Workspace(object):
    """Grow-only device byte buffers keyed by purpose."""

    def __init__(self, device):
        self.device = device
        self._bufs = {}

    def get(self, key, nbytes):
        nbytes = max(int(nbytes), 1)
        t = self._bufs.get(key)
        if t is None or t.numel() < nbytes:
            t = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
            self._bufs[key] = t
        return t


_WS = {}


def _workspace(device):
    key = str(device)
    if key not in _WS:
        _WS[key] = Workspace(device)
    return _WS[key]


def alloc_outputs(arrays, keep_ekb=False):
    ntime, nbl = arrays['antenna1'].shape
    nchan = arrays['frequency'].shape[0]
    dev = arrays['uvw'].device
    ft = arrays['uvw'].dtype
    out = dict(model_vis=torch.empty((ntime, nbl, nchan, 4), dtype=complex_dtype(arrays), device=dev),
               chi_sqrd_result=torch.empty((ntime, nbl, nchan), dtype=ft, device=dev),
               X2=torch.zeros((), dtype=torch.float64, device=dev))
    if keep_ekb:
        out['ekb_vis'] = torch.zeros((ntime, nbl, nchan, 4), dtype=complex_dtype(arrays), device=dev)
    return out


# Default kernel form per dtype: (pb, qb, producer_waves) -- pair blocks of
# pb x qb antenna pairs per consumer thread; producer_waves 0 = the
# double-buffered kernel with an automatic producer count, < 0 = the
# single-role pair-block kernel; 'generic' = per-baseline kernel (measured
# fastest in double, where the pair-block accumulators spill).
DEFAULT_PAIR_BLOCK = {'f32': (2, 2, 0), 'f64': 'generic'}


def build_plan(antenna1, antenna2, na, device, form, point_gemm=True):
    """Launch plan from HOST antenna arrays: (mbx_pair_plan, device buffers
    it points to), or None for the generic kernel without the matrix-core
    point-source path.

    form: (pb, qb[, producer_waves]) pair blocks, or 'generic'.
    point_gemm: use the matrix cores for point sources when the layout is
    time-invariant with one baseline per antenna pair.
    """
    keep = []
    blocks, nblocks, pb, qb, pw = None, 0, 0, 0, 0
    if form != 'generic':
        f = tuple(form)
        pb, qb = f[0], f[1]
        pw = f[2] if len(f) > 2 else 0
        host = _lib.build_pair_plan(antenna1, antenna2, na, pb, qb)
        if host is not None:
            dev = torch.from_numpy(host).to(device)
            keep.append(dev)
            blocks, nblocks = dev.data_ptr(), host.shape[0]
    bmap = None
    if point_gemm:
        try:
            hm = _lib.build_baseline_map(antenna1, antenna2, na)
        except _lib.MontblancHipError as e:
            if e.code != _lib.E_UNSUPPORTED:
                raise
            hm = None
        if hm is not None:
            dm = torch.from_numpy(hm).to(device)
            keep.append(dm)
            bmap = dm.data_ptr()
    if blocks is None and bmap is None:
        return None
    return (_lib.mbx_pair_plan(blocks, nblocks, pb, qb, pw, bmap, 1 if bmap else 0), keep)


def pair_plan(arrays, pair_block=None, point_gemm=True):
    """Device launch plan for the antenna layout of ``arrays`` (cached in
    the dict under '_pair_plan'); None selects the generic kernel.

    pair_block: (pb, qb[, producer_waves]), 'generic', or None for the
    dtype default.
    """
    if pair_block is None:
        pair_block = DEFAULT_PAIR_BLOCK[suffix(arrays)]
    key = ('_pair_plan', pair_block if pair_block == 'generic' else tuple(pair_block),
           bool(point_gemm))
    cached = arrays.get('_pair_plan')
    if cached is not None and cached[0] == key:
        return cached[1]
    plan = build_plan(arrays['antenna1'].cpu().numpy(), arrays['antenna2'].cpu().numpy(),
                      arrays['uvw'].shape[1], arrays['uvw'].device, pair_block, point_gemm)
    arrays['_pair_plan'] = (key, plan)
    return plan


def forward(arrays, counts, residual=False, use_weight_vector=False, sigma_sqrd=None,
            outputs=None, keep_ekb=False, dims=None, stream=None, pair_block=None,
            point_gemm=True, workspace=None):
    """Fused forward on device tensors. Returns the outputs dict
    (model_vis, chi_sqrd_result, X2 [device float64 scalar], ekb_vis?).
    point_gemm=False keeps every source on the vector units (True: point
    sources, and from 40 antennas in float / 48 in float64 the gaussian and
    sersic sources, on the matrix cores where the layout allows);
    workspace (a Workspace) replaces the per-device one, for concurrent
    calls on several streams."""
    dev = arrays['uvw'].device
    out = outputs if outputs is not None else alloc_outputs(arrays, keep_ekb)
    d = make_dims(arrays, counts) if dims is None else dims
    plan = pair_plan(arrays, pair_block, point_gemm)
    ins = _lib.mbx_inputs(**{n: _raw(arrays.get(n)) for n in _lib.INPUT_NAMES})
    outs = _lib.mbx_outputs(_raw(out.get('model_vis')), _raw(out.get('chi_sqrd_result')),
                            _raw(out.get('X2')), _raw(out.get('ekb_vis')))
    lib = _lib.load()
    nb = lib.mbx_forward_workspace_bytes(ctypes.byref(d), int(is_double(arrays)))
    ws = (workspace or _workspace(dev)).get('forward', nb)
    b = beam_extents(arrays)
    k = consts(arrays, sigma_sqrd)
    h = stream_of(dev) if stream is None else ctypes.c_void_p(stream)
    _lib.call('mbx_forward_' + suffix(arrays), ctypes.byref(d), ctypes.byref(ins),
              ctypes.byref(outs), ctypes.byref(b), ctypes.byref(k),
              ctypes.byref(plan[0]) if plan is not None else None,
              options(residual, use_weight_vector), _ptr(ws), nb, h)
    return out


def baseline_map(arrays):
    """Device (na, na) pair -> baseline map of the antenna layout (cached in
    the dict under '_baseline_map')."""
    m = arrays.get('_baseline_map')
    if m is None:
        host = _lib.build_baseline_map(arrays['antenna1'].cpu().numpy(),
                                       arrays['antenna2'].cpu().numpy(), arrays['uvw'].shape[1])
        m = torch.from_numpy(host).to(arrays['uvw'].device)
        arrays['_baseline_map'] = m
    return m


def alloc_grads(arrays, counts):
    """float64 gradient buffers shaped like the parameters (G complex128)."""
    dev = arrays['uvw'].device
    nsrc = sum(counts)
    ntime, na = arrays['uvw'].shape[:2]
    nchan = arrays['frequency'].shape[0]
    f64 = dict(dtype=torch.float64, device=dev)
    return dict(lm=torch.empty((nsrc, 2), **f64), stokes=torch.empty((nsrc, ntime, 4), **f64),
                alpha=torch.empty((nsrc, ntime), **f64),
                gauss_shape=torch.empty((3, counts[1]), **f64),
                sersic_shape=torch.empty((3, counts[2]), **f64),
                G_term=torch.empty((ntime, na, nchan, 4), dtype=torch.complex128, device=dev))


def adjoint(arrays, counts, ekb_vis, grad_scale=1.0, use_weight_vector=False, sigma_sqrd=None,
            grads=None, stream=None, dims=None, workspace=None):
    """d chi^2 / d(lm, stokes, alpha, gauss_shape, sersic_shape, G_term) times
    grad_scale, from the forward's pre-gain sums ``ekb_vis``. Returns a dict
    of float64 tensors (G_term complex128: d/dRe + i d/dIm)."""
    dev = arrays['uvw'].device
    grads = alloc_grads(arrays, counts) if grads is None else grads
    d = make_dims(arrays, counts) if dims is None else dims
    bmap = baseline_map(arrays)
    ins = _lib.mbx_inputs(**{n: _raw(arrays.get(n)) for n in _lib.INPUT_NAMES})
    g = _lib.mbx_grads(**{n: _raw(grads.get(n)) for n in _lib.GRAD_NAMES})
    lib = _lib.load()
    nb = lib.mbx_adjoint_workspace_bytes(ctypes.byref(d), int(is_double(arrays)))
    ws = (workspace or _workspace(dev)).get('adjoint', nb)
    b = beam_extents(arrays)
    k = consts(arrays, sigma_sqrd)
    h = stream_of(dev) if stream is None else ctypes.c_void_p(stream)
    _lib.call('mbx_adjoint_' + suffix(arrays), ctypes.byref(d), ctypes.byref(ins), _ptr(ekb_vis),
              _ptr(bmap), ctypes.byref(g), ctypes.byref(b), ctypes.byref(k),
              options(False, use_weight_vector), float(grad_scale), _ptr(ws), nb, h)
    return grads


def forward_numpy(a, counts, device='cuda:0', **kw):
    """numpy in, numpy out (model_vis, chi_sqrd_result, X2 float)."""
    arrays = to_device(a, device)
    out = forward(arrays, counts, **kw)
    torch.cuda.synchronize(device)
    res = {k: v.cpu().numpy() for k, v in out.items() if k != 'X2'}
    res['X2'] = float(out['X2'].item())
    return res


# ---- stage functions (one v4 node each) ---------------------------------

def b_sqrt(arrays, counts, out=None):
    ns, nt = arrays['stokes'].shape[:2]
    nc = arrays['frequency'].shape[0]
    out = out if out is not None else torch.empty((ns, nt, nc, 4), dtype=complex_dtype(arrays),
                                                  device=arrays['uvw'].device)
    d = make_dims(arrays, counts)
    _lib.call('mbx_b_sqrt_' + suffix(arrays), ctypes.byref(d), _ptr(arrays['stokes']),
              _ptr(arrays['alpha']), _ptr(arrays['frequency']), _ptr(arrays['ref_frequency']),
              _ptr(out), stream_of(out.device))
    return out


def e_beam(arrays, counts, out=None):
    ns = arrays['lm'].shape[0]
    nt, na = arrays['uvw'].shape[:2]
    nc = arrays['frequency'].shape[0]
    out = out if out is not None else torch.empty((ns, nt, na, nc, 4), dtype=complex_dtype(arrays),
                                                  device=arrays['uvw'].device)
    d = make_dims(arrays, counts)
    b = beam_extents(arrays)
    _lib.call('mbx_e_beam_' + suffix(arrays), ctypes.byref(d), _ptr(arrays['lm']),
              _ptr(arrays['parallactic_angles']), _ptr(arrays['point_errors']),
              _ptr(arrays['antenna_scaling']), _ptr(arrays['frequency']), _ptr(arrays['E_beam']),
              _ptr(out), ctypes.byref(b), stream_of(out.device))
    return out


def ekb_sqrt(arrays, counts, B_sqrt, jones):
    d = make_dims(arrays, counts)
    k = consts(arrays)
    _lib.call('mbx_ekb_sqrt_' + suffix(arrays), ctypes.byref(d), _ptr(arrays['uvw']),
              _ptr(arrays['lm']), _ptr(arrays['frequency']), _ptr(B_sqrt), _ptr(jones),
              ctypes.byref(k), stream_of(jones.device))
    return jones


def sum_coherencies(arrays, counts, jones, model_vis, chi_sqrd_result, residual=False,
                    use_weight_vector=False, dims=None):
    d = make_dims(arrays, counts) if dims is None else dims
    k = consts(arrays)
    _lib.call('mbx_sum_coherencies_' + suffix(arrays), ctypes.byref(d), _ptr(arrays['uvw']),
              _ptr(arrays.get('gauss_shape')), _ptr(arrays.get('sersic_shape')),
              _ptr(arrays['frequency']), _ptr(arrays['antenna1']), _ptr(arrays['antenna2']),
              _ptr(jones), _ptr(arrays['flag']), _ptr(arrays.get('weight_vector')),
              _ptr(arrays['observed_vis']), _ptr(arrays['G_term']), _ptr(model_vis),
              _ptr(chi_sqrd_result), ctypes.byref(k), options(residual, use_weight_vector),
              stream_of(model_vis.device))


def chi2_reduce(arrays, counts, chi_sqrd_result, X2, use_weight_vector=False, sigma_sqrd=None,
                dims=None):
    d = make_dims(arrays, counts) if dims is None else dims
    k = consts(arrays, sigma_sqrd)
    lib = _lib.load()
    nb = lib.mbx_chi2_workspace_bytes()
    ws = _workspace(X2.device).get('chi2', nb)
    _lib.call('mbx_chi2_reduce_' + suffix(arrays), ctypes.byref(d), _ptr(chi_sqrd_result),
              _ptr(X2), ctypes.byref(k), options(False, use_weight_vector), _ptr(ws), nb,
              stream_of(X2.device))
    return X2
