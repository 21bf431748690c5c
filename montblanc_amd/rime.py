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
    """dict of numpy arrays or torch tensors (v4 names) -> dict of
    contiguous device tensors.

    dtype: np.float32 / np.float64 (or the torch equivalents) for the float
    arrays, complex arrays following; None keeps each array's own dtype.
    Antenna indices always become int32 and flags uint8, whatever the input
    type (torch's default integer type is int64, which the kernels would
    read as pairs of int32). Scalars (beam extents, constants) are kept as
    Python floats.
    """
    out = {}
    if isinstance(dtype, torch.dtype):
        dtype = {torch.float32: np.float32, torch.float64: np.float64}[dtype]
    ft = None if dtype is None else np.dtype(dtype)
    ct = None if ft is None else (np.complex64 if ft == np.float32 else np.complex128)
    tft = None if ft is None else (torch.float32 if ft == np.float32 else torch.float64)
    tct = None if ft is None else (torch.complex64 if ft == np.float32 else torch.complex128)
    for k, v in a.items():
        if isinstance(v, np.ndarray):
            if k in INT_ARRAYS:
                v = v.astype(np.int32)
            elif k == 'flag':
                v = v.astype(np.uint8)
            elif v.dtype.kind == 'f' and ft is not None:
                v = v.astype(ft)
            elif v.dtype.kind == 'c' and ct is not None:
                v = v.astype(ct)
            out[k] = torch.from_numpy(np.ascontiguousarray(v)).to(device)
        elif isinstance(v, torch.Tensor):
            if k in INT_ARRAYS:
                v = v.to(torch.int32)
            elif k == 'flag':
                v = v.to(torch.uint8)
            elif v.is_floating_point() and tft is not None:
                v = v.to(tft)
            elif v.is_complex() and tct is not None:
                v = v.to(tct)
            out[k] = v.to(device).contiguous()
        else:
            out[k] = v
    return out


# v4 array shapes in terms of the dimension names (v4/config.py:131-247)
ARRAY_SHAPES = {
    'uvw': ('ntime', 'na', 3), 'lm': ('nsrc', 2), 'stokes': ('nsrc', 'ntime', 4),
    'alpha': ('nsrc', 'ntime'), 'frequency': ('nchan',), 'ref_frequency': ('nchan',),
    'parallactic_angles': ('ntime', 'na'), 'point_errors': ('ntime', 'na', 'nchan', 2),
    'antenna_scaling': ('na', 'nchan', 2), 'E_beam': ('beam_lw', 'beam_mh', 'beam_nud', 4),
    'gauss_shape': (3, 'ngsrc'), 'sersic_shape': (3, 'nssrc'),
    'antenna1': ('ntime', 'nbl'), 'antenna2': ('ntime', 'nbl'),
    'flag': ('ntime', 'nbl', 'nchan', 4), 'weight_vector': ('ntime', 'nbl', 'nchan', 4),
    'observed_vis': ('ntime', 'nbl', 'nchan', 4), 'G_term': ('ntime', 'na', 'nchan', 4),
    'model_vis': ('ntime', 'nbl', 'nchan', 4), 'ekb_vis': ('ntime', 'nbl', 'nchan', 4),
    'chi_sqrd_result': ('ntime', 'nbl', 'nchan'),
}


def _sizes(arrays, counts):
    ntime, na = arrays['uvw'].shape[:2]
    lw, mh, nud = arrays['E_beam'].shape[:3] if arrays.get('E_beam') is not None else (0, 0, 0)
    return dict(ntime=ntime, na=na, nbl=arrays['antenna1'].shape[1],
                nchan=arrays['frequency'].shape[0], nsrc=sum(counts), npsrc=counts[0],
                ngsrc=counts[1], nssrc=counts[2], beam_lw=lw, beam_mh=mh, beam_nud=nud)


def check_arrays(arrays, counts, names=None, outputs=None, dims=None):
    """Raise ValueError unless every array the kernels will read or write
    has the run dtype, lives on the device of ``uvw``, is contiguous and
    has the v4 shape implied by uvw (ntime, na), antenna1 (nbl), frequency
    (nchan), E_beam (beam) and counts = (npsrc, ngsrc, nssrc). Runs before
    any launch, so a bad argument never reaches a kernel as a raw pointer.

    names: the input names to check (default: every input present);
    outputs: a dict of output tensors to check the same way;
    dims: an mbx_dims whose local sizes must agree with the arrays."""
    for k in ('uvw', 'antenna1', 'frequency'):
        if not isinstance(arrays.get(k), torch.Tensor):
            raise ValueError('array %r is missing or not a torch tensor' % k)
    ft = arrays['uvw'].dtype
    if ft not in (torch.float32, torch.float64):
        raise ValueError('uvw must be float32 or float64, not %s' % ft)
    ct = torch.complex64 if ft == torch.float32 else torch.complex128
    dev = arrays['uvw'].device
    if len(counts) != 3 or min(counts) < 0:
        raise ValueError('counts must be (npsrc, ngsrc, nssrc) >= 0, got %r' % (counts,))
    sz = _sizes(arrays, counts)
    items = [(n, arrays.get(n)) for n in (names if names is not None else _lib.INPUT_NAMES)]
    items += list((outputs or {}).items())
    for n, t in items:
        if t is None or n not in ARRAY_SHAPES:
            continue
        if not isinstance(t, torch.Tensor):
            raise ValueError('array %r must be a torch tensor, not %s' % (n, type(t).__name__))
        if n in INT_ARRAYS:
            want = torch.int32
        elif n == 'flag':
            want = torch.uint8
        elif n in COMPLEX_ARRAYS or n in ('model_vis', 'ekb_vis'):
            want = ct
        else:
            want = ft
        if t.dtype != want:
            raise ValueError('array %r has dtype %s; this call needs %s (the dtype of uvw '
                             'decides float32 or float64)' % (n, t.dtype, want))
        if t.device != dev:
            raise ValueError('array %r is on %s, uvw on %s' % (n, t.device, dev))
        if not t.is_contiguous():
            raise ValueError('array %r is not contiguous' % n)
        shape = tuple(s if isinstance(s, int) else sz[s] for s in ARRAY_SHAPES[n])
        if n in ('gauss_shape', 'sersic_shape') and shape[1] == 0 and t.numel() == 0:
            continue
        if tuple(t.shape) != shape:
            raise ValueError('array %r has shape %s; the dimensions give %s' %
                             (n, tuple(t.shape), shape))
    x2 = (outputs or {}).get('X2')
    if x2 is not None and (x2.dtype != torch.float64 or x2.device != dev):
        raise ValueError('X2 must be a float64 scalar tensor on %s' % dev)
    if dims is not None:
        # extents against global and local sizes are validated by the library
        for n in ('ntime', 'na', 'nbl', 'nchan', 'npsrc', 'ngsrc', 'nssrc'):
            dm = getattr(dims, n)
            if dm.local_size != sz[n]:
                raise ValueError('dims.%s.local_size = %d but the arrays give %d' %
                                 (n, dm.local_size, sz[n]))


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


def default_pair_block(sfx, na):
    """Default form of the vector-unit forward kernel: (pb, qb,
    producer_waves) -- pair blocks of pb x qb antenna pairs per consumer
    thread; producer_waves 0 = the double-buffered producer / consumer
    kernel with an automatic producer count, < 0 = the single-role
    pair-block kernel; 'generic' = the per-baseline kernel.

    Measured on MI355X (forward, median of 5): up to 64 antennas, where the
    point sources run on the matrix cores and this kernel sums the other
    sources and applies the epilogue, (1, 1, 0) in both dtypes (27
    antennas, 200 times, 128 channels, 300 point + 200 gaussian: float
    10.37 ms against 11.65 for (2, 2, 0) and 10.91 generic; float64 23.17
    against 25.23 generic; also ahead at 14 and 16 antennas). Above 64
    antennas every source runs here: float (2, 2, 0), float64 (2, 2, -1)
    (100 antennas, 320 sources: 1.79 and 5.91 ms against 5.50 and 8.42
    generic, 5.73 and 15.48 for (1, 1, 0))."""
    if na <= 64:
        return (1, 1, 0)
    return (2, 2, 0) if sfx == 'f32' else (2, 2, -1)


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


def _plan_key(sfx, pair_block, point_gemm, na):
    if pair_block is None:
        pair_block = default_pair_block(sfx, na)
    return ('_pair_plan', pair_block if pair_block == 'generic' else tuple(pair_block),
            bool(point_gemm)), pair_block


def plan_entry(antenna1, antenna2, na, device, sfx, pair_block=None, point_gemm=True):
    """The '_pair_plan' cache entry of a forward call, built from HOST
    antenna arrays (no device-to-host copy): callers that hold the host
    layout (the chunked solver) put it in the arrays dict beforehand."""
    key, form = _plan_key(sfx, pair_block, point_gemm, na)
    return (key, build_plan(np.asarray(antenna1), np.asarray(antenna2), na, device, form, point_gemm))


def pair_plan(arrays, pair_block=None, point_gemm=True):
    """Device launch plan for the antenna layout of ``arrays`` (cached in
    the dict under '_pair_plan'); None selects the generic kernel.

    pair_block: (pb, qb[, producer_waves]), 'generic', or None for the
    default of the dtype and antenna count (default_pair_block).
    """
    key, form = _plan_key(suffix(arrays), pair_block, point_gemm, arrays['uvw'].shape[1])
    cached = arrays.get('_pair_plan')
    if cached is not None and cached[0] == key:
        return cached[1]
    plan = build_plan(arrays['antenna1'].cpu().numpy(), arrays['antenna2'].cpu().numpy(),
                      arrays['uvw'].shape[1], arrays['uvw'].device, form, point_gemm)
    arrays['_pair_plan'] = (key, plan)
    return plan


def forward(arrays, counts, residual=False, use_weight_vector=False, sigma_sqrd=None,
            outputs=None, keep_ekb=False, dims=None, stream=None, pair_block=None,
            point_gemm=True, workspace=None, static_scratch=False):
    """Fused forward on device tensors. Returns the outputs dict
    (model_vis, chi_sqrd_result, X2 [device float64 scalar], ekb_vis?).
    point_gemm=False keeps every source on the vector units (True: point
    sources, and from 40 antennas in float / 48 in float64 the gaussian and
    sersic sources, on the matrix cores where the layout allows);
    workspace (a Workspace) replaces the per-device one, for concurrent
    calls on several streams; static_scratch=True calls the entry point
    whose workspace is the library's own device array
    (mbx_forward_static_*, at most _lib.STATIC_SCRATCH_BYTES; calls that
    overlap on one device race on it)."""
    dev = arrays['uvw'].device
    check_arrays(arrays, counts, dims=dims)
    out = outputs if outputs is not None else alloc_outputs(arrays, keep_ekb)
    if outputs is not None:
        check_arrays(arrays, counts, names=(), outputs=out)
    d = make_dims(arrays, counts) if dims is None else dims
    plan = pair_plan(arrays, pair_block, point_gemm)
    ins = _lib.mbx_inputs(**{n: _raw(arrays.get(n)) for n in _lib.INPUT_NAMES})
    outs = _lib.mbx_outputs(_raw(out.get('model_vis')), _raw(out.get('chi_sqrd_result')),
                            _raw(out.get('X2')), _raw(out.get('ekb_vis')))
    b = beam_extents(arrays)
    k = consts(arrays, sigma_sqrd)
    h = stream_of(dev) if stream is None else ctypes.c_void_p(stream)
    if static_scratch:
        _lib.call('mbx_forward_static_' + suffix(arrays), ctypes.byref(d), ctypes.byref(ins),
                  ctypes.byref(outs), ctypes.byref(b), ctypes.byref(k),
                  ctypes.byref(plan[0]) if plan is not None else None,
                  options(residual, use_weight_vector), h)
        return out
    lib = _lib.load()
    nb = lib.mbx_forward_workspace_bytes(ctypes.byref(d), int(is_double(arrays)))
    ws = (workspace or _workspace(dev)).get('forward', nb)
    _lib.call('mbx_forward_' + suffix(arrays), ctypes.byref(d), ctypes.byref(ins),
              ctypes.byref(outs), ctypes.byref(b), ctypes.byref(k),
              ctypes.byref(plan[0]) if plan is not None else None,
              options(residual, use_weight_vector), _ptr(ws), nb, h)
    return out


def baseline_map(arrays):
    """Device (na, na) pair -> baseline map of the antenna layout (cached in
    the dict under '_baseline_map').

    Raises ValueError for layouts the adjoint does not take: antenna pairs
    that change with time, or a pair that appears on two baselines. The
    forward accepts those (on its per-baseline kernel); the adjoint's
    source passes work on the per-(time, channel) antenna-pair matrix,
    which needs one baseline per pair for the whole run."""
    m = arrays.get('_baseline_map')
    if m is None:
        try:
            host = _lib.build_baseline_map(arrays['antenna1'].cpu().numpy(),
                                           arrays['antenna2'].cpu().numpy(),
                                           arrays['uvw'].shape[1])
        except _lib.MontblancHipError as e:
            if e.code != _lib.E_UNSUPPORTED:
                raise
            raise ValueError('gradients need an antenna layout that is the same at every time '
                             'step with at most one baseline per antenna pair (%s)' % e) from None
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


def _check_grads(arrays, counts, grads):
    sz = _sizes(arrays, counts)
    want = dict(lm=(sz['nsrc'], 2), stokes=(sz['nsrc'], sz['ntime'], 4),
                alpha=(sz['nsrc'], sz['ntime']), gauss_shape=(3, sz['ngsrc']),
                sersic_shape=(3, sz['nssrc']), G_term=(sz['ntime'], sz['na'], sz['nchan'], 4))
    dev = arrays['uvw'].device
    for n, shape in want.items():
        t = grads.get(n)
        if t is None:
            continue
        dt = torch.complex128 if n == 'G_term' else torch.float64
        if (not isinstance(t, torch.Tensor) or t.dtype != dt or t.device != dev
                or not t.is_contiguous() or tuple(t.shape) != shape):
            raise ValueError('gradient buffer %r must be a contiguous %s tensor of shape %s on %s'
                             % (n, dt, shape, dev))


def adjoint(arrays, counts, ekb_vis, grad_scale=1.0, use_weight_vector=False, sigma_sqrd=None,
            grads=None, stream=None, dims=None, workspace=None):
    """d chi^2 / d(lm, stokes, alpha, gauss_shape, sersic_shape, G_term) times
    grad_scale, from the forward's pre-gain sums ``ekb_vis``. Returns a dict
    of float64 tensors (G_term complex128: d/dRe + i d/dIm)."""
    dev = arrays['uvw'].device
    check_arrays(arrays, counts, outputs={'ekb_vis': ekb_vis}, dims=dims)
    bmap = baseline_map(arrays)
    grads = alloc_grads(arrays, counts) if grads is None else grads
    _check_grads(arrays, counts, grads)
    d = make_dims(arrays, counts) if dims is None else dims
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

def _check_out(arrays, t, shape):
    """A per-source scratch array (B_sqrt, jones) of the run's complex type."""
    ct = complex_dtype(arrays)
    dev = arrays['uvw'].device
    if (not isinstance(t, torch.Tensor) or t.dtype != ct or t.device != dev
            or not t.is_contiguous() or tuple(t.shape) != tuple(shape)):
        raise ValueError('scratch array must be a contiguous %s tensor of shape %s on %s'
                         % (ct, tuple(shape), dev))


def b_sqrt(arrays, counts, out=None):
    ns, nt = arrays['stokes'].shape[:2]
    nc = arrays['frequency'].shape[0]
    check_arrays(arrays, counts, names=('stokes', 'alpha', 'frequency', 'ref_frequency'))
    out = out if out is not None else torch.empty((ns, nt, nc, 4), dtype=complex_dtype(arrays),
                                                  device=arrays['uvw'].device)
    _check_out(arrays, out, (ns, nt, nc, 4))
    d = make_dims(arrays, counts)
    _lib.call('mbx_b_sqrt_' + suffix(arrays), ctypes.byref(d), _ptr(arrays['stokes']),
              _ptr(arrays['alpha']), _ptr(arrays['frequency']), _ptr(arrays['ref_frequency']),
              _ptr(out), stream_of(out.device))
    return out


def e_beam(arrays, counts, out=None):
    ns = arrays['lm'].shape[0]
    nt, na = arrays['uvw'].shape[:2]
    nc = arrays['frequency'].shape[0]
    check_arrays(arrays, counts, names=('lm', 'parallactic_angles', 'point_errors',
                                        'antenna_scaling', 'frequency', 'E_beam'))
    out = out if out is not None else torch.empty((ns, nt, na, nc, 4), dtype=complex_dtype(arrays),
                                                  device=arrays['uvw'].device)
    _check_out(arrays, out, (ns, nt, na, nc, 4))
    d = make_dims(arrays, counts)
    b = beam_extents(arrays)
    _lib.call('mbx_e_beam_' + suffix(arrays), ctypes.byref(d), _ptr(arrays['lm']),
              _ptr(arrays['parallactic_angles']), _ptr(arrays['point_errors']),
              _ptr(arrays['antenna_scaling']), _ptr(arrays['frequency']), _ptr(arrays['E_beam']),
              _ptr(out), ctypes.byref(b), stream_of(out.device))
    return out


def ekb_sqrt(arrays, counts, B_sqrt, jones):
    check_arrays(arrays, counts, names=('uvw', 'lm', 'frequency'))
    ns, (nt, na), nc = sum(counts), arrays['uvw'].shape[:2], arrays['frequency'].shape[0]
    _check_out(arrays, B_sqrt, (ns, nt, nc, 4))
    _check_out(arrays, jones, (ns, nt, na, nc, 4))
    d = make_dims(arrays, counts)
    k = consts(arrays)
    _lib.call('mbx_ekb_sqrt_' + suffix(arrays), ctypes.byref(d), _ptr(arrays['uvw']),
              _ptr(arrays['lm']), _ptr(arrays['frequency']), _ptr(B_sqrt), _ptr(jones),
              ctypes.byref(k), stream_of(jones.device))
    return jones


def sum_coherencies(arrays, counts, jones, model_vis, chi_sqrd_result, residual=False,
                    use_weight_vector=False, dims=None):
    check_arrays(arrays, counts, names=('uvw', 'gauss_shape', 'sersic_shape', 'frequency',
                                        'antenna1', 'antenna2', 'flag', 'weight_vector',
                                        'observed_vis', 'G_term'),
                 outputs=dict(model_vis=model_vis, chi_sqrd_result=chi_sqrd_result), dims=dims)
    ns, (nt, na), nc = sum(counts), arrays['uvw'].shape[:2], arrays['frequency'].shape[0]
    _check_out(arrays, jones, (ns, nt, na, nc, 4))
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
    check_arrays(arrays, counts, names=(), outputs=dict(chi_sqrd_result=chi_sqrd_result, X2=X2),
                 dims=dims)
    d = make_dims(arrays, counts) if dims is None else dims
    k = consts(arrays, sigma_sqrd)
    lib = _lib.load()
    nb = lib.mbx_chi2_workspace_bytes()
    ws = _workspace(X2.device).get('chi2', nb)
    _lib.call('mbx_chi2_reduce_' + suffix(arrays), ctypes.byref(d), _ptr(chi_sqrd_result),
              _ptr(X2), ctypes.byref(k), options(False, use_weight_vector), _ptr(ws), nb,
              stream_of(X2.device))
    return X2
