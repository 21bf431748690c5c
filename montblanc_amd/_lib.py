"""ctypes binding of the gfx950 RIME library (include/montblanc_amd.h).

The shared library is built in-tree (montblanc_amd/libmontblanc_amd.so) by
``montblanc_amd/csrc/Makefile``. There is no fallback: if the library is
missing every entry point raises, so nothing silently runs on the CPU.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libmontblanc_amd.so')

ABI_VERSION = 3
APPLY_WEIGHTS = 1
RESIDUAL_OUTPUT = 2
# size of the library-held workspace of mbx_forward_static_* per device
STATIC_SCRATCH_BYTES = 64 << 20

ERROR_NAMES = {
    -1: 'MBX_E_NULL', -2: 'MBX_E_DIMS', -3: 'MBX_E_EXTENT',
    -4: 'MBX_E_WORKSPACE', -5: 'MBX_E_UNSUPPORTED', -6: 'MBX_E_NO_DEVICE',
}


E_UNSUPPORTED = -5


class MontblancHipError(RuntimeError):
    """A C-ABI entry point returned a non-zero status."""

    def __init__(self, code, where, message):
        self.code = code
        name = ERROR_NAMES.get(code, 'hipError_t %d' % code if code > 0 else str(code))
        super().__init__('%s failed (%s): %s' % (where, name, message))


class mbx_dim(ctypes.Structure):
    _fields_ = [('global_size', ctypes.c_uint32), ('local_size', ctypes.c_uint32),
                ('lower_extent', ctypes.c_uint32), ('upper_extent', ctypes.c_uint32)]


DIM_NAMES = ('ntime', 'na', 'nbl', 'nchan', 'npol', 'npolchan',
             'nsrc', 'npsrc', 'ngsrc', 'nssrc', 'beam_lw', 'beam_mh', 'beam_nud')


class mbx_dims(ctypes.Structure):
    _fields_ = [(n, mbx_dim) for n in DIM_NAMES]


class mbx_beam_extents(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ('ll', 'lm', 'lfreq', 'ul', 'um', 'ufreq')]


class mbx_consts(ctypes.Structure):
    _fields_ = [('gauss_scale', ctypes.c_double), ('two_pi_over_c', ctypes.c_double),
                ('sigma_sqrd', ctypes.c_double)]


INPUT_NAMES = ('uvw', 'lm', 'stokes', 'alpha', 'frequency', 'ref_frequency',
               'parallactic_angles', 'point_errors', 'antenna_scaling', 'E_beam',
               'gauss_shape', 'sersic_shape', 'antenna1', 'antenna2', 'flag',
               'weight_vector', 'observed_vis', 'G_term')


class mbx_inputs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in INPUT_NAMES]


OUTPUT_NAMES = ('model_vis', 'chi_sqrd_result', 'X2', 'ekb_vis')


class mbx_outputs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in OUTPUT_NAMES]


class mbx_pair_plan(ctypes.Structure):
    _fields_ = [('blocks', ctypes.c_void_p), ('nblocks', ctypes.c_int32),
                ('pb', ctypes.c_int32), ('qb', ctypes.c_int32),
                ('producer_waves', ctypes.c_int32), ('baseline_map', ctypes.c_void_p),
                ('point_gemm', ctypes.c_int32)]


GRAD_NAMES = ('lm', 'stokes', 'alpha', 'gauss_shape', 'sersic_shape', 'G_term')


class mbx_grads(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in GRAD_NAMES]


class mbx_batch(ctypes.Structure):
    _fields_ = [('nsamples', ctypes.c_int32)] + [
        (n, ctypes.c_void_p) for n in ('src_offset', 'npsrc', 'ngsrc', 'nssrc', 'gauss_offset',
                                       'sersic_offset', 'order')]


class mbx_batch_tuning(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ('workgroups', 'channels_per_item',
                                              'sources_per_tile', 'strategy', 'samples_per_group')]


BL_REVERSED = 1 << 30
BL_AUTO = 1 << 29

_P = ctypes.c_void_p
_PD = ctypes.POINTER(mbx_dims)
_PB = ctypes.POINTER(mbx_beam_extents)
_PK = ctypes.POINTER(mbx_consts)
_PI = ctypes.POINTER(mbx_inputs)
_PO = ctypes.POINTER(mbx_outputs)
_I = ctypes.c_int
_SZ = ctypes.c_size_t
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_PP = ctypes.POINTER(mbx_pair_plan)

SIGNATURES = {
    'mbx_abi_version': ([], _I),
    'mbx_last_error': ([], ctypes.c_char_p),
    'mbx_device_count': ([], _I),
    'mbx_chi2_workspace_bytes': ([], _SZ),
    'mbx_forward_workspace_bytes': ([_PD, _I], _SZ),
    'mbx_pair_plan_capacity': ([_I32, _I32, _I32], _I64),
    'mbx_build_pair_plan': ([_P, _P, _I32, _I32, _I32, _I32, _I32, _P, _I64], _I64),
    'mbx_build_baseline_map': ([_P, _P, _I32, _I32, _I32, _P], _I),
    'mbx_adjoint_workspace_bytes': ([_PD, _I], _SZ),
    'mbx_batched_workspace_bytes': ([_PD, _I, ctypes.POINTER(mbx_batch_tuning)], _SZ),
}
for _t in ('f32', 'f64'):
    SIGNATURES.update({
        'mbx_b_sqrt_' + _t: ([_PD, _P, _P, _P, _P, _P, _P], _I),
        'mbx_e_beam_' + _t: ([_PD, _P, _P, _P, _P, _P, _P, _P, _PB, _P], _I),
        'mbx_ekb_sqrt_' + _t: ([_PD, _P, _P, _P, _P, _P, _PK, _P], _I),
        'mbx_sum_coherencies_' + _t: ([_PD] + [_P] * 13 + [_PK, _I, _P], _I),
        'mbx_chi2_reduce_' + _t: ([_PD, _P, _P, _PK, _I, _P, _SZ, _P], _I),
        'mbx_forward_' + _t: ([_PD, _PI, _PO, _PB, _PK, _PP, _I, _P, _SZ, _P], _I),
        'mbx_forward_static_' + _t: ([_PD, _PI, _PO, _PB, _PK, _PP, _I, _P], _I),
        'mbx_adjoint_' + _t: ([_PD, _PI, _P, _P, ctypes.POINTER(mbx_grads), _PB, _PK, _I,
                               ctypes.c_double, _P, _SZ, _P], _I),
        'mbx_batched_' + _t: ([_PD, _PI, ctypes.POINTER(mbx_batch), _P, ctypes.POINTER(mbx_grads),
                               _P, _PB, _PK, ctypes.POINTER(mbx_batch_tuning), _I, _P, _SZ, _P], _I),
    })

_lib = None


def load():
    """Load the library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            'montblanc_amd: %s is missing; build it with '
            '`make -C montblanc_amd/csrc` (hipcc, gfx950). There is no CPU '
            'fallback.' % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    v = lib.mbx_abi_version()
    if v != ABI_VERSION:
        raise ImportError('montblanc_amd: ABI version %d, expected %d' % (v, ABI_VERSION))
    _lib = lib
    return lib


def call(name, *args):
    """Call entry point ``name``; raise MontblancHipError on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.mbx_last_error()
        raise MontblancHipError(rc, name, msg.decode() if msg else '')
    return rc


def make_dims(sizes, extents=None, global_sizes=None):
    """Build an mbx_dims from a dict of local sizes.

    extents: optional dict name -> (lower, upper); default (0, local).
    global_sizes: optional dict name -> global size; default local.
    """
    d = mbx_dims()
    extents = extents or {}
    global_sizes = global_sizes or {}
    for n in DIM_NAMES:
        local = int(sizes.get(n, 0))
        lo, hi = extents.get(n, (0, local))
        g = int(global_sizes.get(n, max(local, hi)))
        setattr(d, n, mbx_dim(g, local, int(lo), int(hi)))
    return d


def build_pair_plan(antenna1, antenna2, na, pb, qb):
    """Pair-block plan (host int32 array) of a (ntime, nbl) antenna layout,
    or None when the layout changes with time (generic kernel)."""
    import numpy as np
    a1 = np.ascontiguousarray(antenna1, dtype=np.int32)
    a2 = np.ascontiguousarray(antenna2, dtype=np.int32)
    ntime, nbl = a1.shape
    lib = load()
    cap = lib.mbx_pair_plan_capacity(nbl, pb, qb)
    out = np.empty(max(int(cap), 1), dtype=np.int32)
    n = lib.mbx_build_pair_plan(a1.ctypes.data, a2.ctypes.data, ntime, nbl, na, pb, qb,
                                out.ctypes.data, cap)
    if n == -5:          # MBX_E_UNSUPPORTED: layout varies with time
        return None
    if n < 0:
        raise MontblancHipError(int(n), 'mbx_build_pair_plan', lib.mbx_last_error().decode())
    return out[: n * (2 + pb * qb)].reshape(n, 2 + pb * qb)


def build_baseline_map(antenna1, antenna2, na):
    """(na, na) int32 pair -> baseline map (host array) for the adjoint
    and the matrix-core point-source path; raises MontblancHipError
    (MBX_E_UNSUPPORTED) for layouts that change with time or repeat a pair."""
    import numpy as np
    a1 = np.ascontiguousarray(antenna1, dtype=np.int32)
    a2 = np.ascontiguousarray(antenna2, dtype=np.int32)
    ntime, nbl = a1.shape
    out = np.empty((na, na), dtype=np.int32)
    call('mbx_build_baseline_map', a1.ctypes.data, a2.ctypes.data, ntime, nbl, na, out.ctypes.data)
    return out
