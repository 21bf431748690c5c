"""CPU tests of the drop-in boundary: the C-ABI library's exports and
argument validation (no kernel is launched), the configuration, the
source-range bookkeeping and the pipeline's error semantics."""

import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, 'include', f) for f in os.listdir(os.path.join(ROOT, 'include'))
           if f.endswith('.h')]


def declared_symbols():
    names = set()
    for h in HEADERS:
        src = open(h).read()
        src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
        for m in re.finditer(r'\b(mbx_[a-z0-9_]+)\s*\(', src):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    from montblanc_amd import _lib
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 19
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert set(_lib.SIGNATURES) == names
    assert lib.mbx_abi_version() == _lib.ABI_VERSION


def test_struct_layouts_match_header():
    from montblanc_amd import _lib
    assert ctypes.sizeof(_lib.mbx_dim) == 16
    assert ctypes.sizeof(_lib.mbx_dims) == 13 * 16
    assert ctypes.sizeof(_lib.mbx_inputs) == 18 * 8
    assert ctypes.sizeof(_lib.mbx_outputs) == 4 * 8
    hdr = open(os.path.join(ROOT, 'include', 'montblanc_amd.h')).read()
    hdr = re.sub(r'/\*.*?\*/', '', hdr, flags=re.S)
    block = re.search(r'typedef struct mbx_inputs \{(.*?)\} mbx_inputs;', hdr, re.S).group(1)
    order = re.findall(r'\*\s*([A-Za-z_0-9]+)', block)
    assert tuple(order) == _lib.INPUT_NAMES


def dims_for(na=4, ntime=2, nchan=3, counts=(1, 0, 0), beam=(3, 3, 3), **over):
    from montblanc_amd import _lib
    sizes = dict(ntime=ntime, na=na, nbl=na * (na - 1) // 2, nchan=nchan, npol=4,
                 npolchan=4 * nchan, nsrc=sum(counts), npsrc=counts[0], ngsrc=counts[1],
                 nssrc=counts[2], beam_lw=beam[0], beam_mh=beam[1], beam_nud=beam[2])
    sizes.update(over)
    return _lib.make_dims(sizes)


def test_validation_errors_without_launching():
    """Bad arguments come back as negative codes before any HIP call."""
    from montblanc_amd import _lib
    lib = _lib.load()
    fake = ctypes.c_void_p(0x1000)    # never dereferenced: validation fails first
    assert lib.mbx_b_sqrt_f32(None, fake, fake, fake, fake, fake, None) == -1
    d = dims_for(npol=3)
    assert lib.mbx_b_sqrt_f32(ctypes.byref(d), fake, fake, fake, fake, fake, None) == -5
    assert b'npol' in lib.mbx_last_error()
    d = dims_for()
    d.nsrc.upper_extent = 5     # beyond global size
    assert lib.mbx_b_sqrt_f64(ctypes.byref(d), fake, fake, fake, fake, fake, None) == -3
    d = dims_for()
    assert lib.mbx_b_sqrt_f64(ctypes.byref(d), None, fake, fake, fake, fake, None) == -1
    d = dims_for(counts=(1, 1, 0))
    k = _lib.mbx_consts(1.0, 1.0, 1.0)
    # gaussian sources but no gauss_shape
    rc = lib.mbx_sum_coherencies_f32(ctypes.byref(d), fake, None, None, fake, fake, fake, fake,
                                     fake, None, fake, fake, fake, fake, ctypes.byref(k), 0, None)
    assert rc == -1 and b'gauss_shape' in lib.mbx_last_error()
    rc = lib.mbx_sum_coherencies_f32(ctypes.byref(dims_for()), fake, None, None, fake, fake,
                                     fake, fake, fake, None, fake, fake, fake, fake,
                                     ctypes.byref(k), 1, None)      # weights but no vector
    assert rc == -1
    rc = lib.mbx_chi2_reduce_f64(ctypes.byref(dims_for()), fake, fake, ctypes.byref(k), 0,
                                 fake, 8, None)
    assert rc == -4
    assert lib.mbx_forward_workspace_bytes(ctypes.byref(dims_for()), 1) > 0
    ins = _lib.mbx_inputs()
    outs = _lib.mbx_outputs()
    b = _lib.mbx_beam_extents(-1, -1, 1, 1, 1, 2)
    rc = lib.mbx_forward_f64(ctypes.byref(dims_for()), ctypes.byref(ins), ctypes.byref(outs),
                             ctypes.byref(b), ctypes.byref(k), None, 0, fake, 1 << 30, None)
    assert rc == -1
    # the no-allocation forward rejects dimensions whose workspace exceeds
    # its static array before touching a device: 256 channels of a
    # 129 x 129 beam are 204 MB of beam planes in float32
    big = dims_for()
    big.nchan = _lib.mbx_dim(256, 256, 0, 256)
    big.npolchan = _lib.mbx_dim(1024, 1024, 0, 1024)
    for n in ('beam_lw', 'beam_mh'):
        setattr(big, n, _lib.mbx_dim(129, 129, 0, 129))
    assert lib.mbx_forward_workspace_bytes(ctypes.byref(big), 0) > _lib.STATIC_SCRATCH_BYTES
    rc = lib.mbx_forward_static_f32(ctypes.byref(big), ctypes.byref(ins), ctypes.byref(outs),
                                    ctypes.byref(b), ctypes.byref(k), None, 0, None)
    assert rc == -4 and b'static scratch' in lib.mbx_last_error()
    assert lib.mbx_forward_static_f64(None, None, None, None, None, None, 0, None) == -1


def test_device_count_never_fails():
    from montblanc_amd import _lib
    assert _lib.load().mbx_device_count() >= 0


# ---- config -------------------------------------------------------------

def test_config_defaults_kwargs_and_ini(tmp_path):
    import montblanc_amd as mb
    cfg = mb.rime_solver_cfg()
    assert cfg['ntime'] == 10 and cfg['na'] == 7 and cfg['nchan'] == 16
    assert cfg['dtype'] == 'double' and cfg['version'] == 'v4'
    assert cfg['beam_lw'] == 50 and cfg['weight_vector'] is False
    assert cfg['vis_output'] == 'model' and cfg['source_batch_size'] == 500
    assert cfg['data_source'] == 'ms'
    assert dict(cfg['sources']) == {'point': 1, 'gaussian': 0, 'sersic': 0}
    ini = tmp_path / 'mb.ini'
    ini.write_text('[montblanc]\nntime = 33\nna = 9\ndtype = float\n'
                   'sources = {"point": 4, "gaussian": 2}\n')
    cfg = mb.rime_solver_cfg(cfg_file=str(ini), na=11, mem_budget=123)
    assert cfg['ntime'] == 33 and cfg['na'] == 11 and cfg['dtype'] == 'float'
    assert cfg['sources'] == {'point': 4, 'gaussian': 2}
    assert cfg['mem_budget'] == 123      # unknown keys pass through
    # like argparse defaults in the reference, choices are not enforced
    # here; the solver rejects the dtype (rime_solver.py:55-62)
    assert mb.rime_solver_cfg(dtype='half')['dtype'] == 'half'


def test_sources_helpers():
    import montblanc_amd as mb
    assert dict(mb.sources(point=10, gaussian=20)) == {'point': 10, 'gaussian': 20, 'sersic': 0}
    assert dict(mb.sources()) == {'point': 1, 'gaussian': 0, 'sersic': 0}
    with pytest.raises(TypeError):
        mb.sources(point='ten')
    assert mb.constants.C == 299792458
    with pytest.raises(SyntaxError):
        mb.constants.C = 3


def test_source_range_known_answers():
    """montblanc/tests/test_source_utils.py:46-106."""
    from montblanc_amd import src_types as mbs
    d = mbs.sources_to_nr_vars({'point': 10, 'gaussian': 20, 'sersic': 40})
    cases = {(0, 45): (10, 20, 15), (3, 45): (7, 20, 15), (13, 45): (0, 17, 15),
             (0, 3): (3, 0, 0), (15, 70): (0, 15, 40)}
    for (s, e), want in cases.items():
        D = mbs.source_range(s, e, d)
        assert (D['npsrc'], D['ngsrc'], D['nssrc']) == want
    assert mbs.source_range(15, 18, d)['ngsrc'] == 3
    tup = {(0, 45): ((0, 10), (0, 20), (0, 15)), (3, 45): ((3, 10), (0, 20), (0, 15)),
           (13, 45): ((0, 0), (3, 20), (0, 15)), (0, 3): ((0, 3), (0, 0), (0, 0)),
           (15, 18): ((0, 0), (5, 8), (0, 0)), (15, 70): ((0, 0), (5, 20), (0, 40))}
    for (s, e), want in tup.items():
        D = mbs.source_range_tuple(s, e, d)
        assert (D['npsrc'], D['ngsrc'], D['nssrc']) == want
    sl = mbs.source_range_slices(15, 18, d)
    assert sl['ngsrc'] == slice(5, 8, 1)


def test_baseline_antenna_counts():
    """montblanc/tests/test_utils.py:109-127."""
    from montblanc_amd.solver import nr_of_antenna, nr_of_baselines
    for na in (7, 14, 27, 64, 3500):
        nbl, nbl_auto = na * (na - 1) // 2, na * (na + 1) // 2
        assert nr_of_baselines(na) == nbl == nr_of_baselines(na, False)
        assert nr_of_baselines(na, True) == nbl_auto
        assert nr_of_antenna(nbl) == na == nr_of_antenna(nbl, False)
        assert nr_of_antenna(nbl_auto, True) == na


# ---- pipeline -----------------------------------------------------------

def test_pipeline_swallows_errors_and_keeps_going_on_shutdown():
    from montblanc_amd.pipeline import Node, NullNode, Pipeline

    calls = []

    class Boom(NullNode):
        def execute(self, solver, stream=None):
            calls.append('boom')
            raise RuntimeError('kernel failed')

        def shutdown(self, solver, stream=None):
            calls.append('shutdown-boom')
            raise RuntimeError('shutdown failed')

    class Ok(NullNode):
        def shutdown(self, solver, stream=None):
            calls.append('shutdown-ok')

    p = Pipeline([Boom(), Ok()])
    assert p.initialise(None) is True
    assert p.execute(None) is False
    assert isinstance(p.last_error, RuntimeError)
    assert p.shutdown(None) is False
    assert calls == ['boom', 'shutdown-boom', 'shutdown-ok']
    with pytest.raises(TypeError):
        Pipeline(Ok())
    with pytest.raises(NotImplementedError):
        Node().execute(None)
    assert str(Pipeline([Ok(), Ok()])) == 'Ok Ok'


# ---- pair-block planner (host code of the library) ------------------------

@pytest.mark.parametrize('na,pb,qb,auto', [(64, 2, 2, False), (27, 2, 4, False), (7, 1, 1, False),
                                           (9, 2, 2, True), (5, 2, 4, True)])
def test_pair_plan_covers_every_baseline_once(na, pb, qb, auto):
    from montblanc_amd import _lib
    from oracle import rime_np
    a1, a2 = rime_np.default_ant_pairs(na, 3, auto_correlations=auto)
    plan = _lib.build_pair_plan(a1, a2, na, pb, qb)
    nbl = a1.shape[1]
    seen = np.zeros(nbl, int)
    for row in plan:
        p0, q0 = row[0], row[1]
        assert p0 % pb == 0 and q0 % qb == 0
        for s, b in enumerate(row[2:]):
            if b < 0:
                continue
            i, j = divmod(s, qb)
            assert (a1[0, b], a2[0, b]) == (p0 + i, q0 + j)
            seen[b] += 1
    assert np.all(seen == 1)
    # blocks are ordered by (p0, q0): neighbouring threads share antennas
    keys = [tuple(r[:2]) for r in plan]
    assert keys == sorted(keys)


def test_pair_plan_duplicates_time_variation_and_bad_indices():
    from montblanc_amd import _lib
    a1 = np.array([[0, 0, 1, 0]], np.int32)
    a2 = np.array([[1, 2, 2, 1]], np.int32)        # baseline (0, 1) twice
    plan = _lib.build_pair_plan(a1, a2, 3, 2, 2)
    got = sorted(int(b) for b in plan[:, 2:].ravel() if b >= 0)
    assert got == [0, 1, 2, 3]
    varying = np.array([[0, 1], [1, 0]], np.int32)
    assert _lib.build_pair_plan(varying, varying + 1, 3, 2, 2) is None
    with pytest.raises(_lib.MontblancHipError):
        _lib.build_pair_plan(a1, a2 + 5, 3, 2, 2)


def test_baseline_map_codes():
    from montblanc_amd import _lib
    from oracle import rime_np
    na = 5
    a1, a2 = rime_np.default_ant_pairs(na, 2, auto_correlations=True)
    m = _lib.build_baseline_map(a1, a2, na)
    for b in range(a1.shape[1]):
        p, q = a1[0, b], a2[0, b]
        if p == q:
            assert m[p, p] == b | _lib.BL_AUTO
        else:
            assert m[p, q] == b and m[q, p] == b | _lib.BL_REVERSED
    assert (m >= 0).all()
    a1, a2 = rime_np.default_ant_pairs(na, 2)
    m = _lib.build_baseline_map(a1, a2, na)
    assert (np.diag(m) == -1).all()
    varying = a1.copy()
    varying[1] = varying[1][::-1]
    with pytest.raises(_lib.MontblancHipError):
        _lib.build_baseline_map(varying, a2, na)
