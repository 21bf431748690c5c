"""Host-side argument checks of the tensor API (montblanc_amd/rime.py).

These run on CPU tensors: every check fires before the library is asked to
launch anything, so a wrong dtype, device, layout or shape never reaches a
kernel as a raw pointer.
"""

import numpy as np
import pytest
import torch

from montblanc_amd import rime
from oracle import synth

COUNTS = (3, 2, 1)


def cpu_arrays(dtype=np.float64):
    a = synth.make_inputs(5, 2, 3, *COUNTS, beam=(6, 6, 4), seed=3)
    return rime.to_device(synth.as_dtype(a, dtype), device='cpu')


def test_to_device_converts_torch_inputs_like_numpy_ones():
    a = synth.make_inputs(5, 2, 3, *COUNTS, beam=(6, 6, 4), seed=3)
    t = {k: (torch.from_numpy(v.astype(np.int64)) if k in ('antenna1', 'antenna2', 'flag')
             else torch.from_numpy(v)) if isinstance(v, np.ndarray) else v
         for k, v in a.items()}
    out = rime.to_device(t, device='cpu', dtype=torch.float32)
    assert out['antenna1'].dtype == torch.int32 and out['antenna2'].dtype == torch.int32
    assert out['flag'].dtype == torch.uint8
    assert out['uvw'].dtype == torch.float32 and out['E_beam'].dtype == torch.complex64
    assert out['G_term'].dtype == torch.complex64
    np.testing.assert_array_equal(out['antenna2'].numpy(), a['antenna2'])
    ref = rime.to_device(a, device='cpu', dtype=np.float32)
    for k, v in ref.items():
        if isinstance(v, torch.Tensor):
            assert v.dtype == out[k].dtype, k
            assert torch.equal(v, out[k]), k
    rime.check_arrays(out, COUNTS)


def test_check_arrays_accepts_consistent_inputs():
    for dt in (np.float32, np.float64):
        rime.check_arrays(cpu_arrays(dt), COUNTS)


@pytest.mark.parametrize('name,bad,match', [
    ('antenna1', lambda t: t.to(torch.int64), 'dtype'),
    ('flag', lambda t: t.to(torch.int32), 'dtype'),
    ('lm', lambda t: t.to(torch.float32), 'dtype'),
    ('observed_vis', lambda t: t.to(torch.complex64), 'dtype'),
    ('stokes', lambda t: t[:-1].contiguous(), 'shape'),
    ('G_term', lambda t: t.transpose(1, 2), 'contiguous'),
    ('point_errors', lambda t: t[:, :, :1].contiguous(), 'shape'),
    ('weight_vector', lambda t: t[:1].contiguous(), 'shape'),
    ('gauss_shape', lambda t: t[:, :1].contiguous(), 'shape'),
    ('uvw', lambda t: t.to(torch.complex128), 'float32 or float64'),
])
def test_check_arrays_rejects(name, bad, match):
    a = cpu_arrays()
    a[name] = bad(a[name])
    with pytest.raises(ValueError, match=match):
        rime.check_arrays(a, COUNTS)


def test_forward_and_adjoint_check_before_any_launch():
    """rime.forward / rime.adjoint raise the ValueError on CPU tensors: the
    check precedes the library call (which would need a GPU)."""
    a = cpu_arrays()
    a['antenna2'] = a['antenna2'].to(torch.int64)
    with pytest.raises(ValueError, match='antenna2'):
        rime.forward(a, COUNTS)
    a = cpu_arrays()
    with pytest.raises(ValueError, match='counts'):
        rime.forward(a, (3, 2))
    a = cpu_arrays()
    ekb = torch.zeros((2, 10, 3, 4), dtype=torch.complex64)
    with pytest.raises(ValueError, match='ekb_vis'):
        rime.adjoint(a, COUNTS, ekb)
    a = cpu_arrays()
    d = rime.make_dims(a, COUNTS)
    d.nchan.local_size = 5
    with pytest.raises(ValueError, match='nchan'):
        rime.forward(a, COUNTS, dims=d)


def test_adjoint_rejects_time_varying_layouts_up_front():
    a = cpu_arrays()
    a1, a2 = a['antenna1'].clone(), a['antenna2'].clone()
    a1[1], a2[1] = a2[1].clone(), a1[1].clone()      # the second time step reversed
    a1[1, 0], a2[1, 0] = 0, 2
    a['antenna1'], a['antenna2'] = a1, a2
    ekb = torch.zeros((2, 10, 3, 4), dtype=torch.complex128)
    with pytest.raises(ValueError, match='same at every time step'):
        rime.adjoint(a, COUNTS, ekb)
