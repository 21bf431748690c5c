"""The no-allocation forward (mbx_forward_static_*), whose workspace is the
library's own device array, against the forward with a caller workspace
(mbx_forward_*): two calls in a row, the second on a smaller problem, both
bit-identical to the caller-workspace forward (same kernels, same inputs),
and no device memory taken by the call."""

import numpy as np
import pytest
import torch

from oracle import synth

pytestmark = pytest.mark.gpu


def run_both(a, counts, dtype, residual=False):
    from montblanc_amd import rime
    arrays = rime.to_device(synth.as_dtype(a, dtype))
    ref = rime.forward(arrays, counts, residual=residual)
    out = rime.forward(arrays, counts, residual=residual, static_scratch=True)
    torch.cuda.synchronize()
    return ref, out


@pytest.mark.parametrize('dtype', [np.float32, np.float64])
def test_static_scratch_forward_matches_caller_workspace(dtype):
    from montblanc_amd import rime
    big = synth.make_inputs(40, 3, 12, 40, 20, 6, beam=(17, 15, 8), seed=71, flag_fraction=0.3)
    small = synth.make_inputs(6, 2, 3, 3, 2, 1, beam=(5, 4, 3), seed=72, flag_fraction=0.3)
    for a, counts, residual in ((big, (40, 20, 6), False), (small, (3, 2, 1), True)):
        print('static scratch', np.dtype(dtype).name, a['uvw'].shape, counts, flush=True)
        ref, out = run_both(a, counts, dtype, residual)
        for k in ('model_vis', 'chi_sqrd_result', 'X2'):
            assert torch.equal(out[k], ref[k]), k
    # a further call takes no device memory: its workspace is the static array
    arrays = rime.to_device(synth.as_dtype(small, dtype))
    out = rime.alloc_outputs(arrays)
    rime.forward(arrays, (3, 2, 1), outputs=out, static_scratch=True)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    rime.forward(arrays, (3, 2, 1), outputs=out, static_scratch=True)
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0


def test_static_scratch_rejects_oversized_workspace():
    from montblanc_amd import _lib, rime
    a = synth.make_inputs(4, 1, 256, 1, 0, 0, beam=(129, 129, 4), seed=73)
    arrays = rime.to_device(synth.as_dtype(a, np.float32))
    with pytest.raises(_lib.MontblancHipError) as e:
        rime.forward(arrays, (1, 0, 0), static_scratch=True)
    assert e.value.code == -4
