"""HIP-graph capture of repeated solves.

``Graphed(fn)`` runs ``fn`` (any sequence of library calls that launch on
torch's current stream, e.g. ``lambda: rime.forward(arrays, counts,
outputs=out)``) a few times eagerly -- which allocates workspaces, builds
plans and fills autotuner caches -- then captures one call into a
torch.cuda.CUDAGraph; ``replay()`` relaunches the whole recorded sequence
with one host call. The library's entry points launch kernels only (no
allocation, no memset: buffers are cleared by fill kernels), so they are
safe to capture; inputs are read from, and results written to, the same
device buffers on every replay.
"""

import torch


class Graphed(object):
    def __init__(self, fn, device=None, warmup=2):
        device = torch.device('cuda', torch.cuda.current_device()) if device is None else device
        side = torch.cuda.Stream(device)
        side.wait_stream(torch.cuda.current_stream(device))
        with torch.cuda.stream(side):
            for _ in range(warmup):
                fn()
        torch.cuda.current_stream(device).wait_stream(side)
        torch.cuda.synchronize(device)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.result = fn()
        self.device = device

    def replay(self):
        self.graph.replay()
        return self.result
