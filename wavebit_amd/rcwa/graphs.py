"""Captured HIP graphs of the batched solve for launch-bound sizes.

A prepared problem has a static structure (layer count, raster shapes, wavelength count, solve
path, order window: the reference's ``Prepared`` contract, ``_lifecycle.py:698-733``), so every
``Prepared.run`` issues the same launch sequence on buffers of the same sizes. For small
problems that sequence is dominated by launch overhead (a few hundred kernels of a few
microseconds), so the solve is captured once into a HIP graph over static device buffers and
later runs copy their inputs into those buffers and replay the graph. Large problems keep the
eager launch sequence: their kernels run for milliseconds, launch overhead is below 1% and a
captured graph would pin a large workspace.

The solve allocates nothing and issues no memset or synchronization (all clears are fill
kernels), which is what makes the sequence capturable.
"""
import torch

from .. import _native

# workspace bytes up to which a prepared solve is captured
GRAPH_WORKSPACE_LIMIT = 256 << 20


class SolveGraph:
    """One captured ``wb_solve`` launch sequence over static buffers.

    ``pts`` (P, PT_STRIDE) float64, ``thick`` (P, L) float64, ``idx`` (P, L) int32 and the conv
    matrices (C, N, N) complex128 fix the shapes; ``run`` takes new tensors of those shapes.
    """

    def __init__(self, pts, thick, idx, epx, epy, epz_i, hu, hv, mode):
        device = epx.device
        self.device = device
        self.hu, self.hv, self.mode = hu, hv, mode
        P, L = thick.shape
        self.key = (tuple(pts.shape), tuple(thick.shape), tuple(epx.shape), hu, hv, mode)
        self.pts = pts.detach().clone()
        self.thick = thick.detach().clone()
        self.idx = idx.detach().clone()
        self.epx = epx.detach().clone()
        self.epy = epy.detach().clone()
        self.epz = epz_i.detach().clone()
        self.ws = torch.empty(max(_native.solve_workspace_bytes(P, L, hu, hv, mode), 1),
                              dtype=torch.uint8, device=device)
        # one eager call on the capture stream sizes every lazily initialised launch parameter
        stream = torch.cuda.Stream(device)
        stream.wait_stream(torch.cuda.current_stream(device))
        with torch.cuda.stream(stream):
            _native.solve(self.pts, self.thick, self.idx, self.epx, self.epy, self.epz, hu, hv,
                          mode, self.ws)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=stream):
            self.out = _native.solve(self.pts, self.thick, self.idx, self.epx, self.epy, self.epz,
                                     hu, hv, mode, self.ws)
        torch.cuda.current_stream(device).wait_stream(stream)

    def matches(self, pts, thick, epx, hu, hv, mode):
        return self.key == (tuple(pts.shape), tuple(thick.shape), tuple(epx.shape), hu, hv, mode)

    def run(self, pts, thick, idx, epx, epy, epz_i):
        """Copy the inputs into the static buffers, replay, return (eff, amp, pe, flags) copies."""
        self.pts.copy_(pts)
        self.thick.copy_(thick)
        self.idx.copy_(idx)
        self.epx.copy_(epx)
        self.epy.copy_(epy)
        self.epz.copy_(epz_i)
        self.graph.replay()
        P, L = self.thick.shape
        N = (2 * self.hu + 1) * (2 * self.hv + 1)
        return _native.packed_views(self.out[0]._wb_packed[0].clone(), P, L, N)


def worth_capturing(P, L, hu, hv, mode):
    """True for launch-bound sizes (small workspaces)."""
    return _native.solve_workspace_bytes(P, L, hu, hv, mode) <= GRAPH_WORKSPACE_LIMIT
