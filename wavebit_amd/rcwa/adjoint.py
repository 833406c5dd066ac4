"""Differentiable raster RCWA: one ``torch.autograd.Function`` over the HIP forward and its
hand-written reverse pass.

The reference differentiates its PyTorch path with stock autograd through FFTs, inverses and
GEMMs, plus one custom rule for the eigensolve (``rcwa/_engine/modal_gradient.py:161-251``,
the generator) and divided differences for the layer exponential (``:446-470``); the backward
retains every n x n intermediate (reference ``Prepared.run`` then ``loss.backward()``,
``_lifecycle.py:785-786``). Here the forward is the batched device solve with a checkpoint
(per layer: eigenvectors, the two basis inverses and the layer S-matrix; the cascade's factors
stay in the solve workspace), and the backward is ``wb_solve_adjoint`` followed by
``wb_assemble_li_adjoint`` (``wb_assemble_laurent_adjoint`` under Laurent's rule): the gradient of the efficiencies with respect to every raster
pixel and every layer thickness, one launch sequence for all wavelengths of the problem.

Leaves: the raster stack ``eps`` (L, Ny, Nx), real or complex, and the thicknesses (L,).
Amplitudes, power entering and flags are returned without gradient.
"""
import numpy as np
import torch

from .. import _native
from . import _engine


class RasterSolve(torch.autograd.Function):
    """eff (P, 3, 4, N) of P plane-wave points through a raster stack, 2D coupled path."""

    @staticmethod
    def forward(ctx, eps, thick, pts, hu, hv, factorization='li'):
        eps_c = eps.detach().to(torch.complex128).contiguous()
        thick_d = thick.detach().to(torch.float64).contiguous()
        L = eps_c.shape[0]
        P = pts.shape[0]
        epx, epy, epz, _, info = _native.assemble(eps_c, hu, hv, factorization)
        thick_pl = thick_d.reshape(1, L).expand(P, L).contiguous()
        conv_idx = torch.arange(L, dtype=torch.int32, device=eps_c.device).reshape(1, L).expand(P, L).contiguous()
        eff, amp, pe, flags, state = _native.solve_keep(pts, thick_pl, conv_idx, epx, epy, epz, hu, hv,
                                                        _native.MODE_2D)
        ctx.state = state
        ctx.hu, ctx.hv = hu, hv
        ctx.factorization = factorization
        ctx.eps_c = eps_c
        ctx.assembly_info = info
        ctx.eps_is_complex = eps.is_complex()
        ctx.mark_non_differentiable(amp, pe, flags)
        return eff, amp, pe, flags

    @staticmethod
    def backward(ctx, g_eff, g_amp, g_pe, g_flags):
        state = ctx.state
        if g_eff is None:
            return None, None, None, None, None, None
        d_epx, d_epy, d_epz, d_thick = _native.solve_adjoint(state, g_eff.contiguous())
        d_eps = None
        if ctx.needs_input_grad[0]:
            if ctx.factorization == 'laurent':
                d_eps = _native.assemble_laurent_adjoint(ctx.eps_c, ctx.hu, ctx.hv, state['epz_i'],
                                                         d_epx, d_epy, d_epz)
            else:
                d_eps = _native.assemble_li_adjoint(ctx.eps_c, ctx.hu, ctx.hv, state['epx'],
                                                    state['epz_i'], d_epx, d_epy, d_epz)
            if not ctx.eps_is_complex:
                d_eps = d_eps.real
        d_t = d_thick.sum(dim=0) if ctx.needs_input_grad[1] else None
        ctx.state = None            # the workspace and checkpoint are released with the graph
        return d_eps, d_t, None, None, None, None


def point_records(points, device):
    """Per-point parameter records (P, 16) float64 on ``device`` for the 2D path."""
    recs = np.array([_engine.point_record(p, _engine.PATH_2D) for p in points], dtype=np.float64)
    return torch.tensor(recs, dtype=torch.float64, device=device)


def solve(eps, thick, points, hu, hv, device, factorization='li'):
    """Differentiable solve: (eff, amp, pe, flags) with ``eff`` carrying the graph to ``eps`` and
    ``thick``. ``eps`` (L, Ny, Nx) and ``thick`` (L,) may be leaves on any device; they are moved
    to ``device`` inside the graph."""
    device = torch.device(device)
    _native.context(device)
    eps_d = eps.to(device)
    thick_d = thick.to(device=device, dtype=torch.float64)
    pts = point_records(points, device)
    return RasterSolve.apply(eps_d, thick_d, pts, hu, hv, factorization)
