"""Differentiable raster RCWA: one ``torch.autograd.Function`` over the HIP forward and its
hand-written reverse pass.

The reference differentiates its PyTorch path with stock autograd through FFTs, inverses and
GEMMs, plus one custom rule for the eigensolve (``rcwa/_engine/modal_gradient.py:161-251``,
the generator) and divided differences for the layer exponential (``:446-470``); the backward
retains every n x n intermediate (reference ``Prepared.run`` then ``loss.backward()``,
``_lifecycle.py:785-786``). Here the forward is the batched device solve with a checkpoint
(per layer: eigenvectors, the two basis inverses and the layer S-matrix; the cascade's factors
stay in the solve workspace), and the backward is ``wb_solve_adjoint`` followed by
``wb_assemble_li_adjoint`` (``wb_assemble_laurent_adjoint`` under Laurent's rule,
``wb_assemble_nvf_adjoint`` under the normal-vector-field rule): the gradient of the
efficiencies with respect to every raster pixel and every layer thickness, one launch sequence
for all wavelengths of the problem.

Under NVF the normal field is held fixed: it is the geometry of the interfaces (where the
raster's values change), which an infinitesimal change of the values does not move -- the
reference's gradient treats it the same way (``tests/test_frozen_geometry.py:644``, "the NVF
gradient holds the interfaces fixed"). Only stacks of passive dielectrics (Re eps > 0,
Im eps >= 0 everywhere) are differentiated there: on the others the reference floors the
singular values of [[1/eps]] (``factorization.py:1968-2160``), a step this reverse pass does
not carry, and ``forward`` refuses them.

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
    def forward(ctx, eps, thick, pts, hu, hv, factorization='li', lattice=None):
        eps_c = eps.detach().to(torch.complex128).contiguous()
        thick_d = thick.detach().to(torch.float64).contiguous()
        L = eps_c.shape[0]
        P = pts.shape[0]
        if factorization == 'nvf':
            if not passive(eps_c):
                raise ValueError("gradients under factorization='nvf' need passive dielectric "
                                 'layers (Re eps > 0, Im eps >= 0 everywhere)')
            epx, epy, epz, epxy, info = _native.assemble_nvf(eps_c, hu, hv, lattice, floor=False)
        else:
            epx, epy, epz, epxy, info = _native.assemble(eps_c, hu, hv, factorization)
        thick_pl = thick_d.reshape(1, L).expand(P, L).contiguous()
        conv_idx = torch.arange(L, dtype=torch.int32, device=eps_c.device).reshape(1, L).expand(P, L).contiguous()
        eff, amp, pe, flags, state = _native.solve_keep(pts, thick_pl, conv_idx, epx, epy, epz, hu, hv,
                                                        _native.MODE_2D, epxy=epxy, epyx=epxy)
        ctx.state = state
        ctx.hu, ctx.hv = hu, hv
        ctx.factorization = factorization
        ctx.lattice = lattice
        ctx.eps_c = eps_c
        ctx.assembly_info = info
        ctx.eps_is_complex = eps.is_complex()
        ctx.mark_non_differentiable(amp, pe, flags)
        return eff, amp, pe, flags

    @staticmethod
    def backward(ctx, g_eff, g_amp, g_pe, g_flags):
        state = ctx.state
        if g_eff is None:
            return None, None, None, None, None, None, None
        d = _native.solve_adjoint(state, g_eff.contiguous())
        d_epx, d_epy, d_epz, d_thick = d[0], d[1], d[2], d[-1]
        d_eps = None
        if ctx.needs_input_grad[0]:
            if ctx.factorization == 'nvf':
                d_eps = _native.assemble_nvf_adjoint(ctx.eps_c, ctx.hu, ctx.hv, d_epx, d_epy, d_epz,
                                                     d[3] + d[4], ctx.lattice)
            elif ctx.factorization == 'laurent':
                d_eps = _native.assemble_laurent_adjoint(ctx.eps_c, ctx.hu, ctx.hv, state['epz_i'],
                                                         d_epx, d_epy, d_epz)
            else:
                d_eps = _native.assemble_li_adjoint(ctx.eps_c, ctx.hu, ctx.hv, state['epx'],
                                                    state['epz_i'], d_epx, d_epy, d_epz)
            if not ctx.eps_is_complex:
                d_eps = d_eps.real
        d_t = d_thick.sum(dim=0) if ctx.needs_input_grad[1] else None
        ctx.state = None            # the workspace and checkpoint are released with the graph
        return d_eps, d_t, None, None, None, None, None


def passive(eps):
    """True when every value of ``eps`` (tensor or array) is a passive dielectric, Re > 0 and
    Im >= 0: the stacks NVF's singular-value floor leaves alone."""
    if torch.is_tensor(eps):
        e = eps.detach()
        if e.is_complex():
            return bool(torch.all((e.real > 0) & (e.imag >= 0)))
        return bool(torch.all(e > 0))
    e = np.asarray(eps)
    return bool(np.all((e.real > 0) & (e.imag >= 0)))


def point_records(points, device):
    """Per-point parameter records (P, 16) float64 on ``device`` for the 2D path."""
    recs = np.array([_engine.point_record(p, _engine.PATH_2D) for p in points], dtype=np.float64)
    return torch.tensor(recs, dtype=torch.float64, device=device)


def solve(eps, thick, points, hu, hv, device, factorization='li', lattice=None):
    """Differentiable solve: (eff, amp, pe, flags) with ``eff`` carrying the graph to ``eps`` and
    ``thick``. ``eps`` (L, Ny, Nx) and ``thick`` (L,) may be leaves on any device; they are moved
    to ``device`` inside the graph. ``lattice``: the cell basis the NVF normal field is measured
    in (2 x 2, columns a1 and a2; square pixels when None)."""
    device = torch.device(device)
    _native.context(device)
    eps_d = eps.to(device)
    thick_d = thick.to(device=device, dtype=torch.float64)
    pts = point_records(points, device)
    return RasterSolve.apply(eps_d, thick_d, pts, hu, hv, factorization, lattice)
