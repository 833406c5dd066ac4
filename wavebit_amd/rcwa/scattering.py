"""The stack's scattering matrix (``Outputs(scattering=True)``): reference
``src/wavebit/rcwa/_engine/scattering.py:371-604`` (``Port``, ``ScatteringMatrix``,
``cascade``) and ``:658-697`` (``_port_disagreement``).

The device solve (``wb_solve_scattering``) returns the four blocks already published: top
incidence, incidence-relative order (``S11`` reflects back into the top medium, ``S21``
transmits top to bottom, ``S12`` bottom to top, ``S22`` reflects into the bottom medium), every
port amplitude referenced at the cell centre, the 2D blocks over ``n_orders`` s amplitudes then
``n_orders`` p port variables (the p wave's magnetic field), the pure 1D blocks in the same
port variables (``channel`` 'te' or 'tm'). This module carries them with their ports and joins
two of them with the Redheffer star product after checking that the ports meet.
"""
from typing import NamedTuple

import numpy as np

_CHANNELS = {
    'te': 'the pure 1D TE path, one s amplitude per order',
    'tm': 'the pure 1D TM path, one p port variable per order',
    'sp': 'a coupled path, s amplitudes over p port variables',
}


class Port(NamedTuple):
    """The plane waves one end of a ``ScatteringMatrix`` counts: the retained orders' in-plane
    wavevectors over k0 (``ky_over_k0`` is None on the pure 1D paths), the half-space's
    refractive index and permeability, and the polarization channel."""
    kx_over_k0: object
    ky_over_k0: object
    refractive_index: object
    mu: object
    channel: str


def _differ(x, y, tol=1e-12):
    if (x is None) != (y is None):
        return True
    if x is None:
        return False
    xa, ya = np.asarray(x), np.asarray(y)
    return xa.shape != ya.shape or bool(np.max(np.abs(xa - ya)) > tol)


def port_disagreement(a, b):
    """Why ports ``a`` and ``b`` are not the same basis, or None when they are."""
    if a.channel != b.channel:
        words = lambda c: _CHANNELS.get(c, 'a channel no solve produces')   # noqa: E731
        return (f'they count different polarizations: {a.channel!r} ({words(a.channel)}) against '
                f'{b.channel!r} ({words(b.channel)})')
    if _differ(a.kx_over_k0, b.kx_over_k0) or _differ(a.ky_over_k0, b.ky_over_k0):
        return ('their retained orders are different in-plane wavevectors; `theta` is measured in '
                'the incidence medium, so choose the angles for the same transverse wavevector on '
                'both sides of the cut')
    if _differ(a.refractive_index, b.refractive_index) or _differ(a.mu, b.mu):
        return 'they open into different half-spaces; both pieces have to name the same joint medium'
    return None


def redheffer_star(S_a, S_b):
    """``S_a`` followed by ``S_b`` (blocks ``(S11, S12, S21, S22)``; reference
    ``scattering.py:191-200``)."""
    a11, a12, a21, a22 = S_a
    b11, b12, b21, b22 = S_b
    I = np.eye(a11.shape[0], dtype=a11.dtype)
    D = np.linalg.solve((I - b11 @ a22).T, a12.T).T           # a12 (I - b11 a22)^-1
    F = np.linalg.solve((I - a22 @ b11).T, b21.T).T           # b21 (I - a22 b11)^-1
    return a11 + D @ b11 @ a21, D @ b12, F @ a21, b22 + F @ a22 @ b12


class ScatteringMatrix(NamedTuple):
    """The stack's scattering matrix, both sides, every retained order, with its two ports."""
    S11: object
    S12: object
    S21: object
    S22: object
    entry_port: Port
    exit_port: Port

    @property
    def blocks(self):
        return (self.S11, self.S12, self.S21, self.S22)

    def cascade(self, beyond):
        """This stack followed by ``beyond`` -- the piece the light reaches next -- as one
        scattering matrix; refused when this exit port and ``beyond``'s entry port differ."""
        if not isinstance(beyond, ScatteringMatrix):
            raise TypeError(f'cascade() takes a ScatteringMatrix, got {type(beyond).__name__}.')
        why = port_disagreement(self.exit_port, beyond.entry_port)
        if why is not None:
            raise ValueError(f'cascade() cannot join these two: {why}.')
        return ScatteringMatrix(*redheffer_star(self.blocks, beyond.blocks),
                                entry_port=self.entry_port, exit_port=beyond.exit_port)
