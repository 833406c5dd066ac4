"""Warnings of the RCWA path (reference public path ``wavebit.rcwa.diagnostics``)."""


class ExceptionalPointWarning(UserWarning):
    """A 2D layer's modes hold a grazing mode or a near-defective pair that the reference would
    rebuild (reference ``_common/backend.py:1120``, ``rcwa/_engine/solve_2d.py:583-660,
    2463-2485``); this engine keeps the modes as computed and says so."""


class RayleighPointWarning(UserWarning):
    """A layer mode sits numerically on its light line and a detuning ladder did not move it off
    (reference ``_common/backend.py:1151``)."""


class AutoDetuneWarning(UserWarning):
    """``Numerics(detuning=step)`` moved some points off the requested wavelength (and angle)
    -- this is the record (reference ``_common/backend.py:1253``)."""
