"""Warnings of the RCWA path (reference public path ``wavebit.rcwa.diagnostics``)."""


class ExceptionalPointWarning(UserWarning):
    """A 2D layer's modes hold a grazing mode or a near-defective pair that the reference would
    rebuild (reference ``_common/backend.py:1120``, ``rcwa/_engine/solve_2d.py:583-660,
    2463-2485``); this engine keeps the modes as computed and says so."""
