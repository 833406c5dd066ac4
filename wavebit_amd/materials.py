"""Constant-permittivity materials (the subset of reference ``src/wavebit/materials/`` the
raster RCWA configurations use): ``Material(epsilon)`` evaluated ``.at(wavelength)``.

Dispersive and tabulated models are out of scope for this engine.
"""
from dataclasses import dataclass

from .errors import ValidationError, UnsupportedFeatureError


@dataclass(frozen=True)
class Response:
    epsilon: complex
    mu: object = None


@dataclass(frozen=True)
class Material:
    """A non-dispersive, non-magnetic medium of relative permittivity ``epsilon``."""
    epsilon: object = 1.0
    mu: object = None

    def __post_init__(self):
        try:
            complex(self.epsilon)
        except (TypeError, ValueError):
            raise ValidationError('Material epsilon must be a number.', code='material.epsilon',
                                  path='material.epsilon', actual=repr(self.epsilon)) from None
        if self.mu is not None and complex(self.mu) != 1:
            raise UnsupportedFeatureError('Magnetic materials are not implemented by this engine.',
                                          code='rcwa.unsupported', path='material.mu')

    def at(self, wavelength, units=None):
        return Response(epsilon=complex(self.epsilon), mu=None)
