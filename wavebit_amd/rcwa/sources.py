"""Plane-wave source declaration (reference ``src/wavebit/rcwa/sources.py:7-44``).

``polarization`` is ``'s'``, ``'p'`` or a complex ``(s, p)`` Jones pair normalised to unit
electric magnitude; ``theta``/``phi`` in radians, or direct ``kx``/``ky`` in radians per length
unit; ``side='bottom'`` lights the stack from the bottom half-space; ``amplitude`` scales the
sampled fields (V/m).
"""
from dataclasses import dataclass

from ..errors import ValidationError


@dataclass(frozen=True)
class PlaneWave:
    side: str = 'top'
    polarization: object = 's'
    amplitude: object = 1.
    theta: object = None
    phi: object = None
    kx: object = None
    ky: object = None

    def __post_init__(self):
        if self.side not in ('top', 'bottom'):
            raise ValidationError("side must be 'top' or 'bottom'.", code='source.side',
                                  path='source.side')
        if (self.theta is not None or self.phi is not None) and (self.kx is not None or self.ky is not None):
            raise ValidationError('Use theta/phi or kx/ky, not both.',
                                  code='source.direction_ambiguous', path='source')
        if isinstance(self.polarization, str) and self.polarization not in ('s', 'p'):
            raise ValidationError("polarization must be 's', 'p', or an (s,p) Jones pair.",
                                  code='source.polarization', path='source.polarization')
        if isinstance(self.polarization, list):
            object.__setattr__(self, 'polarization', tuple(self.polarization))
