"""Length units at the public boundary (reference ``src/wavebit/units.py``)."""
from dataclasses import dataclass

from .errors import ValidationError

_LENGTH_METRES = {'m': 1.0, 'mm': 1e-3, 'um': 1e-6, 'nm': 1e-9}
# SI constants as the reference states them (units.py:7-9): the vacuum impedance converts the
# engine's H (in units of E) to A/m
C0 = 299792458.0
EPSILON0 = 8.8541878188e-12
Z0 = 1.0 / (EPSILON0 * C0)


@dataclass(frozen=True, kw_only=True)
class Units:
    length: str = 'm'

    def __post_init__(self):
        if not isinstance(self.length, str) or self.length not in _LENGTH_METRES:
            raise ValidationError("length must be 'm', 'mm', 'um', or 'nm'", code='units.length',
                                  path='units.length', actual=self.length)

    @property
    def length_unit_meters(self):
        return _LENGTH_METRES[self.length]
