"""Immutable physical declarations of a periodic layered RCWA problem.

Same names, fields and meaning as reference ``src/wavebit/rcwa/problem.py:341-494``: layers are
listed bottom-to-top and the bottom interface is z=0; ``Layer.epsilon`` is a sampled relative-
permittivity raster ``(Ny, Nx)`` (rows along y, columns along x); arrays are borrowed leaves
that callers must not mutate during a calculation. This engine solves raster layers and
uniform (background-only) layers; region/shape layers are declared but refused at admission
as unsupported.
"""
from dataclasses import dataclass, field

from ..errors import ValidationError
from ..materials import Material
from ..units import Units


def _freeze(value):
    return tuple(_freeze(v) for v in value) if isinstance(value, (tuple, list)) else value


@dataclass(frozen=True)
class Region:
    """A shape region of one material (declared for API parity; refused by this engine)."""
    geometry: object
    material: object


@dataclass(frozen=True)
class Layer:
    """A layer of given thickness: a sampled ``epsilon`` raster, or a uniform background."""
    thickness: object
    background: object = None
    regions: tuple = ()
    epsilon: object = None
    mu: object = None

    def __post_init__(self):
        object.__setattr__(self, 'regions', tuple(self.regions))
        object.__setattr__(self, 'epsilon', _freeze(self.epsilon))
        if not all(isinstance(r, Region) for r in self.regions):
            raise ValidationError('regions must contain Region declarations.',
                                  code='layer.regions', path='layer.regions')
        if self.epsilon is not None:
            if self.background is not None or self.regions:
                raise ValidationError('A sampled layer cannot also declare a background or regions.',
                                      code='layer.ambiguous', path='layer')
        elif self.mu is not None:
            raise ValidationError('Put permeability in each Material; layer.mu belongs to sampled epsilon.',
                                  code='layer.mu_without_raster', path='layer.mu')
        elif self.background is None:
            object.__setattr__(self, 'background', Material(epsilon=1.))


@dataclass(frozen=True)
class Stack:
    """Bottom-to-top finite layers between two half-spaces; ``period=(Lx, Ly)``."""
    layers: tuple = ()
    top: object = field(default_factory=lambda: Material(epsilon=1.))
    bottom: object = field(default_factory=lambda: Material(epsilon=1.))
    period: object = None
    lattice: object = None

    def __post_init__(self):
        object.__setattr__(self, 'layers', tuple(self.layers))
        if not all(isinstance(layer, Layer) for layer in self.layers):
            raise ValidationError('layers must contain rcwa.Layer declarations.',
                                  code='stack.layers', path='stack.layers')
        if self.period is not None and self.lattice is not None:
            raise ValidationError('Supply either period or lattice, not both.',
                                  code='stack.cell_ambiguous', path='stack')
        if self.period is None and self.lattice is None:
            object.__setattr__(self, 'period', (1., 1.))
        object.__setattr__(self, 'period', _freeze(self.period))
        object.__setattr__(self, 'lattice', _freeze(self.lattice))


@dataclass(frozen=True)
class Excitation:
    """Vacuum wavelength(s) and the sources that own direction and polarization."""
    wavelength: object
    sources: tuple

    def __post_init__(self):
        from .sources import PlaneWave
        object.__setattr__(self, 'sources', tuple(self.sources))
        object.__setattr__(self, 'wavelength', _freeze(self.wavelength))
        if not self.sources or not all(isinstance(s, PlaneWave) for s in self.sources):
            raise ValidationError('sources must be a nonempty tuple of RCWA source declarations.',
                                  code='excitation.sources', path='excitation.sources')


@dataclass(frozen=True)
class Outputs:
    """Requested observations; fields and scattering matrices are opt-in."""
    diffraction: bool = True
    fields: bool = False
    scattering: bool = False
    power_region: object = None

    def __post_init__(self):
        for name in ('diffraction', 'fields', 'scattering'):
            if type(getattr(self, name)) is not bool:
                raise ValidationError(f'{name} must be a boolean.', code='outputs.boolean',
                                      path=f'outputs.{name}')


@dataclass(frozen=True)
class Problem:
    """The complete physical RCWA problem; no backend or mutable solver state."""
    stack: Stack
    excitation: Excitation
    units: Units = field(default_factory=Units)
    outputs: Outputs = field(default_factory=Outputs)

    def __post_init__(self):
        for name, cls in (('stack', Stack), ('excitation', Excitation), ('units', Units),
                          ('outputs', Outputs)):
            if not isinstance(getattr(self, name), cls):
                raise ValidationError(f'{name} must be {cls.__name__}.',
                                      code='problem.type', path=name)
