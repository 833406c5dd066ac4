"""Approximation policy (reference ``src/wavebit/rcwa/numerics.py:63-164``).

``orders`` is the retained Fourier window ``(Mu, Mv)``. The engine implements the reference's
defaults for raster layers -- SMM, sampled coefficients, Li's rules with fixed nesting, the
second-order formulation -- plus Laurent's rule (``factorization='laurent'``) and integrated
coefficients (``fourier_coefficients='integrated'``: on a raster the closed-form segment
integral of the same staircase, which the pixel series already evaluates exactly), and refuses
other choices at admission with ``UnsupportedFeatureError``.
"""
from dataclasses import dataclass

from ..errors import ValidationError


def order_pair(value, path='numerics.orders'):
    if isinstance(value, (int,)) and not isinstance(value, bool):
        value = (value, 0)
    try:
        pair = tuple(value)
    except TypeError:
        raise ValidationError('orders must be a pair of nonnegative integers.',
                              code='numerics.orders', path=path) from None
    if len(pair) != 2 or not all(type(v) is int and v >= 0 for v in pair):
        raise ValidationError('orders must be a pair of nonnegative integers.',
                              code='numerics.orders', path=path, actual=repr(value))
    return pair


@dataclass(frozen=True)
class Numerics:
    orders: tuple = (0, 0)
    algorithm: str = 'smm'
    fourier_coefficients: object = None
    factorization: object = None
    formulation: str = 'auto'
    raster_normal_method: object = None
    nvf_projector: object = None
    shape_nvf_grid_size: object = None
    li_nesting: object = None
    detuning: object = None
    topology_check: str = 'device'
    half_space_merge_tolerance: float = 0.0
    exceptional_points: str = 'rebuild'
    wall_resolution: object = None
    shape_normal_method: object = None
    slicing: object = None
    # written to a job only off their default (reference numerics.py:104-105)
    _FIELDS_OMITTED_AT_DEFAULT = ('half_space_merge_tolerance', 'exceptional_points',
                                  'wall_resolution', 'shape_normal_method', 'slicing')

    def __post_init__(self):
        object.__setattr__(self, 'orders', order_pair(self.orders))
        choices = {'algorithm': ('smm', 'etm'),
                   'fourier_coefficients': (None, 'sampled', 'integrated'),
                   'factorization': (None, 'li', 'nvf', 'laurent'),
                   'formulation': ('auto', 'first_order', 'second_order'),
                   'raster_normal_method': (None, 'distance', 'contour', 'spectral', 'smooth'),
                   'nvf_projector': (None, 'linear', 'circular'),
                   'li_nesting': (None, 'fixed', 'symmetric'),
                   'shape_normal_method': (None, 'per_shape', 'interfaces'),
                   'topology_check': ('device', 'host'),
                   'exceptional_points': ('rebuild', 'off')}
        for name, allowed in choices.items():
            if getattr(self, name) not in allowed:
                raise ValidationError(f'{name} must be one of {allowed}.',
                                      code='numerics.choice', path=f'numerics.{name}')
        if self.detuning is not None:
            import math
            if type(self.detuning) not in (float, int) or not math.isfinite(self.detuning) or self.detuning <= 0:
                raise ValidationError('detuning is None (exact requested problem) or a positive relative step.',
                                      code='numerics.detuning', path='numerics.detuning')

    def unsupported(self):
        """The first non-default choice this engine does not implement, as (field, value)."""
        for name, ok in (('algorithm', ('smm',)), ('fourier_coefficients', (None, 'sampled', 'integrated')),
                         ('factorization', (None, 'li', 'laurent', 'nvf')),
                         ('formulation', ('auto', 'second_order')),
                         ('li_nesting', (None, 'fixed')),
                         ('raster_normal_method', (None, 'distance')), ('nvf_projector', (None, 'linear')),
                         ('shape_nvf_grid_size', (None,)), ('half_space_merge_tolerance', (0.0,)),
                         ('wall_resolution', (None,)), ('shape_normal_method', (None,)),
                         ('slicing', (None,))):
            if getattr(self, name) not in ok:
                return name, getattr(self, name)
        return None
