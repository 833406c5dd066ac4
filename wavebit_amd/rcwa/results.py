"""Result containers (subset of reference ``src/wavebit/rcwa/results.py:93-487`` and
``src/wavebit/observations/contracts.py:288``).

``Result.R`` / ``Result.T`` are per-order efficiencies of shape (wavelength, order_v, order_u);
``Result.reflection`` / ``Result.transmission`` are ``DiffractionOrders`` with the Fresnel
amplitude ratios in each order's own s/p frame, centre-referenced, as the reference reports
them. Every array is an independent host (NumPy) snapshot: a later solve never changes an
earlier result.
"""
from dataclasses import dataclass, field, replace
from types import MappingProxyType

import numpy as np

from ..errors import UnsupportedFeatureError


_ORDER_ARRAYS = ('amplitude_ratio_s', 'amplitude_ratio_p', 'efficiency', 'efficiency_s', 'efficiency_p',
                 'k_parallel')


@dataclass(frozen=True)
class DiffractionOrder:
    """One order ``(order_u, order_v)`` of a ``DiffractionOrders`` table: each array at that
    order, over the wavelength axis."""
    wavelengths: object
    order_u: int
    order_v: int
    amplitude_ratio_s: object
    amplitude_ratio_p: object
    efficiency: object
    efficiency_s: object
    efficiency_p: object
    k_parallel: object = None
    reference: dict = field(default_factory=dict)


@dataclass(frozen=True)
class DiffractionOrders:
    wavelengths: object
    orders_u: object
    orders_v: object
    amplitude_ratio_s: object
    amplitude_ratio_p: object
    efficiency: object
    efficiency_s: object
    efficiency_p: object
    k_parallel: object = None
    reference: dict = field(default_factory=dict)

    def order(self, m, n=0):
        """Order ``(m, n)`` as one record (reference ``observations/contracts.py:257-264``):
        indexed by the order labels, not the array position; ``KeyError`` when not retained."""
        try:
            i = [int(v) for v in self.orders_u].index(int(m))
            j = [int(v) for v in self.orders_v].index(int(n))
        except ValueError as exc:
            raise KeyError((m, n)) from exc
        data = {name: None if getattr(self, name) is None else getattr(self, name)[:, j, i]
                for name in _ORDER_ARRAYS}
        return DiffractionOrder(wavelengths=self.wavelengths, order_u=int(m), order_v=int(n),
                                reference=self.reference, **data)

    def at_wavelength_index(self, index):
        """These orders at the wavelength in position ``index``, as a length-1 axis (reference
        ``observations/contracts.py:266-283``)."""
        i = int(index)
        if not 0 <= i < len(self.wavelengths):
            raise IndexError(i)
        data = {name: None if getattr(self, name) is None else getattr(self, name)[i:i + 1]
                for name in _ORDER_ARRAYS}
        return replace(self, wavelengths=self.wavelengths[i:i + 1], **data)


@dataclass(frozen=True)
class Validity:
    checks: dict
    convergence: dict


@dataclass(frozen=True)
class Result:
    """Independent result of one prepared run."""
    outputs: object
    wavelengths: object
    validity: object
    diagnostics: object
    provenance: object
    _scattering: object = None       # tuple of ScatteringMatrix, one per wavelength
    _fields: object = None           # fields.FieldRecord of an Outputs(fields=True) run

    def __post_init__(self):
        for name in ('outputs', 'diagnostics', 'provenance'):
            object.__setattr__(self, name, MappingProxyType(dict(getattr(self, name))))

    def _efficiency(self, port):
        observation = self.outputs.get(port)
        if observation is None or observation.efficiency is None:
            raise UnsupportedFeatureError(
                'No single-incident-port diffraction efficiency was requested or is available for this result.',
                code='outputs.unavailable', path=f'outputs.{port}.efficiency')
        return observation.efficiency

    @property
    def R(self):
        """Reflected efficiencies, shape (wavelength, order_v, order_u)."""
        return self._efficiency('reflected')

    @property
    def T(self):
        """Transmitted efficiencies, shape (wavelength, order_v, order_u)."""
        return self._efficiency('transmitted')

    @property
    def reflection(self):
        return self.outputs['reflected']

    @property
    def transmission(self):
        return self.outputs['transmitted']

    @property
    def absorption(self):
        """Net power absorbed in the finite layers over the incident power, per wavelength:
        ``power_entering - T.sum()`` (reference ``results.py:266-272``,
        ``observations/diffraction.py:92-104``); half-space absorption excluded, not clipped."""
        if 'absorption' not in self.outputs:
            raise UnsupportedFeatureError('Absorption is unavailable for this observation geometry.',
                                          code='outputs.unavailable', path='outputs.absorption')
        return self.outputs['absorption']

    def field(self, *, x, y, z, reconstruction='plain'):
        """Collocated complex phasors at Cartesian sample points, shape
        ``(wavelength, z, y, x, component)``, E in V/m and H in A/m (reference
        ``results.py:149-227``). ``x``, ``y`` from the cell centre and ``z`` up from the bottom
        interface, in the problem's length unit; an integer count samples uniformly instead (the
        cell across, every layer face to face) and the returned ``coordinates`` hold the points
        sampled. Interfaces select the upper layer. Only ``reconstruction='plain'`` is
        implemented."""
        from .fields import field
        return field(self, x, y, z, reconstruction)

    @property
    def scattering(self):
        """The stack's ``ScatteringMatrix`` per wavelength (reference ``results.py:136-141``)."""
        if self._scattering is None:
            raise UnsupportedFeatureError('Request Outputs(scattering=True) with the SMM algorithm before running.',
                                          code='outputs.unavailable', path='outputs.scattering')
        return self._scattering


def pack(per_wavelength, hu, hv, wavelengths, checks, diagnostics, provenance, scattering=None,
         side='top', fields=None):
    """Stack per-wavelength solve outputs (dicts from the engine) into a ``Result``.

    NaN-masks every masked output when a validity check failed (reference
    ``results.py:441-464`` ``admission_scale``), the scattering blocks included.
    """
    orders_u = np.arange(-hu, hu + 1)
    orders_v = np.arange(-hv, hv + 1)
    ws = np.asarray(wavelengths, dtype=np.float64)

    stacked = {}

    def stack(key):
        if key not in stacked:
            stacked[key] = np.stack([np.asarray(d['incident'][key]) for d in per_wavelength])
        return stacked[key]

    finite = all(np.all(np.isfinite(stack(k))) for k in ('r_s', 'r_p', 't_s', 't_p', 'R', 'T'))
    combined = bool(finite and all(bool(v) for v in checks.values()))
    scale = 1.0 if combined else float('nan')
    outputs = {}
    other = 'bottom' if side == 'top' else 'top'
    for name, prefix, eff in (('reflected', 'r', 'R'), ('transmitted', 't', 'T')):
        outputs[name] = DiffractionOrders(
            wavelengths=ws, orders_u=orders_u, orders_v=orders_v,
            amplitude_ratio_s=stack(prefix + '_s') * scale,
            amplitude_ratio_p=stack(prefix + '_p') * scale,
            efficiency=stack(eff) * scale,
            efficiency_s=stack(eff + '_s') * scale,
            efficiency_p=stack(eff + '_p') * scale,
            reference={'port': side if (name == 'reflected') else other, 'wave': name,
                       'incident_port': side, 'polarization_frame': "each order's outgoing s/p frame",
                       'time_convention': 'exp(-i*omega*t)'})
    outputs['power_entering'] = np.array([float(d['incident']['power_entering']) for d in per_wavelength]) * scale
    outputs['absorption'] = outputs['power_entering'] - outputs['transmitted'].efficiency.sum(axis=(1, 2))
    variants = {}
    for v in ('s_inc', 'p_inc'):
        if all(v in d for d in per_wavelength):
            variants[v] = {k: np.stack([np.asarray(d[v][k]) for d in per_wavelength]) * scale
                           for k in ('R', 'T', 'r_s', 'r_p', 't_s', 't_p')}
    diagnostics = dict(diagnostics)
    diagnostics['incidence_variants'] = variants
    if scattering is not None:
        scattering = tuple(m._replace(S11=m.S11 * scale, S12=m.S12 * scale, S21=m.S21 * scale,
                                      S22=m.S22 * scale) for m in scattering)
    if fields is not None and not combined:
        fields.amplitude = complex('nan')      # sampled fields follow the admission like the rest
    return Result(outputs=outputs, wavelengths=ws, _scattering=scattering, _fields=fields,
                  validity=Validity(checks={**checks, 'numerical_finite': combined},
                                    convergence={'status': 'not assessed',
                                                 'requires': 'independent refinement study'}),
                  diagnostics=diagnostics, provenance=provenance)
