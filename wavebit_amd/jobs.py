"""Portable RCWA jobs in the reference's ``wavebit.job/v2`` wire format (reference
``src/wavebit/jobs.py:1-264``), so the serving loop and ``run`` accept the reference's own
job files.

The envelope is ``{"schema": "wavebit.job/v2", "solver", "problem", "numerics", "execution"}``
with every value tagged: scalars as JSON, complex numbers ``{"type": "complex", ...}``, arrays
``{"type": "array", "dtype", "shape", "data": base64}``, tuples / lists / mappings, and each
declaration ``{"type": "<public path>", "fields": {...}}`` under the reference's public paths
(``wavebit.rcwa.Problem``, ``wavebit.materials.Material``, ``wavebit.Execution`` ...), which
are the names written here too. Decoding never imports a class named by its input: the tags
resolve through a fixed table. Declarations of the reference that this engine does not
implement (shape regions, other sources, dispersive materials, the FDTD solvers) are refused
by name; ``Numerics`` fields off their defaults are refused later, at admission, as for any
problem.
"""
import base64
import dataclasses
import hashlib
import json
import math
from collections.abc import Mapping

import numpy as np

from .errors import UnsupportedFeatureError, ValidationError
from .materials import Material
from .units import Units
from . import rcwa
from .rcwa._lifecycle import Execution

SCHEMA = 'wavebit.job/v2'

# public path -> class (reference jobs.py:22-47, the subset this engine implements)
_REGISTRY = {
    'wavebit.units.Units': Units,
    'wavebit.materials.Material': Material,
    'wavebit.rcwa.Problem': rcwa.Problem,
    'wavebit.rcwa.Stack': rcwa.Stack,
    'wavebit.rcwa.Layer': rcwa.Layer,
    'wavebit.rcwa.Region': rcwa.Region,
    'wavebit.rcwa.Excitation': rcwa.Excitation,
    'wavebit.rcwa.Outputs': rcwa.Outputs,
    'wavebit.rcwa.Numerics': rcwa.Numerics,
    'wavebit.rcwa.PlaneWave': rcwa.PlaneWave,
    'wavebit.Execution': Execution,
}
_TAG = {cls: tag for tag, cls in _REGISTRY.items()}

# declarations the reference serializes that this engine does not implement
_KNOWN_ELSEWHERE = (
    'wavebit.lattice.Lattice', 'wavebit.materials.Bruggeman', 'wavebit.materials.ConductiveMaterial',
    'wavebit.materials.DispersiveMaterial', 'wavebit.materials.Drude', 'wavebit.materials.Lorentz',
    'wavebit.materials.Debye', 'wavebit.rcwa.ProfiledLayer', 'wavebit.rcwa.Rectangle',
    'wavebit.rcwa.Circle', 'wavebit.rcwa.Ellipse', 'wavebit.rcwa.Polygon', 'wavebit.rcwa.Trapezoid',
    'wavebit.rcwa.Loft', 'wavebit.rcwa.Profile', 'wavebit.rcwa.Graded', 'wavebit.rcwa.Slicing',
    'wavebit.rcwa.IncidentOrders', 'wavebit.rcwa.GaussianBeam', 'wavebit.rcwa.ElectricDipole')


def _error(message, path):
    return ValidationError(message, code='job.encoding', path=path)


def _encode(value, path):
    if value is None or type(value) in (bool, str, int):
        return value
    if type(value) is float:
        if not math.isfinite(value):
            raise _error('Portable scalar values must be finite.', path)
        return value
    if isinstance(value, complex):
        return {'type': 'complex', 'real': _encode(float(value.real), path),
                'imag': _encode(float(value.imag), path)}
    if isinstance(value, np.generic):
        return _encode(np.asarray(value), path)
    if isinstance(value, np.ndarray):
        if value.dtype.hasobject or value.dtype.kind not in 'biufc':
            raise _error('Only numeric or boolean arrays are portable job data.', path)
        a = np.ascontiguousarray(value)
        return {'type': 'array', 'dtype': value.dtype.str, 'shape': list(value.shape),
                'data': base64.b64encode(a.tobytes()).decode('ascii')}
    if dataclasses.is_dataclass(value) and not isinstance(value, type):
        tag = _TAG.get(type(value))
        if tag is None:
            raise _error(f'{type(value).__name__} is not a portable public job declaration.', path)
        later = getattr(type(value), '_FIELDS_OMITTED_AT_DEFAULT', ())
        return {'type': tag, 'fields': {f.name: _encode(getattr(value, f.name), f'{path}.{f.name}')
                                      for f in dataclasses.fields(value)
                                      if f.init and not f.name.startswith('_')
                                      and not (f.name in later and getattr(value, f.name) == f.default)}}
    if isinstance(value, tuple):
        return {'type': 'tuple', 'items': [_encode(v, f'{path}[{i}]') for i, v in enumerate(value)]}
    if isinstance(value, list):
        return {'type': 'list', 'items': [_encode(v, f'{path}[{i}]') for i, v in enumerate(value)]}
    if isinstance(value, Mapping):
        pairs = [(_encode(k, path), _encode(v, f'{path}[{k!r}]')) for k, v in value.items()]
        pairs.sort(key=lambda pair: json.dumps(pair[0], sort_keys=True, allow_nan=False))
        return {'type': 'mapping', 'items': pairs}
    if callable(value):
        raise _error('Callables cannot be saved as jobs; first construct a concrete Problem.', path)
    raise _error(f'{type(value).__name__} is not portable job data. Convert concrete device arrays '
                 'to explicit host arrays before saving.', path)


def _decode(value, path):
    if value is None or type(value) in (bool, str, int):
        return value
    if type(value) is float:
        if not math.isfinite(value):
            raise _error('Nonfinite scalar in saved job.', path)
        return value
    if not isinstance(value, dict) or not isinstance(value.get('type'), str):
        raise _error('Expected a tagged job value.', path)
    tag = value['type']
    if tag == 'complex' and set(value) == {'type', 'real', 'imag'}:
        re_, im_ = value['real'], value['imag']
        if (type(re_) not in (int, float) or type(im_) not in (int, float)
                or not math.isfinite(re_) or not math.isfinite(im_)):
            raise _error('Complex parts must be finite real numbers.', path)
        return complex(re_, im_)
    if tag == 'array' and set(value) == {'type', 'dtype', 'shape', 'data'}:
        try:
            dtype = np.dtype(value['dtype'])
            shape = value['shape']
            if not isinstance(shape, list) or any(type(n) is not int or n < 0 for n in shape):
                raise ValueError('invalid shape')
            if dtype.hasobject or dtype.kind not in 'biufc':
                raise ValueError('non-numeric dtype')
            raw = base64.b64decode(value['data'], validate=True)
            if math.prod(shape) * dtype.itemsize != len(raw):
                raise ValueError('shape and byte count disagree')
            return np.frombuffer(raw, dtype=dtype).reshape(shape).copy()
        except (TypeError, ValueError, OverflowError) as exc:
            raise _error(f'Invalid array encoding: {exc}', path) from exc
    if tag in ('tuple', 'list', 'mapping') and set(value) == {'type', 'items'}:
        if not isinstance(value['items'], list):
            raise _error('items must be a list.', path)
        if tag == 'mapping':
            out = {}
            for i, item in enumerate(value['items']):
                if not isinstance(item, list) or len(item) != 2:
                    raise _error('Each mapping item must be a key/value pair.', path)
                key, entry = (_decode(v, f'{path}[{i}]') for v in item)
                try:
                    if key in out:
                        raise _error('Duplicate mapping key.', path)
                    out[key] = entry
                except TypeError as exc:
                    raise _error('Mapping keys must be immutable hashable values.', path) from exc
            return out
        items = [_decode(v, f'{path}[{i}]') for i, v in enumerate(value['items'])]
        return tuple(items) if tag == 'tuple' else items
    if tag in _REGISTRY and set(value) == {'type', 'fields'} and isinstance(value['fields'], dict):
        cls = _REGISTRY[tag]
        known = {f.name for f in dataclasses.fields(cls) if f.init and not f.name.startswith('_')}
        extra = set(value['fields']) - known
        if extra:
            raise _error(f'Unknown declaration fields: {sorted(extra)}.', path)
        kwargs = {k: _decode(v, f'{path}.{k}') for k, v in value['fields'].items()}
        try:
            return cls(**kwargs)
        except (TypeError, ValueError, NotImplementedError) as exc:
            raise _error(f'Invalid {tag}: {exc}', path) from exc
    if tag in _KNOWN_ELSEWHERE or tag.startswith('wavebit.fdtd'):
        raise UnsupportedFeatureError(f'{tag} is a declaration of the reference this engine does not '
                                      'implement (it runs raster RCWA layers lit by one plane wave).',
                                      code='rcwa.unsupported', path=path)
    raise _error(f'Unknown or malformed job type {tag!r}.', path)


@dataclasses.dataclass(frozen=True)
class Job:
    """A solver-tagged problem, numerical policy and execution settings (reference
    ``jobs.py:205-250``); only ``solver='rcwa'`` runs here."""
    solver: str
    problem: object
    numerics: object
    execution: Execution = dataclasses.field(default_factory=Execution)

    def __post_init__(self):
        if self.solver in ('fdtd', 'fdtd.native', 'fdtd.periodic'):
            raise UnsupportedFeatureError(f"solver={self.solver!r}: this engine runs 'rcwa' jobs only.",
                                          code='rcwa.unsupported', path='solver')
        if self.solver != 'rcwa':
            raise _error("solver must be 'rcwa', 'fdtd', 'fdtd.native', or 'fdtd.periodic'.", 'solver')
        if not isinstance(self.problem, rcwa.Problem) or not isinstance(self.numerics, rcwa.Numerics):
            raise _error('Problem and Numerics must belong to the selected solver.', 'problem')
        if not isinstance(self.execution, Execution):
            raise _error('execution must be Execution.', 'execution')

    def replace(self, **changes):
        return dataclasses.replace(self, **changes)

    def to_json(self, *, indent=None):
        body = {'schema': SCHEMA, 'solver': self.solver,
                'problem': _encode(self.problem, 'problem'),
                'numerics': _encode(self.numerics, 'numerics'),
                'execution': _encode(self.execution, 'execution')}
        return json.dumps(body, sort_keys=True, indent=indent, allow_nan=False,
                          separators=(',', ':') if indent is None else None)

    @classmethod
    def from_json(cls, text):
        try:
            body = json.loads(text, parse_constant=lambda name: (_ for _ in ()).throw(ValueError(name)))
        except (TypeError, ValueError) as exc:
            raise _error(f'Invalid JSON: {exc}', 'job') from exc
        if isinstance(body, dict) and body.get('schema') == 'wavebit.job/v1':
            raise UnsupportedFeatureError('This is a wavebit.job/v1 file; v1 jobs are refused, not '
                                          'converted: build the Job again and save it.',
                                          code='job.schema', path='schema')
        if not isinstance(body, dict) or body.get('schema') != SCHEMA:
            raise UnsupportedFeatureError('Expected wavebit.job/v2.', code='job.schema', path='schema')
        if set(body) != {'schema', 'solver', 'problem', 'numerics', 'execution'}:
            raise _error('The job envelope contains missing or unknown fields.', 'job')
        return cls(solver=body['solver'], problem=_decode(body['problem'], 'problem'),
                   numerics=_decode(body['numerics'], 'numerics'),
                   execution=_decode(body['execution'], 'execution'))

    @property
    def key(self):
        """SHA-256 of the canonical envelope."""
        return hashlib.sha256(self.to_json().encode('utf-8')).hexdigest()


def run(job):
    """Execute one Job (reference ``jobs.py:256-264``)."""
    if not isinstance(job, Job):
        raise _error('run accepts a Job; use rcwa.run for a Problem.', 'job')
    return rcwa.run(job.problem, numerics=job.numerics, backend=job.execution.backend,
                    device=job.execution.device, precision=job.execution.precision)
