"""Portable RCWA jobs: the ``wavebit.job/v2`` document (reference ``src/wavebit/jobs.py``), so the
server and ``run`` take the reference's own job files, and write files the reference reads.

Wire format (the interface both sides share; reference ``jobs.py:22-47`` for the tags):

* envelope: a JSON object with exactly ``schema`` (``"wavebit.job/v2"``), ``solver``,
  ``problem``, ``numerics`` and ``execution``; canonical text is key-sorted with ``,``/``:``
  separators, and a job's ``key`` is the SHA-256 of that text;
* plain values: ``null``, booleans, strings, integers and finite floats as themselves;
* tagged values, objects with a ``type``: ``complex`` (``real``, ``imag``), ``array``
  (``dtype`` as NumPy's type string, ``shape``, ``data`` the base64 of the C-ordered bytes),
  ``tuple`` / ``list`` (``items``), ``mapping`` (``items``: ``[key, value]`` pairs ordered by the
  key-sorted JSON text of the encoded key), and a declaration (its public path, e.g.
  ``wavebit.rcwa.Problem``, with ``fields``: every init field, less those a class lists in
  ``_FIELDS_OMITTED_AT_DEFAULT`` while they hold their default).

Which Python values take which form is part of the format: exact ``int``/``float``/``bool``/
``str`` are plain, any ``complex`` (NumPy's included) is ``complex``, every other NumPy scalar is
a 0-d ``array``. Reading resolves tags through the fixed table below and never imports a name
found in a document; declarations of the reference this engine does not run are refused by name.
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
_ENVELOPE = ('schema', 'solver', 'problem', 'numerics', 'execution')

# the declarations this engine reads and writes, under the reference's public paths
DECLARATIONS = {
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
_PATH_OF = {cls: path for path, cls in DECLARATIONS.items()}

# public paths of the reference's other declarations: documents naming them are understood and
# refused, not mistaken for malformed input
NOT_RUN_HERE = frozenset((
    'wavebit.lattice.Lattice', 'wavebit.materials.Bruggeman', 'wavebit.materials.ConductiveMaterial',
    'wavebit.materials.DispersiveMaterial', 'wavebit.materials.Drude', 'wavebit.materials.Lorentz',
    'wavebit.materials.Debye', 'wavebit.rcwa.ProfiledLayer', 'wavebit.rcwa.Rectangle',
    'wavebit.rcwa.Circle', 'wavebit.rcwa.Ellipse', 'wavebit.rcwa.Polygon', 'wavebit.rcwa.Trapezoid',
    'wavebit.rcwa.Loft', 'wavebit.rcwa.Profile', 'wavebit.rcwa.Graded', 'wavebit.rcwa.Slicing',
    'wavebit.rcwa.IncidentOrders', 'wavebit.rcwa.GaussianBeam', 'wavebit.rcwa.ElectricDipole'))


def _bad(where, message):
    return ValidationError(message, code='job.encoding', path=where)


def _canonical(node):
    return json.dumps(node, sort_keys=True, separators=(',', ':'), allow_nan=False)


def _field_items(decl):
    """(name, value) of the fields a declaration writes."""
    skip_at_default = getattr(type(decl), '_FIELDS_OMITTED_AT_DEFAULT', ())
    for f in dataclasses.fields(decl):
        if not f.init or f.name.startswith('_'):
            continue
        value = getattr(decl, f.name)
        if f.name in skip_at_default and value == f.default:
            continue
        yield f.name, value


def encode(value, where='job'):
    """The tagged JSON tree of ``value`` (plain Python data, ready for ``json.dumps``)."""
    kind = type(value)
    if value is None or kind in (bool, str, int):
        return value
    if kind is float:
        if not math.isfinite(value):
            raise _bad(where, f'{value!r} cannot be written: job numbers are finite')
        return value
    match value:
        case complex():
            return {'type': 'complex', 'real': encode(float(value.real), where),
                    'imag': encode(float(value.imag), where)}
        case np.generic():
            return encode(np.asarray(value), where)
        case np.ndarray():
            if value.dtype.hasobject or value.dtype.kind not in 'biufc':
                raise _bad(where, f'an array of dtype {value.dtype} has no portable form; jobs carry '
                                  'boolean and numeric arrays')
            raw = np.ascontiguousarray(value).tobytes()
            return {'type': 'array', 'dtype': value.dtype.str, 'shape': list(value.shape),
                    'data': base64.b64encode(raw).decode('ascii')}
        case tuple() | list():
            return {'type': 'tuple' if isinstance(value, tuple) else 'list',
                    'items': [encode(v, f'{where}[{i}]') for i, v in enumerate(value)]}
        case Mapping():
            pairs = [[encode(k, f'{where}<key>'), encode(v, f'{where}[{k!r}]')]
                     for k, v in value.items()]
            # ordered by the key's JSON text as json.dumps writes it by default (sorted keys)
            return {'type': 'mapping',
                    'items': sorted(pairs, key=lambda kv: json.dumps(kv[0], sort_keys=True))}
    if dataclasses.is_dataclass(value) and not isinstance(value, type):
        path = _PATH_OF.get(kind)
        if path is None:
            raise _bad(where, f'{kind.__name__} has no public job declaration path')
        return {'type': path, 'fields': {name: encode(v, f'{where}.{name}')
                                         for name, v in _field_items(value)}}
    if callable(value):
        raise _bad(where, 'a function cannot travel in a job document: evaluate it into a concrete '
                          'Problem first')
    raise _bad(where, f'{kind.__name__} is not job data (device tensors must be copied to host '
                      'arrays before the job is written)')


def _finite_number(node):
    return type(node) in (int, float) and math.isfinite(node)


def _decode_array(node, where):
    try:
        dtype = np.dtype(node['dtype'])
    except (TypeError, ValueError) as exc:
        raise _bad(where, f'unreadable array dtype {node["dtype"]!r}') from exc
    if dtype.hasobject or dtype.kind not in 'biufc':
        raise _bad(where, f'array dtype {dtype} is not boolean or numeric')
    shape = node['shape']
    if not isinstance(shape, list) or not all(type(d) is int and d >= 0 for d in shape):
        raise _bad(where, 'array shape must be a list of non-negative integers')
    if not isinstance(node['data'], str):
        raise _bad(where, 'array data must be a base64 string')
    try:
        raw = base64.b64decode(node['data'], validate=True)
    except ValueError as exc:
        raise _bad(where, 'array data is not valid base64') from exc
    if len(raw) != math.prod(shape) * dtype.itemsize:
        raise _bad(where, f'array data holds {len(raw)} bytes, shape {shape} of {dtype} needs '
                          f'{math.prod(shape) * dtype.itemsize}')
    return np.frombuffer(raw, dtype=dtype).reshape(shape).copy()


def _decode_mapping(items, where):
    out = {}
    for i, pair in enumerate(items):
        if not (isinstance(pair, list) and len(pair) == 2):
            raise _bad(where, f'mapping item {i} is not a [key, value] pair')
        key = decode(pair[0], f'{where}<key {i}>')
        try:
            hash(key)
        except TypeError as exc:
            raise _bad(where, f'mapping key {i} is not hashable') from exc
        if key in out:
            raise _bad(where, f'mapping key {key!r} appears twice')
        out[key] = decode(pair[1], f'{where}[{key!r}]')
    return out


def _decode_declaration(path, fields, where):
    cls = DECLARATIONS[path]
    accepted = {f.name for f in dataclasses.fields(cls) if f.init and not f.name.startswith('_')}
    unknown = sorted(set(fields) - accepted)
    if unknown:
        raise _bad(where, f'{path} has no field(s) {unknown}')
    args = {name: decode(node, f'{where}.{name}') for name, node in fields.items()}
    try:
        return cls(**args)
    except (TypeError, ValueError, NotImplementedError) as exc:
        raise _bad(where, f'{path} refused its fields: {exc}') from exc


def decode(node, where='job'):
    """The Python value of a tagged JSON tree (the inverse of ``encode``)."""
    if node is None or type(node) in (bool, str, int):
        return node
    if type(node) is float:
        if not math.isfinite(node):
            raise _bad(where, 'job numbers must be finite')
        return node
    if not (isinstance(node, dict) and isinstance(node.get('type'), str)):
        raise _bad(where, 'expected a plain value or an object with a "type" tag')
    tag, keys = node['type'], set(node)
    if tag == 'complex' and keys == {'type', 'real', 'imag'}:
        if not (_finite_number(node['real']) and _finite_number(node['imag'])):
            raise _bad(where, 'complex parts must be finite numbers')
        return complex(node['real'], node['imag'])
    if tag == 'array' and keys == {'type', 'dtype', 'shape', 'data'}:
        return _decode_array(node, where)
    if tag in ('tuple', 'list', 'mapping') and keys == {'type', 'items'}:
        if not isinstance(node['items'], list):
            raise _bad(where, f'a {tag} carries its items as a list')
        if tag == 'mapping':
            return _decode_mapping(node['items'], where)
        items = [decode(v, f'{where}[{i}]') for i, v in enumerate(node['items'])]
        return tuple(items) if tag == 'tuple' else items
    if tag in DECLARATIONS and keys == {'type', 'fields'} and isinstance(node['fields'], dict):
        return _decode_declaration(tag, node['fields'], where)
    if tag in NOT_RUN_HERE or tag.startswith('wavebit.fdtd'):
        raise UnsupportedFeatureError(f'{tag} is a reference declaration this engine does not run: '
                                      'it solves raster RCWA stacks lit by one plane wave',
                                      code='rcwa.unsupported', path=where)
    raise _bad(where, f'unrecognised tagged value {tag!r} (or wrong keys for that tag)')


def _strict_json(text):
    def refuse(literal):
        raise ValueError(f'{literal} is not a finite JSON number')
    return json.loads(text, parse_constant=refuse)


@dataclasses.dataclass(frozen=True)
class Job:
    """One solver-tagged calculation: a problem, its numerical policy and where it runs
    (reference ``jobs.py:205-250``). Only ``solver='rcwa'`` runs on this engine."""
    solver: str
    problem: object
    numerics: object
    execution: Execution = dataclasses.field(default_factory=Execution)

    def __post_init__(self):
        if self.solver in ('fdtd', 'fdtd.native', 'fdtd.periodic'):
            raise UnsupportedFeatureError(f'solver {self.solver!r} is not run by this engine, which '
                                          "runs 'rcwa' jobs", code='rcwa.unsupported', path='solver')
        if self.solver != 'rcwa':
            raise _bad('solver', f"unknown solver {self.solver!r}; the reference's solvers are "
                                 "'rcwa', 'fdtd', 'fdtd.native' and 'fdtd.periodic'")
        if not isinstance(self.problem, rcwa.Problem):
            raise _bad('problem', 'an rcwa job needs an rcwa.Problem')
        if not isinstance(self.numerics, rcwa.Numerics):
            raise _bad('numerics', 'an rcwa job needs rcwa.Numerics')
        if not isinstance(self.execution, Execution):
            raise _bad('execution', 'execution must be an Execution')

    def replace(self, **changes):
        return dataclasses.replace(self, **changes)

    def _tree(self):
        return {'schema': SCHEMA, 'solver': self.solver,
                'problem': encode(self.problem, 'problem'),
                'numerics': encode(self.numerics, 'numerics'),
                'execution': encode(self.execution, 'execution')}

    def to_json(self, *, indent=None):
        """The document; without ``indent`` its canonical (key-sorted, compact) text."""
        tree = self._tree()
        if indent is None:
            return _canonical(tree)
        return json.dumps(tree, sort_keys=True, indent=indent, allow_nan=False)

    @classmethod
    def from_json(cls, text):
        try:
            tree = _strict_json(text)
        except (TypeError, ValueError) as exc:
            raise _bad('job', f'not a JSON document: {exc}') from exc
        schema = tree.get('schema') if isinstance(tree, dict) else None
        if schema == 'wavebit.job/v1':
            raise UnsupportedFeatureError('wavebit.job/v1 documents are not read: rebuild the Job '
                                          'from current declarations and write it again',
                                          code='job.schema', path='schema')
        if schema != SCHEMA:
            raise UnsupportedFeatureError(f'expected a {SCHEMA} document, found schema {schema!r}',
                                          code='job.schema', path='schema')
        if set(tree) != set(_ENVELOPE):
            raise _bad('job', f'the envelope must hold exactly {list(_ENVELOPE)}, found {sorted(tree)}')
        return cls(solver=tree['solver'], problem=decode(tree['problem'], 'problem'),
                   numerics=decode(tree['numerics'], 'numerics'),
                   execution=decode(tree['execution'], 'execution'))

    @property
    def key(self):
        """SHA-256 (hex) of the canonical document: equal jobs, equal keys."""
        return hashlib.sha256(self.to_json().encode('utf-8')).hexdigest()


def run(job):
    """Run one Job on its execution (reference ``jobs.py:256-264``)."""
    if not isinstance(job, Job):
        raise _bad('job', 'run takes a Job (rcwa.run takes a Problem)')
    ex = job.execution
    return rcwa.run(job.problem, numerics=job.numerics, backend=ex.backend, device=ex.device,
                    precision=ex.precision)
