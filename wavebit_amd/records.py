"""Run archives in the reference's ``wavebit.run/v2`` layout (reference ``src/wavebit/records.py``,
``_common/record_io.py``, ``_common/provenance.py``): a directory holding what a run was asked
(``job.json``, a ``wavebit.job/v2`` document), what it measured (``observations.json`` plus the
numeric arrays in ``arrays.npz``), optionally the engine's source (``source.zip``), and last of all
``manifest.json``, whose presence marks the archive complete and which lists every file with its
SHA-256 and size.

Reading an archive executes nothing from it: arrays load without pickle, recorded objects come
back as ``RecordedObject`` (a type name and its fields) rather than as live classes, and every file
is checked against the manifest before use. Archives of the reference load too (their source
inventory lists ``wavebit/...`` files; this engine's lists ``wavebit_amd/...``).

Observation tree (``observations.json``), a data-only encoding whose nodes carry a ``kind``:
plain JSON scalars; ``float`` (``value`` "nan" / "inf" / "-inf") for non-finite floats;
``complex``; ``dtype``; ``array`` (``name`` of its ``arrays.npz`` member, ``dtype``, ``shape``);
``record`` (``type``: the producing class's ``module.qualname``, ``fields``); ``mapping``
(``items``: ``[key, value]`` pairs); ``tuple`` / ``list`` (``items``).
"""
import hashlib
import io
import json
import math
import platform
import subprocess
import sys
import time
import zipfile
from collections.abc import Mapping
from dataclasses import dataclass, fields, is_dataclass
from datetime import datetime, timezone
from enum import Enum
from importlib import metadata
from pathlib import Path, PurePosixPath
from types import MappingProxyType

import numpy as np

from .errors import ValidationError
from .jobs import Job, run

SCHEMA = 'wavebit.run/v2'
PACKAGE = Path(__file__).resolve().parent
_REQUIRED = frozenset({'job.json', 'observations.json', 'arrays.npz'})
_ALLOWED = _REQUIRED | {'source.zip', 'source.patch'}
_SOURCE_ROOTS = ('wavebit/', 'wavebit_amd/')
_SOURCE_SUFFIXES = ('.py', '.pyi', '.hip', '.h')
_ENVELOPE = frozenset({'outputs', 'samples', 'validity', 'diagnostics', 'provenance', 'extra',
                       'deferred'})


def _bad(message, where='record'):
    return ValidationError(message, code='record.format', path=where)


def _digest(data):
    return hashlib.sha256(data).hexdigest()


@dataclass(frozen=True)
class RecordedObject(Mapping):
    """An archived object as data: the name of the type that produced it and its fields."""
    type_name: str
    fields: Mapping

    def __post_init__(self):
        object.__setattr__(self, 'fields', MappingProxyType(dict(self.fields)))

    def __getitem__(self, key):
        return self.fields[key]

    def __iter__(self):
        return iter(self.fields)

    def __len__(self):
        return len(self.fields)


@dataclass(frozen=True)
class RunRecord:
    """One archive's contents. ``outputs['transmitted']['efficiency']`` is a read-only NumPy
    array; re-running is explicit (``jobs.run(record.job)``)."""
    path: Path
    job: Job
    outputs: Mapping
    samples: Mapping
    validity: RecordedObject
    diagnostics: Mapping
    provenance: Mapping
    extra: Mapping
    deferred: Mapping
    manifest: Mapping


# ------------------------------------------------------------------------- observation tree

class _Writer:
    """Encodes observations; numeric arrays go to ``self.arrays`` under generated names."""

    def __init__(self):
        self.arrays = {}

    def _array(self, value, where):
        if type(value).__module__.split('.')[0] == 'torch':
            value = value.detach().resolve_conj().resolve_neg().cpu().numpy()
        try:
            a = np.array(value, copy=True, order='C')
        except Exception as exc:
            raise _bad(f'{where} is not a concrete numeric array; save outside of autodiff '
                       'tracing', where) from exc
        if a.dtype.hasobject or a.dtype.kind not in 'biufc':
            raise _bad(f'{where} has dtype {a.dtype}; archives hold numeric arrays only', where)
        name = f'a{len(self.arrays):06d}'
        self.arrays[name] = a
        return {'kind': 'array', 'name': name, 'dtype': a.dtype.str, 'shape': list(a.shape)}

    def _record(self, value, names, where):
        cls = type(value)
        return {'kind': 'record', 'type': f'{cls.__module__}.{cls.__qualname__}',
                'fields': {n: self.write(getattr(value, n), f'{where}.{n}') for n in names}}

    def write(self, value, where='observations'):
        kind = type(value)
        if value is None or kind in (str, bool, int):
            return value
        if kind is float:
            return value if math.isfinite(value) else {'kind': 'float', 'value': repr(value)}
        if kind is complex:
            return {'kind': 'complex', 'real': self.write(value.real, where),
                    'imag': self.write(value.imag, where)}
        if isinstance(value, Enum):
            return {'kind': 'record', 'type': f'{kind.__module__}.{kind.__qualname__}',
                    'fields': {'value': self.write(value.value, f'{where}.value')}}
        if isinstance(value, np.dtype):
            if value.hasobject or value.kind not in 'biufc':
                raise _bad(f'{where}: only numeric dtypes are archived', where)
            return {'kind': 'dtype', 'value': value.str}
        if isinstance(value, (np.ndarray, np.generic)) or (hasattr(value, 'shape') and
                                                           hasattr(value, 'dtype')):
            return self._array(value, where)
        if is_dataclass(value) and not isinstance(value, type):
            names = [f.name for f in fields(value)]
            if any(n.startswith('_') for n in names):
                raise _bad(f'{where} ({kind.__name__}) holds private state; archive the public '
                           'data it produced instead', where)
            return self._record(value, names, where)
        if isinstance(value, tuple) and hasattr(value, '_fields'):
            return self._record(value, value._fields, where)
        if isinstance(value, Mapping):
            def scalar_key(k):
                return k is None or type(k) in (str, bool, int, float, complex) or (
                    type(k) is tuple and all(scalar_key(p) for p in k))
            if not all(scalar_key(k) for k in value):
                raise _bad(f'{where}: mapping keys must be Python scalars or tuples of them', where)
            return {'kind': 'mapping',
                    'items': [[self.write(k, f'{where}<key>'), self.write(v, f'{where}[{k!r}]')]
                              for k, v in value.items()]}
        if isinstance(value, (tuple, list)):
            return {'kind': 'tuple' if isinstance(value, tuple) else 'list',
                    'items': [self.write(v, f'{where}[{i}]') for i, v in enumerate(value)]}
        raise _bad(f'{where} ({kind.__name__}) is not archive data; evaluate or convert it first',
                   where)


class _Reader:
    """Decodes an observation tree against the archive's arrays, noting which were used."""

    def __init__(self, arrays):
        self.arrays = arrays
        self.used = set()

    def read(self, node, where='observations'):
        if node is None or type(node) in (str, bool, int):
            return node
        if type(node) is float:
            if math.isfinite(node):
                return node
            raise _bad(f'{where}: non-finite floats are written as kind "float"', where)
        if not isinstance(node, dict):
            raise _bad(f'{where}: unreadable observation node', where)
        kind, keys = node.get('kind'), set(node)
        if kind == 'array' and keys == {'kind', 'name', 'dtype', 'shape'}:
            return self._array(node, where)
        if kind == 'float' and keys == {'kind', 'value'} and node['value'] in ('nan', 'inf', '-inf'):
            return float(node['value'])
        if kind == 'complex' and keys == {'kind', 'real', 'imag'}:
            re_, im_ = self.read(node['real'], where), self.read(node['imag'], where)
            if type(re_) not in (int, float) or type(im_) not in (int, float):
                raise _bad(f'{where}: complex parts must be numbers', where)
            return complex(re_, im_)
        if kind == 'dtype' and keys == {'kind', 'value'}:
            try:
                dt = np.dtype(node['value'])
            except (TypeError, ValueError) as exc:
                raise _bad(f'{where}: unreadable dtype', where) from exc
            if dt.hasobject or dt.kind not in 'biufc':
                raise _bad(f'{where}: only numeric dtypes are archived', where)
            return dt
        if kind == 'record' and keys == {'kind', 'type', 'fields'}:
            if not (isinstance(node['type'], str) and isinstance(node['fields'], dict)):
                raise _bad(f'{where}: a record needs a type name and a field object', where)
            return RecordedObject(node['type'], {k: self.read(v, f'{where}.{k}')
                                                 for k, v in node['fields'].items()})
        if kind in ('tuple', 'list') and keys == {'kind', 'items'} and isinstance(node['items'], list):
            items = [self.read(v, f'{where}[{i}]') for i, v in enumerate(node['items'])]
            return tuple(items) if kind == 'tuple' else items
        if kind == 'mapping' and keys == {'kind', 'items'} and isinstance(node['items'], list):
            out = {}
            for i, pair in enumerate(node['items']):
                if not (isinstance(pair, list) and len(pair) == 2):
                    raise _bad(f'{where}: mapping item {i} is not a [key, value] pair', where)
                key = self.read(pair[0], where)
                try:
                    hash(key)
                except TypeError as exc:
                    raise _bad(f'{where}: mapping key {i} is not hashable', where) from exc
                if key in out:
                    raise _bad(f'{where}: mapping key {key!r} appears twice', where)
                out[key] = self.read(pair[1], f'{where}[{key!r}]')
            return MappingProxyType(out)
        raise _bad(f'{where}: unknown observation kind {kind!r} (or wrong keys for it)', where)

    def _array(self, node, where):
        name = node['name']
        if not isinstance(name, str) or name not in self.arrays:
            raise _bad(f'{where}: array {name!r} is not in arrays.npz', where)
        a = self.arrays[name]
        if a.dtype.hasobject or a.dtype.kind not in 'biufc':
            raise _bad(f'{where}: array {name} is not numeric', where)
        shape = node['shape']
        if (not isinstance(shape, list) or list(a.shape) != shape or a.dtype.str != node['dtype']):
            raise _bad(f'{where}: array {name} differs from its declared dtype or shape', where)
        self.used.add(name)
        out = a.copy()
        out.setflags(write=False)
        return out


# ------------------------------------------------------------------------- provenance

def _inventory_identity(files):
    h = hashlib.sha256()
    for name in sorted(files):
        h.update(name.encode('utf-8') + b'\0' + files[name]['sha256'].encode('ascii') + b'\n')
    return h.hexdigest()


def environment():
    """Interpreter, platform and the versions of the packages a run depends on."""
    versions = {}
    for name in ('numpy', 'scipy', 'torch'):
        try:
            versions[name] = metadata.version(name)
        except metadata.PackageNotFoundError:
            versions[name] = None
    return {'python': platform.python_version(), 'implementation': sys.implementation.name,
            'system': platform.system(), 'machine': platform.machine(), 'packages': versions,
            'engine': 'wavebit_amd'}


def _git_commit(root):
    try:
        out = subprocess.run(['git', '--no-optional-locks', '-C', str(root), 'rev-parse',
                              '--verify', 'HEAD'], capture_output=True, timeout=10)
    except (OSError, subprocess.TimeoutExpired):
        return None
    return out.stdout.decode().strip() if out.returncode == 0 else None


def capture_source(include_snapshot=True):
    """``(info, snapshot_zip_or_None)`` for this package's source: Python modules, the HIP sources
    and headers, each by SHA-256 and size; symbolic links are refused."""
    blobs = {}
    for path in sorted(PACKAGE.rglob('*')):
        if not path.is_file() or path.suffix not in _SOURCE_SUFFIXES or '__pycache__' in path.parts:
            continue
        if path.is_symlink():
            raise ValueError(f'source capture does not follow symbolic links ({path.name})')
        blobs['wavebit_amd/' + path.relative_to(PACKAGE).as_posix()] = path.read_bytes()
    files = {name: {'sha256': _digest(data), 'bytes': len(data)} for name, data in blobs.items()}
    snapshot = None
    if include_snapshot:
        buf = io.BytesIO()
        with zipfile.ZipFile(buf, 'w', compression=zipfile.ZIP_DEFLATED) as zf:
            for name, data in blobs.items():
                entry = zipfile.ZipInfo(name, date_time=(1980, 1, 1, 0, 0, 0))
                entry.compress_type = zipfile.ZIP_DEFLATED
                entry.external_attr = 0o100644 << 16
                zf.writestr(entry, data)
        snapshot = buf.getvalue()
    info = {'identity': _inventory_identity(files), 'files': files,
            'git': {'commit': _git_commit(PACKAGE)},
            'scope': 'wavebit_amd Python modules, HIP sources and headers',
            'snapshot_included': include_snapshot,
            'observation': 'files on disk at capture time'}
    return info, snapshot


# ------------------------------------------------------------------------- write

def _observations(job, result, samples):
    from .rcwa.results import Result
    if not isinstance(result, Result) or result.provenance.get('solver') != job.solver:
        raise _bad("the result must be an rcwa Result of the job's solver", 'result')
    if samples is not None and not isinstance(samples, Mapping):
        raise _bad('samples must map names to explicitly evaluated observations', 'samples')
    extra = {'wavelengths': result.wavelengths}
    if result._scattering is not None:
        extra['scattering'] = result._scattering
    return {'outputs': dict(result.outputs), 'samples': dict(samples or {}),
            'validity': result.validity, 'diagnostics': dict(result.diagnostics),
            'provenance': dict(result.provenance), 'extra': extra, 'deferred': {}}


def _json_bytes(tree):
    return (json.dumps(tree, sort_keys=True, indent=2, allow_nan=False) + '\n').encode('utf-8')


def _write(path, job, result, samples, source, env, association, execution=None):
    writer = _Writer()
    tree = writer.write(_observations(job, result, samples))
    npz = io.BytesIO()
    np.savez_compressed(npz, **writer.arrays)
    info, snapshot = source
    payloads = {'job.json': job.to_json(indent=2).encode('utf-8'),
                'observations.json': _json_bytes(tree), 'arrays.npz': npz.getvalue()}
    if snapshot is not None:
        payloads['source.zip'] = snapshot
    manifest = {'schema': SCHEMA, 'created_utc': datetime.now(timezone.utc).isoformat(),
                'job_key': job.key, 'solver': job.solver, 'environment': env, 'source': info,
                'association': association, 'execution': execution,
                'contents': 'public outputs, explicit samples and metadata as data; no live '
                            'solver or autodiff state',
                'files': {name: {'sha256': _digest(data), 'bytes': len(data)}
                          for name, data in payloads.items()}}
    path = Path(path)
    path.mkdir(parents=True, exist_ok=False)       # never into an existing file, folder or link
    for name, data in payloads.items():
        with (path / name).open('xb') as fh:
            fh.write(data)
    with (path / 'manifest.json').open('xb') as fh:  # last: the completion marker
        fh.write(_json_bytes(manifest))
    return load_run(path)


def save_run(path, *, job, result, samples=None, include_source=True):
    """Archive an existing result with the Job the caller says produced it, into a new directory
    (reference ``records.py:129-146``). Arrays are copied to the host."""
    if not isinstance(job, Job):
        raise _bad('job must be a Job', 'job')
    if type(include_source) is not bool:
        raise _bad('include_source must be True or False', 'include_source')
    job.to_json()                                    # portable inputs before any file is written
    return _write(path, job, result, samples, capture_source(include_source), environment(),
                  'caller-supplied job/result; source and environment observed at save time')


def run_recorded(job, path, *, sample=None, include_source=True):
    """Run a Job and archive its observations with the run's own provenance (reference
    ``records.py:149-180``). ``sample(result)`` may return extra evaluated observations; it runs
    here only, never when the archive is read."""
    if not isinstance(job, Job):
        raise _bad('job must be a Job', 'job')
    if sample is not None and not callable(sample):
        raise _bad('sample must be callable', 'sample')
    if type(include_source) is not bool:
        raise _bad('include_source must be True or False', 'include_source')
    if Path(path).exists() or Path(path).is_symlink():
        raise FileExistsError(path)
    document = job.to_json()
    source = capture_source(include_source)
    env = environment()
    started = datetime.now(timezone.utc).isoformat()
    t0 = time.monotonic()
    result = run(job)
    samples = sample(result) if sample is not None else None
    seconds = time.monotonic() - t0
    if (job.to_json() != document or capture_source(False)[0]['identity'] != source[0]['identity']
            or environment() != env):
        raise _bad('the job, the engine source or the environment changed while the job ran; '
                   'nothing was archived', 'execution')
    return _write(path, job, result, samples, source, env, 'executed by wavebit_amd.records.run_recorded',
                  {'started_utc': started, 'seconds_including_sampling': seconds,
                   'inputs_unchanged': True, 'source_unchanged': True})


# ------------------------------------------------------------------------- read

def _strict_json(data, where):
    def no_duplicates(pairs):
        out = {}
        for k, v in pairs:
            if k in out:
                raise ValueError(f'key {k!r} appears twice')
            out[k] = v
        return out

    def no_constants(literal):
        raise ValueError(f'{literal} is not a finite JSON number')
    try:
        return json.loads(data, object_pairs_hook=no_duplicates, parse_constant=no_constants)
    except (ValueError, UnicodeError) as exc:
        raise _bad(f'{where} is not valid JSON: {exc}', where) from exc


def _check_source(source, payloads):
    if not (isinstance(source, dict) and isinstance(source.get('files'), dict) and source['files']):
        raise _bad('the manifest has no source inventory', 'manifest.source')
    for name, row in source['files'].items():
        member = PurePosixPath(name)
        ok = (isinstance(name, str) and name.startswith(_SOURCE_ROOTS) and '..' not in member.parts
              and '\\' not in name and member.as_posix() == name and isinstance(row, dict)
              and set(row) == {'sha256', 'bytes'} and type(row['bytes']) is int and row['bytes'] >= 0
              and isinstance(row['sha256'], str) and len(row['sha256']) == 64
              and all(c in '0123456789abcdef' for c in row['sha256']))
        if not ok:
            raise _bad(f'source inventory entry {name!r} is malformed', 'manifest.source.files')
    if source.get('identity') != _inventory_identity(source['files']):
        raise _bad('the source identity does not match its inventory', 'manifest.source.identity')
    if type(source.get('snapshot_included')) is not bool:
        raise _bad('snapshot_included must be a boolean', 'manifest.source.snapshot_included')
    if source['snapshot_included'] != ('source.zip' in payloads):
        raise _bad('the manifest and the files disagree on the source snapshot', 'source.zip')
    if 'source.zip' not in payloads:
        return
    try:
        with zipfile.ZipFile(io.BytesIO(payloads['source.zip'])) as zf:
            names = zf.namelist()
            if len(set(names)) != len(names) or set(names) != set(source['files']):
                raise _bad('the source snapshot and its inventory list different files', 'source.zip')
            for name in names:
                if source['files'][name] != {'sha256': _digest(zf.read(name)),
                                             'bytes': len(zf.read(name))}:
                    raise _bad(f'snapshot member {name} differs from its inventory entry', name)
    except (zipfile.BadZipFile, KeyError) as exc:
        raise _bad('the source snapshot is not a readable zip archive', 'source.zip') from exc


def load_run(path):
    """Verify an archive against its manifest and read it (reference ``records.py:183-271``);
    nothing in it is executed."""
    path = Path(path)
    mpath = path / 'manifest.json'
    if mpath.is_symlink():
        raise _bad('archive files may not be symbolic links', 'manifest.json')
    try:
        manifest = _strict_json(mpath.read_bytes(), 'manifest.json')
    except FileNotFoundError as exc:
        raise _bad('no manifest.json: the archive is incomplete or not a run archive', str(path)) from exc
    schema = manifest.get('schema') if isinstance(manifest, dict) else None
    if schema == 'wavebit.run/v1':
        raise _bad('wavebit.run/v1 archives are not read: run the job again from current '
                   'declarations', 'manifest.schema')
    if schema != SCHEMA:
        raise _bad(f'expected a {SCHEMA} archive, found schema {schema!r}', 'manifest.schema')
    listed = manifest.get('files')
    if not (isinstance(listed, dict) and _REQUIRED <= set(listed) <= _ALLOWED):
        raise _bad('the manifest lists missing or unexpected files', 'manifest.files')
    payloads = {}
    for name, row in listed.items():
        target = path / name
        if target.is_symlink():
            raise _bad('archive files may not be symbolic links', name)
        try:
            data = target.read_bytes()
        except OSError as exc:
            raise _bad(f'{name} is missing or unreadable', name) from exc
        if row != {'sha256': _digest(data), 'bytes': len(data)}:
            raise _bad(f'{name} does not match its manifest checksum and size', name)
        payloads[name] = data
    job = Job.from_json(payloads['job.json'])
    if job.key != manifest.get('job_key') or job.solver != manifest.get('solver'):
        raise _bad('job.json is not the job the manifest names', 'job.json')
    _check_source(manifest.get('source'), payloads)
    tree = _strict_json(payloads['observations.json'], 'observations.json')
    try:
        with np.load(io.BytesIO(payloads['arrays.npz']), allow_pickle=False) as npz:
            if len(set(npz.files)) != len(npz.files):
                raise _bad('arrays.npz repeats a member name', 'arrays.npz')
            arrays = {name: npz[name] for name in npz.files}
    except (TypeError, ValueError, OSError, zipfile.BadZipFile) as exc:
        raise _bad('arrays.npz is not a numeric array archive (pickled or object arrays are not '
                   'read)', 'arrays.npz') from exc
    reader = _Reader(arrays)
    body = reader.read(tree)
    if reader.used != set(arrays):
        raise _bad('arrays.npz holds arrays no observation refers to', 'arrays.npz')
    if not (isinstance(body, Mapping) and set(body) == _ENVELOPE
            and all(isinstance(body[k], Mapping) for k in _ENVELOPE)):
        raise _bad(f'observations.json must be a mapping of exactly {sorted(_ENVELOPE)}',
                   'observations.json')
    if not isinstance(body['validity'], RecordedObject):
        raise _bad('the validity declaration is missing', 'observations.json')
    return RunRecord(path=path.resolve(), job=job, manifest=MappingProxyType(manifest),
                     **{k: body[k] for k in _ENVELOPE})
