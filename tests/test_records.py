"""Run archives (``wavebit.run/v2``, reference ``src/wavebit/records.py``) on the CPU: a result
built from the oracle's solve is archived and read back, and every kind of tampering, an
incomplete archive and an occupied target are refused."""
import io
import json
import zipfile

import numpy as np
import pytest

import wavebit_amd
from oracle import rcwa_oracle as orc
from wavebit_amd import rcwa, records
from wavebit_amd.errors import ValidationError
from wavebit_amd.jobs import Job
from wavebit_amd.rcwa._lifecycle import Execution
from wavebit_amd.rcwa.results import pack


def _job_and_result():
    eps = np.array([[1.0, 12.25, 12.25], [4.0, 1.0, 2.0 + 0.3j]])
    wls = [0.8, 0.93]
    problem = rcwa.Problem(
        stack=rcwa.Stack(layers=[rcwa.Layer(thickness=0.25, epsilon=eps)],
                         top=wavebit_amd.Material(1.0), bottom=wavebit_amd.Material(2.25),
                         period=(0.7, 0.8)),
        excitation=rcwa.Excitation(wavelength=wls,
                                   sources=(rcwa.PlaneWave(polarization='p', theta=0.2, phi=0.4),)),
        units=wavebit_amd.Units(length='um'))
    job = Job(solver='rcwa', problem=problem, numerics=rcwa.Numerics(orders=(2, 1)),
              execution=Execution(backend='torch', device='cuda:0'))
    per = [orc.solve_raster(eps[None], [0.25], 1.0, 2.25, (0.7, 0.8), wl, 0.2, 0.4, 'p', (2, 1))
           for wl in wls]
    result = pack(per, 2, 1, wls, {'finite_inputs': True},
                  {'solve_path': ('2d', '2d'), 'note': float('nan'), (1, 'x'): 3 + 4j},
                  {'solver': 'rcwa', 'engine': 'oracle-built for the archive test', 'length_units': 'um'})
    return job, result, per


def test_archive_round_trip(tmp_path):
    job, result, per = _job_and_result()
    rec = records.save_run(tmp_path / 'run', job=job, result=result, samples={'probe': np.arange(3.0)})
    assert rec.job.key == job.key
    eff = rec.outputs['transmitted']['efficiency']
    assert isinstance(eff, np.ndarray) and not eff.flags.writeable
    np.testing.assert_array_equal(eff, result.T)
    np.testing.assert_array_equal(rec.outputs['reflected']['amplitude_ratio_p'],
                                  result.reflection.amplitude_ratio_p)
    np.testing.assert_array_equal(rec.extra['wavelengths'], [0.8, 0.93])
    np.testing.assert_array_equal(rec.samples['probe'], np.arange(3.0))
    assert isinstance(rec.validity, records.RecordedObject)
    assert rec.validity.type_name == 'wavebit_amd.rcwa.results.Validity'
    assert rec.validity['checks']['numerical_finite'] is True
    assert np.isnan(rec.diagnostics['note']) and rec.diagnostics[(1, 'x')] == 3 + 4j
    assert rec.provenance['solver'] == 'rcwa'
    m = rec.manifest
    assert m['schema'] == 'wavebit.run/v2' and m['job_key'] == job.key and m['solver'] == 'rcwa'
    assert set(m['files']) == {'job.json', 'observations.json', 'arrays.npz', 'source.zip'}
    assert all(name.startswith('wavebit_amd/') for name in m['source']['files'])
    assert 'wavebit_amd/csrc/wb_rcwa.hip' in m['source']['files']
    with zipfile.ZipFile(tmp_path / 'run' / 'source.zip') as zf:
        assert set(zf.namelist()) == set(m['source']['files'])
    # the job document is the wire format the reference reads
    assert json.loads((tmp_path / 'run' / 'job.json').read_text())['schema'] == 'wavebit.job/v2'


def test_occupied_target_and_incomplete_archive_are_refused(tmp_path):
    job, result, _ = _job_and_result()
    (tmp_path / 'taken').mkdir()
    with pytest.raises(FileExistsError):
        records.save_run(tmp_path / 'taken', job=job, result=result)
    records.save_run(tmp_path / 'run', job=job, result=result, include_source=False)
    assert not (tmp_path / 'run' / 'source.zip').exists()
    (tmp_path / 'run' / 'manifest.json').unlink()
    with pytest.raises(ValidationError) as e:
        records.load_run(tmp_path / 'run')
    assert e.value.code == 'record.format' and 'manifest' in str(e.value)


def _tampered(tmp_path, edit):
    job, result, _ = _job_and_result()
    path = tmp_path / 'run'
    records.save_run(path, job=job, result=result)
    edit(path)
    with pytest.raises(ValidationError) as e:
        records.load_run(path)
    assert e.value.code == 'record.format'
    return str(e.value)


def _rewrite_manifest(path, change):
    m = json.loads((path / 'manifest.json').read_text())
    change(m)
    (path / 'manifest.json').write_text(json.dumps(m))


def _resign(path, name):
    data = (path / name).read_bytes()
    import hashlib
    _rewrite_manifest(path, lambda m: m['files'].__setitem__(
        name, {'sha256': hashlib.sha256(data).hexdigest(), 'bytes': len(data)}))


def test_a_changed_byte_is_refused(tmp_path):
    def flip(path):
        raw = bytearray((path / 'arrays.npz').read_bytes())
        raw[len(raw) // 2] ^= 1
        (path / 'arrays.npz').write_bytes(bytes(raw))
    assert 'checksum' in _tampered(tmp_path, flip)


def test_an_unreferenced_array_is_refused(tmp_path):
    def add(path):
        with np.load(path / 'arrays.npz') as npz:
            arrays = {k: npz[k] for k in npz.files}
        arrays['extra'] = np.zeros(2)
        buf = io.BytesIO()
        np.savez(buf, **arrays)
        (path / 'arrays.npz').write_bytes(buf.getvalue())
        _resign(path, 'arrays.npz')
    assert 'no observation refers' in _tampered(tmp_path, add)


def test_a_foreign_file_a_wrong_job_and_a_bad_source_inventory_are_refused(tmp_path):
    assert 'unexpected' in _tampered(tmp_path / 'a', lambda p: _rewrite_manifest(
        p, lambda m: m['files'].__setitem__('notes.txt', {'sha256': '0' * 64, 'bytes': 0})))
    assert 'job' in _tampered(tmp_path / 'b', lambda p: _rewrite_manifest(
        p, lambda m: m.__setitem__('job_key', '0' * 64)))
    assert 'identity' in _tampered(tmp_path / 'c', lambda p: _rewrite_manifest(
        p, lambda m: m['source'].__setitem__('identity', '0' * 64)))

    def escape(p):
        def change(m):
            files = m['source']['files']
            files['wavebit_amd/../evil.py'] = files.pop(next(iter(files)))
        _rewrite_manifest(p, change)
    assert 'malformed' in _tampered(tmp_path / 'd', escape)


def test_pickled_arrays_and_non_finite_json_are_refused(tmp_path):
    def pickled(path):
        buf = io.BytesIO()
        np.savez(buf, a000000=np.array([object()], dtype=object))
        (path / 'arrays.npz').write_bytes(buf.getvalue())
        _resign(path, 'arrays.npz')
    assert 'pickled' in _tampered(tmp_path / 'a', pickled)

    def nan_literal(path):
        text = (path / 'observations.json').read_text()
        assert '"rcwa"' in text
        text = text.replace('"rcwa"', 'NaN', 1)
        (path / 'observations.json').write_text(text)
        _resign(path, 'observations.json')
    assert 'JSON' in _tampered(tmp_path / 'b', nan_literal)


def test_a_reference_style_source_inventory_is_read(tmp_path):
    """Archives of the reference list their source as wavebit/...; such an inventory (with its
    identity recomputed by the shared rule) reads."""
    def rename(path):
        def change(m):
            files = {'wavebit/' + k.split('/', 1)[1]: v for k, v in m['source']['files'].items()}
            m['source']['files'] = files
            m['source']['identity'] = records._inventory_identity(files)
            m['source']['snapshot_included'] = False
            m['files'].pop('source.zip')
        _rewrite_manifest(path, change)
        (path / 'source.zip').unlink()
    job, result, _ = _job_and_result()
    path = tmp_path / 'run'
    records.save_run(path, job=job, result=result)
    rename(path)
    rec = records.load_run(path)
    assert rec.job.key == job.key
