"""The ``wavebit.job/v2`` wire format on the CPU (reference ``src/wavebit/jobs.py``; the cases
mirror ``tests/api/test_unified_jobs.py``): round trips, keys, refusals, and the envelope's
exact shape for a small problem, restated from the reference's encoding rules."""
import json

import numpy as np
import pytest

import wavebit_amd
from wavebit_amd import rcwa
from wavebit_amd.errors import UnsupportedFeatureError, ValidationError
from wavebit_amd.jobs import Job, SCHEMA
from wavebit_amd.rcwa._lifecycle import Execution


def _problem(eps=None):
    eps = np.array([[1.0, 12.25], [12.25, 1.0 + 0.5j]]) if eps is None else eps
    return rcwa.Problem(
        stack=rcwa.Stack(layers=[rcwa.Layer(thickness=0.25, epsilon=eps)], top=wavebit_amd.Material(1.0),
                         bottom=wavebit_amd.Material(2.25 + 0.1j), period=(0.7, 0.8)),
        excitation=rcwa.Excitation(wavelength=[0.8, 0.9],
                                   sources=(rcwa.PlaneWave(polarization=(0.6, 0.8j), theta=0.2, phi=0.3),)),
        units=wavebit_amd.Units(length='um'))


def job(**kw):
    return Job(solver='rcwa', problem=kw.get('problem', _problem()),
               numerics=kw.get('numerics', rcwa.Numerics(orders=(3, 2))),
               execution=Execution(backend='torch', device='cuda:0'))


def test_round_trip_preserves_every_declaration_and_array():
    original = job()
    restored = Job.from_json(original.to_json())
    assert restored.to_json() == original.to_json() and restored.key == original.key
    eps = restored.problem.stack.layers[0].epsilon
    assert eps.dtype == np.complex128 and eps.shape == (2, 2)
    np.testing.assert_array_equal(eps, original.problem.stack.layers[0].epsilon)
    assert restored.problem.stack.bottom.epsilon == 2.25 + 0.1j
    assert restored.numerics == original.numerics and restored.execution == original.execution
    assert Job.from_json(original.to_json(indent=2)).key == original.key


def test_the_envelope_has_the_reference_shape():
    """Tags are the reference's public paths; a Numerics field listed as omitted-at-default is
    absent until it is set (reference numerics.py:104-105, jobs.py:103-110)."""
    body = json.loads(job().to_json())
    assert set(body) == {'schema', 'solver', 'problem', 'numerics', 'execution'}
    assert body['schema'] == SCHEMA == 'wavebit.job/v2' and body['solver'] == 'rcwa'
    assert body['execution'] == {'type': 'wavebit.Execution',
                                 'fields': {'backend': 'torch', 'device': 'cuda:0', 'precision': 'double'}}
    p = body['problem']
    assert p['type'] == 'wavebit.rcwa.Problem' and set(p['fields']) == {'stack', 'excitation', 'units', 'outputs'}
    layer = p['fields']['stack']['fields']['layers']['items'][0]
    assert layer['type'] == 'wavebit.rcwa.Layer'
    assert set(layer['fields']) == {'thickness', 'background', 'regions', 'epsilon', 'mu'}
    assert layer['fields']['epsilon']['type'] == 'array' and layer['fields']['epsilon']['dtype'] == '<c16'
    assert p['fields']['units'] == {'type': 'wavebit.units.Units', 'fields': {'length': 'um'}}
    assert p['fields']['stack']['fields']['bottom'] == {
        'type': 'wavebit.materials.Material',
        'fields': {'epsilon': {'type': 'complex', 'real': 2.25, 'imag': 0.1}, 'mu': None}}
    n = body['numerics']['fields']
    assert n['orders'] == {'type': 'tuple', 'items': [3, 2]}
    assert {'raster_normal_method', 'nvf_projector', 'shape_nvf_grid_size', 'topology_check'} <= set(n)
    assert not {'half_space_merge_tolerance', 'exceptional_points', 'wall_resolution', 'shape_normal_method',
                'slicing'} & set(n)
    off = json.loads(job(numerics=rcwa.Numerics(orders=(3, 2), exceptional_points='off')).to_json())
    assert off['numerics']['fields']['exceptional_points'] == 'off'


def test_key_changes_with_physics_and_execution():
    a, b = job(), job(numerics=rcwa.Numerics(orders=(3, 3)))
    assert a.key != b.key
    assert a.key != a.replace(execution=Execution(backend='torch', device='cuda:1')).key


@pytest.mark.parametrize('mutate,error', [
    (lambda b: b.update(schema='wavebit.job/v1'), UnsupportedFeatureError),
    (lambda b: b.update(schema='wavebit.job/v3'), UnsupportedFeatureError),
    (lambda b: b.update(extra=1), ValidationError),
    (lambda b: b['problem'].update(type='os.system'), ValidationError),
    (lambda b: b['problem']['fields'].update(bogus=1), ValidationError),
    (lambda b: b['numerics']['fields'].update(orders={'type': 'tuple', 'items': [1.5, 2]}), ValidationError),
    (lambda b: b['problem']['fields']['stack']['fields']['layers']['items'][0]['fields']['epsilon'].update(
        shape=[3, 3]), ValidationError),
    (lambda b: b['problem'].update(type='wavebit.rcwa.GaussianBeam'), UnsupportedFeatureError),
])
def test_malformed_jobs_are_refused(mutate, error):
    body = json.loads(job().to_json())
    mutate(body)
    with pytest.raises(error):
        Job.from_json(json.dumps(body))


def test_nonportable_values_are_refused():
    for bad in (float('inf'), lambda: 0.2, object()):
        p = _problem()
        object.__setattr__(p.stack.layers[0], 'thickness', bad)     # past the constructor's checks
        with pytest.raises(ValidationError):
            job(problem=p).to_json()
    with pytest.raises(ValidationError):
        Job.from_json('{"schema": "wavebit.job/v2", "solver": "rcwa", "problem": NaN}')
    with pytest.raises(UnsupportedFeatureError):
        Job(solver='fdtd', problem=_problem(), numerics=rcwa.Numerics())


def test_unsupported_numerics_decode_and_are_refused_at_admission():
    body = json.loads(job().to_json())
    body['numerics']['fields']['raster_normal_method'] = 'distance'
    j = Job.from_json(json.dumps(body))
    assert j.numerics.raster_normal_method == 'distance'
    with pytest.raises(UnsupportedFeatureError) as e:
        rcwa.prepare(j.problem, numerics=j.numerics, backend='torch', device='cuda')
    assert e.value.code == 'rcwa.unsupported'


def test_a_job_saved_for_the_cpu_becomes_a_server_request():
    """The reference's default execution is numpy on the CPU; the server takes the problem and
    numerics and runs them on its own device."""
    from wavebit_amd.rcwa import serve
    text = Job(solver='rcwa', problem=_problem(), numerics=rcwa.Numerics(orders=(2, 2))).to_json()
    assert json.loads(text)['execution']['fields']['backend'] == 'numpy'
    req = serve.Request.from_job(text, arrival=0.5, id=7)
    assert req.id == 7 and req.arrival == 0.5 and req.numerics.orders == (2, 2)
    assert isinstance(req.problem, rcwa.Problem)
