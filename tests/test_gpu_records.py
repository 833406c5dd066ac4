"""A job run on the GPU and archived by ``records.run_recorded`` (``wavebit.run/v2``): the archive
reads back with the run's efficiencies to the bit, and re-running its job gives them again."""
import numpy as np
import pytest

import wavebit_amd
from wavebit_amd import jobs, rcwa, records
from wavebit_amd.rcwa._lifecycle import Execution

pytestmark = pytest.mark.gpu


def test_run_recorded_archives_the_device_result(tmp_path):
    rng = np.random.default_rng(4)
    eps = rng.uniform(1, 9, (2, 8, 8)) + 1j * rng.uniform(0, 0.1, (2, 8, 8))
    problem = rcwa.Problem(
        stack=rcwa.Stack(layers=[rcwa.Layer(thickness=t, epsilon=e) for t, e in zip((0.2, 0.31), eps)],
                         top=wavebit_amd.Material(1.0), bottom=wavebit_amd.Material(2.25),
                         period=(0.9, 0.8)),
        excitation=rcwa.Excitation(wavelength=[0.71, 0.83],
                                   sources=(rcwa.PlaneWave(polarization='s', theta=0.3, phi=0.6),)),
        units=wavebit_amd.Units(length='um'))
    job = jobs.Job(solver='rcwa', problem=problem, numerics=rcwa.Numerics(orders=(3, 3)),
                   execution=Execution(backend='torch', device='cuda:0'))
    rec = records.run_recorded(job, tmp_path / 'run',
                               sample=lambda r: {'absorption': r.absorption})
    assert rec.manifest['execution']['inputs_unchanged'] is True
    again = jobs.run(rec.job)
    np.testing.assert_array_equal(rec.outputs['transmitted']['efficiency'], again.T)
    np.testing.assert_array_equal(rec.outputs['reflected']['efficiency'], again.R)
    np.testing.assert_array_equal(rec.samples['absorption'], again.absorption)
    assert rec.provenance['engine'] == 'wavebit_amd/hip-gfx950'
