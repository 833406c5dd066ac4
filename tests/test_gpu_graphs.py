"""Captured HIP graphs of the prepared solve: every replay equals the eager launch sequence.

A prepared problem of launch-bound size runs eagerly once, is captured the second time and
replayed afterwards. Each replay (with new live values: wavelengths, thicknesses, rasters) is
checked against an eager solve of the same live problem, with other eager device work and
synchronizations between the replays; the 1D configuration-1 grating is also checked against
the oracle through the replayed graph."""
import numpy as np
import pytest
import torch

import wavebit_amd
from wavebit_amd import rcwa, workloads
from wavebit_amd.rcwa import graphs
from oracle import rcwa_oracle as orc

pytestmark = pytest.mark.gpu
RUN = dict(backend='torch', device='cuda:0')


def _problem(rasters, thickness, wavelengths, pol='p', theta=0.3, phi=0.7, period=(600.0, 650.0)):
    layers = [rcwa.Layer(thickness=float(t), epsilon=e) for t, e in zip(thickness, rasters)]
    return rcwa.Problem(
        stack=rcwa.Stack(layers=layers, top=wavebit_amd.Material(1.0),
                         bottom=wavebit_amd.Material(2.25), period=period),
        excitation=rcwa.Excitation(wavelength=wavelengths,
                                   sources=(rcwa.PlaneWave(polarization=pol, theta=theta, phi=phi),)),
        units=wavebit_amd.Units(length='nm'))


def _eager(problem, numerics):
    # a fresh preparation runs its first solve eagerly
    return rcwa.prepare(problem, numerics=numerics, **RUN).run(problem)


def test_prepared_replays_match_eager_solves():
    rng = np.random.default_rng(11)
    numerics = rcwa.Numerics(orders=(3, 3))
    rasters = rng.uniform(1.0, 9.0, size=(2, 24, 24)) + 1j * rng.uniform(0.0, 0.05, size=(2, 24, 24))
    thick = np.array([140.0, 90.0])
    wls = [520.0, 610.0, 700.0]
    prepared = rcwa.prepare(_problem(rasters, thick, wls), numerics=numerics, **RUN)
    assert prepared._cache.graph[0] is None
    prepared.run(_problem(rasters, thick, wls))                  # eager
    assert prepared._cache.graph[0] is None
    for rep in range(5):
        # new live values every run: thicknesses, wavelengths and (from rep 2) the rasters
        t_rep = thick * (1.0 + 0.05 * rep)
        w_rep = [w + 7.0 * rep for w in wls]
        r_rep = rasters if rep < 2 else rasters * (1.0 + 0.02 * rep)
        problem = _problem(r_rep, t_rep, w_rep)
        got = prepared.run(problem)
        assert prepared._cache.graph[0] is not None               # captured on the second run
        # eager work of the program between replays, then a full synchronization
        x = torch.randn(512, 512, dtype=torch.float64, device='cuda:0')
        (x @ x).sum().item()
        torch.cuda.synchronize()
        want = _eager(problem, numerics)
        assert np.array_equal(got.R, want.R), rep
        assert np.array_equal(got.T, want.T), rep
        assert np.array_equal(got.reflection.amplitude_ratio_p, want.reflection.amplitude_ratio_p)


def test_config1_grating_through_the_replayed_graph_matches_the_oracle():
    w = workloads.binary_grating_1d()
    p = w['points'][0]
    numerics = rcwa.Numerics(orders=w['orders'])

    def problem(wl):
        layers = [rcwa.Layer(thickness=float(t), epsilon=e) for t, e in zip(w['thickness'], w['rasters'])]
        return rcwa.Problem(
            stack=rcwa.Stack(layers=layers, top=wavebit_amd.Material(w['eps_top']),
                             bottom=wavebit_amd.Material(w['eps_bottom']), period=w['period']),
            excitation=rcwa.Excitation(wavelength=wl, sources=(rcwa.PlaneWave(polarization='s',
                                                                                theta=p['theta'], phi=None),)),
            units=wavebit_amd.Units(length='nm'))

    prepared = rcwa.prepare(problem(p['wavelength']), numerics=numerics, **RUN)
    for i, wl in enumerate([p['wavelength'], 880.0, 905.0, 930.0]):
        res = prepared.run(problem(wl))
        if i >= 1:
            assert prepared._cache.graph[0] is not None
        ref = orc.solve_raster(w['rasters'], w['thickness'], w['eps_top'], w['eps_bottom'], w['period'],
                               wl, p['theta'], None, 's', w['orders'])['incident']
        assert np.abs(res.R[0] - ref['R']).max() < 1e-10
        assert np.abs(res.T[0] - ref['T']).max() < 1e-10
        torch.cuda.synchronize()


def test_large_preparations_stay_eager():
    # (9,9) with four layers: the workspace is far above the capture limit
    assert not graphs.worth_capturing(8, 4, 9, 9, 0)
    assert graphs.worth_capturing(3, 2, 3, 3, 0)
