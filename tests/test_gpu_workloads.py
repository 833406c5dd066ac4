"""The crossed-grating and uneven-batch workloads on the GPU.

* Crossed grating: 3 layers of 256 x 256 binary {1, 12.25} rasters of random rectangles and
  discs, square period 1000 nm, orders (7, 7) (450 x 450 layer eigenproblems), 512 points
  (wavelength U[400, 800] nm, theta U[0, 30] deg, phi U[0, 360) deg, s/p) in one batched
  launch sequence. Full size: every point conserves energy (|R + T - 1| <= 1e-10, the stack is
  lossless); a sample of points matches the oracle per order within 1e-10.
* Uneven batch: the first problems of the seeded 2048-problem set (orders per axis 3..12,
  1-8 layers of 128 x 128 two-material lossy rasters, substrate U[1.5, 4]) served at once,
  size-bucketed by (orders, layer count): every result is passive (R, T >= 0, R + T <= 1) and
  equals a solve of the same problem alone; the smaller problems match the oracle.
"""
import numpy as np
import pytest
import torch

from oracle import rcwa_oracle as orc
from wavebit_amd import _native, rcwa, workloads
from wavebit_amd.rcwa import _engine as eng
from wavebit_amd.rcwa import serve

pytestmark = pytest.mark.gpu
DEV = torch.device('cuda:0')


def test_crossed_grating_512_points_conserve_energy_and_match_the_oracle():
    w = workloads.crossed_grating_2d()
    hu, hv = w['orders']
    L = w['rasters'].shape[0]
    b = eng.Bucket(hu=hu, hv=hv, path=eng.PATH_2D, n_layers=L)
    for p in w['points']:
        b.points.append(eng.Point(stack=0, wavelength=p['wavelength'], theta=p['theta'], phi=p['phi'],
                                  polarization=p['polarization'], eps_top=w['eps_top'],
                                  eps_bottom=w['eps_bottom'], period=w['period'],
                                  thickness=tuple(w['thickness'])))
    conv = [_native.assemble_li(torch.from_numpy(w['rasters']).to(DEV), hu, hv)[:3]]
    eff, amp, pe, flags, N = eng.run_bucket(b, conv, DEV)
    assert int((flags & 3).sum()) == 0          # no failed eigensolve or factorization
    eff_h, amp_h, pe_h = eff.cpu().numpy(), amp.cpu().numpy(), pe.cpu().numpy()
    worst = 0.0
    for i in range(len(w['points'])):
        got = eng.unpack(eff_h, amp_h, pe_h, i, hu, hv, eng.PATH_2D)['incident']
        worst = max(worst, abs(got['R'].sum() + got['T'].sum() - 1.0))
    assert worst <= 1e-10, worst
    for i in (0, 1, 257, 511):
        p = w['points'][i]
        got = eng.unpack(eff_h, amp_h, pe_h, i, hu, hv, eng.PATH_2D)
        want = orc.solve_raster(w['rasters'], w['thickness'], w['eps_top'], w['eps_bottom'], w['period'],
                                p['wavelength'], p['theta'], p['phi'], p['polarization'], (hu, hv))
        for v in got:
            for k in ('R', 'T'):
                assert np.abs(got[v][k] - want[v][k]).max() < 1e-10, (i, v, k)


def test_uneven_batch_is_passive_bucketed_and_matches_solves_alone():
    specs = workloads.uneven_problems(count=2048)[:40]
    reqs = []
    for i, s in enumerate(specs):
        problem, numerics = workloads.to_problem(s)
        reqs.append(serve.Request(problem=problem, numerics=numerics, arrival=0.0, id=i))
    srv = serve.Server(streams=3, max_points_per_launch=64)
    results, report = srv.serve(reqs, realtime=False)
    keys = {(s['orders'], s['rasters'].shape[0]) for s in specs}
    assert report['launches'] <= len(keys)            # one launch per (orders, layers) bucket
    for i, s in enumerate(specs):
        res = results[i]
        assert not isinstance(res, Exception), (i, res)
        R, T = np.asarray(res.R), np.asarray(res.T)
        assert R.min() >= -1e-12 and T.min() >= -1e-12
        assert float(R.sum() + T.sum()) <= 1.0 + 1e-10
        if i % 5 == 0:
            problem, numerics = workloads.to_problem(s)
            alone = rcwa.run(problem, numerics=numerics, backend='torch', device='cuda:0')
            assert np.abs(R - alone.R).max() < 1e-10
            assert np.abs(T - alone.T).max() < 1e-10
        if s['orders'][0] <= 4:
            p = s['points'][0]
            want = orc.solve_raster(s['rasters'], s['thickness'], s['eps_top'], s['eps_bottom'],
                                    s['period'], p['wavelength'], p['theta'], p['phi'],
                                    p['polarization'], s['orders'])['incident']
            assert np.abs(R[0] - want['R']).max() < 1e-10, i
            assert np.abs(T[0] - want['T']).max() < 1e-10, i
