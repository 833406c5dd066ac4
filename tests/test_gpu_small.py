"""The fused solve of 1D points with at most 32 harmonics (one workgroup per point) against the
CPU oracle and against the batched kernels it replaces (WAVEBIT_AMD_FUSED_SMALL=0)."""
import os

import numpy as np
import pytest
import torch

from oracle import rcwa_oracle as orc
from wavebit_amd import _native
from wavebit_amd.rcwa import _engine as eng

pytestmark = pytest.mark.gpu
DEV = torch.device('cuda:0')


def _bucket(eps, thick, et, eb, period, points, pol, hu):
    path = eng.route(1, pol, None, 0)
    b = eng.Bucket(hu=hu, hv=0, path=path, n_layers=eps.shape[0])
    for wl, theta in points:
        b.points.append(eng.Point(stack=0, wavelength=wl, theta=theta, phi=None, polarization=pol,
                                  eps_top=et, eps_bottom=eb, period=period, thickness=tuple(thick)))
    return b, path


def _run(b, conv, fused):
    old = os.environ.get('WAVEBIT_AMD_FUSED_SMALL')
    os.environ['WAVEBIT_AMD_FUSED_SMALL'] = '1' if fused else '0'
    try:
        eff, amp, pe, flags, _ = eng.run_bucket(b, conv, DEV)
        torch.cuda.synchronize()
    finally:
        if old is None:
            del os.environ['WAVEBIT_AMD_FUSED_SMALL']
        else:
            os.environ['WAVEBIT_AMD_FUSED_SMALL'] = old
    return eff.cpu().numpy(), amp.cpu().numpy(), pe.cpu().numpy(), flags.cpu().numpy()


# (layers, pixels, hu, pol, lossy, eps_top, eps_bottom)
CASES = [
    (1, 4, 10, 's', False, 1.0, 2.25),        # the 21-harmonic binary grating's shape
    (1, 4, 10, 'p', False, 1.0, 2.25),
    (3, 7, 5, 's', True, 1.0, 2.25 + 0.05j),
    (2, 9, 15, 'p', True, 1.44, 4.0),         # 31 harmonics: the largest fused order
    (4, 5, 3, 'p', False, 2.25, 1.0),
    (1, 3, 0, 's', True, 1.0, 2.25),          # a single harmonic
    (2, 6, 1, 'p', True, 1.0, 1.0),
]


@pytest.mark.parametrize('case', CASES)
def test_fused_small_matches_oracle_and_batched(case):
    L, X, hu, pol, lossy, et, eb = case
    rng = np.random.default_rng(X * 13 + L * 7 + hu)
    eps = rng.uniform(1.0, 12.0, (L, 1, X))
    if lossy:
        eps = eps + 1j * rng.uniform(0.0, 0.8, (L, 1, X))
    eps = eps.astype(np.complex128)
    thick = list(rng.uniform(0.1, 0.5, L))
    period = (0.7, 0.7)
    points = [(0.9, 0.0), (0.61, 0.35), (1.3, -0.2)]
    b, path = _bucket(eps, thick, et, eb, period, points, pol, hu)
    conv = [_native.assemble_li(torch.from_numpy(eps).to(DEV), hu, 0)[:3]]
    ef, af, pf, ff = _run(b, conv, True)
    eb_, ab, pb, fb = _run(b, conv, False)
    assert int(ff.sum()) == 0 and int(fb.sum()) == 0
    # the fused and the batched kernels compute the same algebra in another order
    assert np.abs(ef - eb_).max() < 1e-11
    assert np.abs(af - ab).max() < 1e-10
    assert np.abs(pf - pb).max() < 1e-11
    for i, (wl, theta) in enumerate(points):
        got = eng.unpack(ef, af, pf, i, hu, 0, path)
        want = orc.solve_raster(eps, thick, et, eb, period, wl, theta, None, pol, (hu, 0))
        for v in got:
            for k in ('R_s', 'R_p', 'T_s', 'T_p', 'R', 'T'):
                assert np.abs(got[v][k] - want[v][k]).max() < 1e-10, (i, v, k)
            for k in ('r_s', 'r_p', 't_s', 't_p'):
                assert np.abs(got[v][k] - want[v][k]).max() < 1e-9, (i, v, k)
            assert abs(got[v]['power_entering'] - want[v]['power_entering']) < 1e-9


def test_fused_small_conserves_energy_on_lossless_stacks():
    rng = np.random.default_rng(5)
    L, X, hu = 3, 8, 12
    eps = rng.uniform(1.0, 9.0, (L, 1, X)).astype(np.complex128)
    thick = list(rng.uniform(0.2, 0.6, L))
    points = [(wl, th) for wl in np.linspace(0.5, 1.4, 9) for th in (0.0, 0.4)]
    for pol in ('s', 'p'):
        b, path = _bucket(eps, thick, 1.0, 2.25, (0.8, 0.8), points, pol, hu)
        conv = [_native.assemble_li(torch.from_numpy(eps).to(DEV), hu, 0)[:3]]
        ef, af, pf, ff = _run(b, conv, True)
        assert int(ff.sum()) == 0
        for i in range(len(points)):
            got = eng.unpack(ef, af, pf, i, hu, 0, path)
            tot = got['incident']['R'].sum() + got['incident']['T'].sum()
            assert abs(tot - 1.0) < 1e-11, (pol, i, tot)
