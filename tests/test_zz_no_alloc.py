"""The allocation-free batch solve (wb_solve_noalloc) against wb_solve on the same inputs.

Two calls in a row on the library's one device array, the second smaller than the first (fewer
points, fewer layers, a smaller order window): each must equal wb_solve bit for bit, since
both run the same launch sequence and only the scratch differs. A problem larger than the fixed
array is refused before any launch.
"""
import numpy as np
import pytest
import torch

from wavebit_amd import _native
from wavebit_amd.rcwa import _engine as eng

pytestmark = pytest.mark.gpu
DEV = torch.device('cuda:0')


def _inputs(seed, P, L, hu, hv, n_raster=12, path=eng.PATH_2D):
    rng = np.random.default_rng(seed)
    ny = 1 if path != eng.PATH_2D else n_raster
    eps = (rng.uniform(1, 12, (L, ny, n_raster)) + 1j * rng.uniform(0, 0.5, (L, ny, n_raster)))
    epx, epy, epz, info = _native.assemble_li(torch.from_numpy(eps).to(DEV), hu, hv)
    assert int(info.abs().sum()) == 0
    pts = []
    for i in range(P):
        pol = ('s' if i % 2 else 'p') if path != eng.PATH_2D else (0.6, 0.8j if i % 2 else 0.8)
        phi = None if path != eng.PATH_2D else float(rng.uniform(0, 6.28))
        pt = eng.Point(stack=0, wavelength=float(rng.uniform(0.6, 1.0)),
                       theta=float(rng.uniform(0, 0.5)), phi=phi, polarization=pol,
                       eps_top=1.0, eps_bottom=2.25 + 0.05j, period=(0.9, 1.1),
                       thickness=tuple(rng.uniform(0.1, 0.4, L)))
        pts.append(eng.point_record(pt, path))
    thick = rng.uniform(0.1, 0.4, (P, L))
    idx = np.tile(np.arange(L, dtype=np.int32), (P, 1))
    return (torch.tensor(np.array(pts), dtype=torch.float64, device=DEV),
            torch.tensor(thick, dtype=torch.float64, device=DEV),
            torch.tensor(idx, device=DEV), epx, epy, epz)


@pytest.mark.parametrize('calls', [
    # (seed, P, L, hu, hv, path): the second call of each pair is the smaller one
    ((1, 6, 3, 3, 3, eng.PATH_2D), (2, 2, 1, 2, 1, eng.PATH_2D)),
    ((3, 5, 2, 10, 0, eng.PATH_TE), (4, 3, 1, 6, 0, eng.PATH_TM)),
])
def test_noalloc_solve_equals_solve_on_two_calls(calls):
    for seed, P, L, hu, hv, path in calls:
        mode = eng._MODE[path]
        assert _native.solve_workspace_bytes(P, L, hu, hv, mode) <= _native.SOLVE_NOALLOC_BYTES
        args = _inputs(seed, P, L, hu, hv, path=path)
        want = _native.solve(*args, hu, hv, mode)
        got = _native.solve_noalloc(*args, hu, hv, mode)
        torch.cuda.synchronize()
        # the 1D paths fill the first incidence variant only (the rest is not written)
        nv = 3 if path == eng.PATH_2D else 1
        for name, g, w in zip(('eff', 'amp', 'pe', 'flags'), got, want):
            if name != 'flags':
                g, w = g[:, :nv], w[:, :nv]
            assert torch.equal(g, w), (seed, name, (g - w).abs().max().item())
        assert int((got[3] & (_native.FLAG_NOCONVERGE | _native.FLAG_SINGULAR)).sum()) == 0
        assert float(got[0][:, :nv].sum()) > 0.0


def test_noalloc_refuses_a_problem_above_its_array():
    hu = hv = 9
    P, L = 4, 4
    need = _native.solve_workspace_bytes(P, L, hu, hv, _native.MODE_2D)
    assert need > _native.SOLVE_NOALLOC_BYTES
    args = _inputs(5, 1, 1, 1, 1)
    pts = args[0].repeat(P, 1)
    thick = torch.ones((P, L), dtype=torch.float64, device=DEV)
    idx = torch.zeros((P, L), dtype=torch.int32, device=DEV)
    conv = torch.zeros((1, 361, 361), dtype=torch.complex128, device=DEV)
    with pytest.raises(_native.NativeError) as exc:
        _native.solve_noalloc(pts, thick, idx, conv, conv, conv, hu, hv, _native.MODE_2D)
    assert exc.value.status == _native.WB_OOM
    assert 'scratch' in str(exc.value)
