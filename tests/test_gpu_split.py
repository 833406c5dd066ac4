"""The split solve on the GPU: ``wb_layer_modes`` then ``wb_solve_from_modes`` equal ``wb_solve`` to
the bit (2D and both 1D paths); one modes batch over buckets of different layer counts serves each
bucket's cascade exactly as its own fused solve; modes kept from one run serve a run with new
thicknesses exactly as a fresh solve of those thicknesses (the resident eigensystems)."""
import numpy as np
import pytest
import torch

from wavebit_amd import _native
from wavebit_amd.rcwa import _engine as eng

pytestmark = pytest.mark.gpu
DEV = torch.device('cuda:0')


def _bucket(rng, L, P, hu, hv, path, n_raster=10):
    ny = n_raster if path == eng.PATH_2D else 1
    eps = rng.uniform(1, 9, (L, ny, n_raster)) + 1j * rng.uniform(0, 0.2, (L, ny, n_raster))
    conv = [_native.assemble_li(torch.from_numpy(eps).to(DEV), hu, hv)[:3]]
    b = eng.Bucket(hu=hu, hv=hv, path=path, n_layers=L)
    for _ in range(P):
        pol = (0.6, 0.8j) if path == eng.PATH_2D else ('s' if path == eng.PATH_TE else 'p')
        b.points.append(eng.Point(stack=0, wavelength=float(rng.uniform(0.6, 1.1)),
                                  theta=float(rng.uniform(0, 0.4)),
                                  phi=float(rng.uniform(0, 6)) if path == eng.PATH_2D else None,
                                  polarization=pol, eps_top=1.0, eps_bottom=2.25 + 0.1j,
                                  period=(0.9, 1.0), thickness=tuple(rng.uniform(0.05, 0.3, L))))
    return b, conv


def _equal(a, b, path):
    nv = 3 if path == eng.PATH_2D else 1        # the 1D paths fill the first variant only
    for x, y in zip(a[:3], b[:3]):
        assert torch.equal(x[:, :nv], y[:, :nv]), (x[:, :nv] - y[:, :nv]).abs().max()
    assert torch.equal(a[3], b[3])


def _close(a, b, path):
    """The ragged cascade runs each point's operations in batches of other sizes: equal within
    rounding (1e-12 of each output's scale), flags equal."""
    nv = 3 if path == eng.PATH_2D else 1
    for x, y in zip(a[:3], b[:3]):
        x, y = x[:, :nv], y[:, :nv]
        scale = float(y.abs().max())
        assert float((x - y).abs().max()) <= 1e-12 * max(scale, 1e-300), float((x - y).abs().max())
    assert torch.equal(a[3], b[3])


@pytest.mark.parametrize('ragged', [False, True])
@pytest.mark.parametrize('path,hu,hv', [(eng.PATH_2D, 3, 2), (eng.PATH_TE, 8, 0), (eng.PATH_TM, 8, 0)])
def test_window_of_mixed_layer_counts_equals_fused_solves(path, hu, hv, ragged):
    """One modes batch over buckets of three layer counts; per-bucket cascades equal each bucket's
    fused solve to the bit, the one ragged cascade within rounding. (A 1D bucket of at most 32
    harmonics solves alone in the one-workgroup kernel, the same algebra in another order:
    within rounding there too.)"""
    rng = np.random.default_rng(hu * 10 + len(path))
    buckets = [_bucket(rng, L, P, hu, hv, path) for L, P in ((1, 3), (3, 2), (2, 4))]
    out, modes = eng.run_window([b for b, _ in buckets], [c for _, c in buckets], DEV,
                                ragged=ragged)
    small = path != eng.PATH_2D and 2 * hu + 1 <= 32
    for (b, conv), got in zip(buckets, out):
        want = eng.run_bucket(b, conv, DEV)
        (_close if ragged or small else _equal)(got[:4], want[:4], path)
    W, V, q, flags, tables = modes
    assert W.shape[0] == sum(b.n_layers * len(b.points) for b, _ in buckets)


def test_ragged_cascade_over_eight_layer_counts_with_shared_counts():
    """Layer counts 1..8 (two buckets share a count, buckets given in no order): one ragged cascade
    equals the per-bucket cascades of the same modes within rounding, and every point's R + T
    is the same."""
    rng = np.random.default_rng(21)
    spec = [(3, 2), (8, 1), (1, 4), (5, 3), (3, 3), (2, 2), (7, 1), (4, 2), (6, 1)]
    buckets = [_bucket(rng, L, P, 3, 3, eng.PATH_2D, n_raster=12) for L, P in spec]
    bs, cs = [b for b, _ in buckets], [c for _, c in buckets]
    got, modes = eng.run_window(bs, cs, DEV, ragged=True)
    want, _ = eng.run_window(bs, cs, DEV, modes=modes, ragged=False)
    for g, w in zip(got, want):
        _close(g[:4], w[:4], eng.PATH_2D)


def test_ragged_entry_refuses_unordered_layer_counts():
    rng = np.random.default_rng(3)
    b, conv = _bucket(rng, 2, 2, 2, 2, eng.PATH_2D)
    _, modes = eng.run_window([b], [conv], DEV)
    W, V, q = modes[:3]
    pts = torch.zeros((2, 16), dtype=torch.float64, device=DEV)
    thick = torch.full((3,), 0.1, dtype=torch.float64, device=DEV)
    with pytest.raises(Exception) as e:
        _native.solve_from_modes_ragged(pts, thick, W, V, q, None, [1, 2], 2, 2, eng._MODE[eng.PATH_2D])
    assert 'non-increasing' in str(e.value)


def test_resident_modes_serve_new_thicknesses_exactly():
    rng = np.random.default_rng(5)
    b, conv = _bucket(rng, 3, 4, 4, 4, eng.PATH_2D)
    _, modes = eng.run_window([b], [conv], DEV)
    import dataclasses
    b2 = dataclasses.replace(b, points=[dataclasses.replace(p, thickness=tuple(np.asarray(p.thickness) * 1.7))
                                        for p in b.points])
    got, _ = eng.run_window([b2], [conv], DEV, modes=modes)
    want = eng.run_bucket(b2, conv, DEV)
    _equal(got[0][:4], want[:4], eng.PATH_2D)


def test_prepared_keeps_layer_eigensystems_for_a_thickness_sweep():
    """A preparation above the captured-graph size keeps its layer modes on the device: a run that
    changes only the thicknesses reuses them (no eigensolve) and equals a fresh preparation's run
    of the same thicknesses to the bit; new wavelengths recompute them."""
    import wavebit_amd
    from wavebit_amd import rcwa
    rng = np.random.default_rng(9)
    eps = rng.uniform(1, 9, (3, 16, 16)) + 1j * rng.uniform(0, 0.1, (3, 16, 16))

    def problem(thick, wls):
        return rcwa.Problem(
            stack=rcwa.Stack(layers=[rcwa.Layer(thickness=t, epsilon=e) for t, e in zip(thick, eps)],
                             top=wavebit_amd.Material(1.0), bottom=wavebit_amd.Material(2.25),
                             period=(0.8, 0.9)),
            excitation=rcwa.Excitation(wavelength=wls, sources=(rcwa.PlaneWave(polarization='p',
                                                                               theta=0.2, phi=0.5),)),
            units=wavebit_amd.Units(length='um'))
    wls = list(np.linspace(0.7, 1.0, 8))
    num = rcwa.Numerics(orders=(6, 6))
    prep = rcwa.prepare(problem([0.1, 0.2, 0.15], wls), numerics=num, backend='torch', device='cuda:0')
    prep.run(problem([0.1, 0.2, 0.15], wls))
    first = prep._cache.modes
    assert first is not None
    res = prep.run(problem([0.12, 0.25, 0.11], wls))
    assert prep._cache.modes[1] is first[1]           # the same device buffers: no eigensolve
    fresh = rcwa.run(problem([0.12, 0.25, 0.11], wls), numerics=num, backend='torch', device='cuda:0')
    assert np.array_equal(res.R, fresh.R) and np.array_equal(res.T, fresh.T)
    prep.run(problem([0.12, 0.25, 0.11], [w * 1.01 for w in wls]))
    assert prep._cache.modes[1] is not first[1]
