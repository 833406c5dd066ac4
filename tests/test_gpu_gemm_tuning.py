"""GEMM tile configurations agree bit for bit; the autotuner tunes recorded shapes without
changing any result."""
import numpy as np
import pytest
import torch

from wavebit_amd import _native

pytestmark = pytest.mark.gpu
DEV = torch.device('cuda:0')


def crand(g, *shape):
    return torch.complex(torch.randn(shape, dtype=torch.float64, device=DEV, generator=g),
                         torch.randn(shape, dtype=torch.float64, device=DEV, generator=g))


SHAPES = [(3, 130, 70, 45, 0, 0), (2, 722, 2, 722, 0, 0), (2, 64, 300, 64, 1, 0), (2, 97, 33, 80, 0, 1),
          (1, 40, 130, 17, 1, 1)]


@pytest.mark.parametrize('batch,m,n,k,opa,opb', SHAPES)
def test_every_tile_configuration_gives_identical_bits(batch, m, n, k, opa, opb):
    g = torch.Generator(device=DEV)
    g.manual_seed(m * 31 + n)
    A = crand(g, batch, k, m) if opa else crand(g, batch, m, k)
    B = crand(g, batch, n, k) if opb else crand(g, batch, k, n)
    C0 = crand(g, batch, m, n)
    outs = []
    try:
        for cfg in range(len(_native.gemm_tuning_configs())):
            _native.gemm_tuning_set(opa, opb, m, n, k, batch, cfg)
            outs.append(_native.zgemm_ops(A, B, opa, opb, C=C0.clone(), alpha=0.5 - 0.25j,
                                          beta=-1.0 + 0.5j).cpu().numpy())
    finally:
        _native.gemm_tuning_clear()
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    An, Bn = A.cpu().numpy(), B.cpu().numpy()
    opA = np.conj(np.swapaxes(An, 1, 2)) if opa else An
    opB = np.conj(np.swapaxes(Bn, 1, 2)) if opb else Bn
    want = (0.5 - 0.25j) * (opA @ opB) + (-1.0 + 0.5j) * C0.cpu().numpy()
    assert np.abs(outs[0] - want).max() <= 1e-13 * np.abs(want).max() * max(1, k) ** 0.5


def test_autotune_tunes_recorded_shapes_and_keeps_results():
    g = torch.Generator(device=DEV)
    g.manual_seed(5)
    A, B = crand(g, 4, 200, 150), crand(g, 4, 150, 3)
    A2, B2 = crand(g, 2, 96, 96), crand(g, 2, 96, 96)
    _native.gemm_tuning_clear()
    try:
        ref1 = _native.zgemm_ops(A, B).cpu().numpy()
        ref2 = _native.zgemm_ops(A2, B2, 1, 0).cpu().numpy()
        _native.gemm_tuning_record(True)
        _native.zgemm_ops(A, B)
        _native.zgemm_ops(A2, B2, 1, 0)
        _native.gemm_tuning_record(False)
        assert _native.gemm_autotune(DEV, reps=2) == 2
        lines = sorted(_native.gemm_tuning_table().splitlines())
        assert [ln.rsplit(' ', 1)[0] for ln in lines] == ['0 0 200 3 150 4', '1 0 96 96 96 2']
        assert all(0 <= int(ln.rsplit(' ', 1)[1]) < 4 for ln in lines)
        assert _native.gemm_autotune(DEV, reps=2) == 0        # nothing left to tune
        assert np.array_equal(_native.zgemm_ops(A, B).cpu().numpy(), ref1)
        assert np.array_equal(_native.zgemm_ops(A2, B2, 1, 0).cpu().numpy(), ref2)
    finally:
        _native.gemm_tuning_record(False)
        _native.gemm_tuning_clear()


def test_problem_autotune_keeps_the_results():
    import wavebit_amd
    from wavebit_amd import rcwa
    rng = np.random.default_rng(2)
    rasters = rng.uniform(1.0, 9.0, size=(2, 24, 24)) + 0j
    layers = [rcwa.Layer(thickness=t, epsilon=e) for t, e in zip((120.0, 80.0), rasters)]
    prob = rcwa.Problem(
        stack=rcwa.Stack(layers=layers, top=wavebit_amd.Material(1.0), bottom=wavebit_amd.Material(2.25),
                         period=(600.0, 600.0)),
        excitation=rcwa.Excitation(wavelength=[550.0, 650.0],
                                   sources=(rcwa.PlaneWave(polarization='s', theta=0.2, phi=0.1),)),
        units=wavebit_amd.Units(length='nm'))
    num = rcwa.Numerics(orders=(4, 4))
    _native.gemm_tuning_clear()
    try:
        before = rcwa.run(prob, numerics=num, backend='torch', device='cuda:0')
        n = rcwa.autotune(prob, numerics=num, reps=1)
        assert n > 0 and len(_native.gemm_tuning_table().splitlines()) == n
        after = rcwa.run(prob, numerics=num, backend='torch', device='cuda:0')
        assert np.array_equal(before.R, after.R) and np.array_equal(before.T, after.T)
    finally:
        _native.gemm_tuning_clear()
