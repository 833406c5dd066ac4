"""GPU parity of the engine's batched primitives against NumPy FP64 on the same inputs."""
import numpy as np
import pytest
import torch

from wavebit_amd import _native

pytestmark = pytest.mark.gpu
DEV = torch.device('cuda:0')


def crand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize('m,n,k,batch', [(16, 16, 4, 1), (64, 64, 64, 2), (67, 45, 130, 3),
                                         (1, 300, 17, 2), (200, 1, 65, 1), (722, 722, 361, 1)])
def test_zgemm_matches_numpy(m, n, k, batch):
    rng = np.random.default_rng(m * 7 + n)
    A, B, C = crand(rng, batch, m, k), crand(rng, batch, k, n), crand(rng, batch, m, n)
    alpha, beta = 0.7 - 0.2j, -1.3 + 0.5j
    got = _native.zgemm(torch.from_numpy(A).to(DEV), torch.from_numpy(B).to(DEV),
                        torch.from_numpy(C.copy()).to(DEV), alpha, beta).cpu().numpy()
    want = alpha * (A @ B) + beta * C
    assert np.abs(got - want).max() <= 1e-13 * max(1.0, np.abs(want).max()) * max(1, k) ** 0.5


# nrhs <= 4 (and n x nrhs within 64 KiB of LDS) takes the one-workgroup-per-matrix solve
@pytest.mark.parametrize('n,nrhs,batch', [(1, 1, 1), (5, 3, 2), (33, 40, 3), (97, 97, 2), (450, 8, 2),
                                          (1000, 5, 1), (722, 2, 3), (722, 1, 2), (31, 4, 2), (64, 4, 1),
                                          (1000, 4, 1)])
def test_lu_solve_matches_numpy(n, nrhs, batch):
    rng = np.random.default_rng(n)
    A = crand(rng, batch, n, n)
    B = crand(rng, batch, n, nrhs)
    LU, ipiv, info = _native.zgetrf(torch.from_numpy(A.copy()).to(DEV))
    X = _native.zgetrs(LU, ipiv, torch.from_numpy(B.copy()).to(DEV)).cpu().numpy()
    assert int(info.abs().sum()) == 0
    want = np.linalg.solve(A, B)
    cond = np.linalg.cond(A).max()
    assert np.abs(X - want).max() <= 1e-15 * cond * n * np.abs(want).max()
    # residual check independent of conditioning
    assert np.abs(A @ X - B).max() <= 1e-12 * n * np.abs(A).max() * np.abs(X).max()


def test_lu_reports_singular():
    A = np.zeros((1, 8, 8), dtype=np.complex128)
    A[0, :4, :4] = np.eye(4)
    _, _, info = _native.zgetrf(torch.from_numpy(A).to(DEV))
    assert int(info[0]) == 5


def _check_eig(A, w, W):
    n = A.shape[-1]
    scale = np.abs(A).max()
    res = np.abs(A @ W - W * w[None, :]).max()
    assert res <= 1e-12 * n * scale, res
    ref = np.linalg.eigvals(A)
    # eigenvalue sets agree (greedy nearest matching)
    left = list(ref)
    for x in w:
        j = int(np.argmin(np.abs(np.asarray(left) - x)))
        assert abs(left[j] - x) <= 1e-9 * scale, (x, left[j])
        left.pop(j)
    assert np.allclose(np.linalg.norm(W, axis=0), 1.0, atol=1e-12)


@pytest.mark.parametrize('n,batch', [(1, 1), (2, 2), (3, 1), (10, 3), (50, 2), (98, 2), (257, 1),
                                     (450, 2), (722, 1), (1250, 1)])
def test_zgeev_random(n, batch):
    rng = np.random.default_rng(100 + n)
    A = crand(rng, batch, n, n)
    w, W, info = _native.zgeev(torch.from_numpy(A.copy()).to(DEV))
    assert int(info.abs().sum()) == 0
    w, W = w.cpu().numpy(), W.cpu().numpy()
    for b in range(batch):
        _check_eig(A[b], w[b], W[b])


def test_zgeev_diagonal_and_repeated():
    """Uniform layers give diagonal operators with repeated eigenvalues."""
    d = np.repeat(np.array([-3.0 + 0.1j, 2.0, 2.0, -1.5j]), 5)
    A = np.diag(d)[None].astype(np.complex128)
    w, W, info = _native.zgeev(torch.from_numpy(A.copy()).to(DEV))
    assert int(info[0]) == 0
    _check_eig(A[0], w.cpu().numpy()[0], W.cpu().numpy()[0])
    assert abs(abs(np.linalg.det(W.cpu().numpy()[0])) - 1.0) < 1e-12


def test_zgeev_repeated_blocks_keep_a_full_basis():
    """diag(B, B) has every eigenvalue twice (as a symmetric grating's operator does); the
    eigenvector matrix must still be a basis."""
    rng = np.random.default_rng(21)
    Bm = crand(rng, 60, 60)
    A = np.zeros((1, 120, 120), dtype=np.complex128)
    A[0, :60, :60] = Bm
    A[0, 60:, 60:] = Bm
    w, W, info = _native.zgeev(torch.from_numpy(A.copy()).to(DEV))
    assert int(info[0]) == 0
    w, W = w.cpu().numpy()[0], W.cpu().numpy()[0]
    _check_eig(A[0], w, W)
    assert np.linalg.cond(W) < 1e10
