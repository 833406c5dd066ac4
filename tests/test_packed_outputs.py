"""The solve's four outputs as views of one buffer, and their one-copy transfer to the host
(host logic only: CPU tensors stand in for device memory)."""
import numpy as np
import pytest
import torch

from wavebit_amd import _native


@pytest.mark.parametrize('P, L, N', [(1, 1, 21), (3, 2, 25), (2, 3, 1), (5, 7, 9)])
def test_packed_views_round_trip_through_host_copy(P, L, N):
    eff, amp, pe, flags = _native._packed_outputs(P, L, N, torch.device('cpu'))
    assert eff.shape == (P, 3, 4, N) and amp.shape == (P, 3, 4, N)
    assert pe.shape == (P, 3) and flags.shape == (P, L)
    assert amp.dtype == torch.complex128 and flags.dtype == torch.int32
    for t in (eff, amp, pe, flags):
        assert t.is_contiguous()
    # the views do not overlap: each keeps what was written to it
    rng = np.random.default_rng(P * 100 + L * 10 + N)
    e = rng.normal(size=eff.shape)
    a = rng.normal(size=amp.shape) + 1j * rng.normal(size=amp.shape)
    p = rng.normal(size=pe.shape)
    f = rng.integers(-5, 9, flags.shape).astype(np.int32)
    eff.copy_(torch.from_numpy(e))
    amp.copy_(torch.from_numpy(a))
    pe.copy_(torch.from_numpy(p))
    flags.copy_(torch.from_numpy(f))
    he, ha, hp, hf = _native.host_copy(eff, amp, pe, flags)
    assert np.array_equal(he, e) and np.array_equal(ha, a)
    assert np.array_equal(hp, p) and np.array_equal(hf, f)
    # a clone of the buffer re-viewed (what a replayed graph returns) copies back the same
    buf = eff._wb_packed[0].clone()
    got = _native.host_copy(*_native.packed_views(buf, P, L, N))
    for x, y in zip(got, (e, a, p, f)):
        assert np.array_equal(x, y)


def test_unpacked_outputs_copy_one_by_one():
    eff = torch.arange(24, dtype=torch.float64).reshape(1, 3, 4, 2)
    amp = torch.ones((1, 3, 4, 2), dtype=torch.complex128)
    pe = torch.zeros((1, 3), dtype=torch.float64)
    flags = torch.full((1, 2), 3, dtype=torch.int32)
    he, ha, hp, hf = _native.host_copy(eff, amp, pe, flags)
    assert np.array_equal(he, eff.numpy()) and np.array_equal(hf, flags.numpy())
