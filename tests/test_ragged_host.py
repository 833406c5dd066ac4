"""Host side of the ragged cascade on the CPU: the layer-major pair numbering of
``wb_solve_from_modes_ragged`` (points by non-increasing layer count, layer l of point p at
base_l + p) and the staging that maps each bucket's (P, L) pairs into it."""
import numpy as np
import torch

from wavebit_amd import _native
from wavebit_amd.rcwa import _engine as eng


def test_ragged_order_numbers_pairs_layer_major():
    order, (pp, ll) = _native.ragged_order([2, 3, 1, 3])
    assert list(order) == [1, 3, 0, 2]          # stable, non-increasing
    # steps: 4 points have layer 0, 3 have layer 1, 2 have layer 2
    assert list(ll) == [0, 0, 0, 0, 1, 1, 1, 2, 2]
    assert list(pp) == [0, 1, 2, 3, 0, 1, 2, 0, 1]


def test_ragged_inputs_map_every_bucket_pair_once():
    rng = np.random.default_rng(0)
    buckets, tables = [], []
    spec = [(2, 3), (4, 1), (1, 2), (4, 2)]
    base = 0
    for L, P in spec:
        b = eng.Bucket(hu=1, hv=1, path=eng.PATH_2D, n_layers=L)
        for _ in range(P):
            b.points.append(eng.Point(stack=0, wavelength=0.7, theta=0.1, phi=0.2,
                                      polarization=(1.0, 0.0), eps_top=1.0, eps_bottom=2.0,
                                      period=(1.0, 1.0), thickness=tuple(rng.uniform(0.1, 1, L))))
        buckets.append(b)
        tables.append(np.arange(base, base + P * L, dtype=np.int32).reshape(P, L))
        base += P * L
    (pts, thick, idx), lays, starts, pairs = eng._ragged_inputs(buckets, tables, torch.device('cpu'))
    assert lays == sorted(lays, reverse=True) and len(lays) == 8
    B = sum(lays)
    assert thick.shape == (B,) and idx.shape == (B,) and pts.shape == (8, 16)
    seen = set()
    for b, s, pair, tab in zip(buckets, starts, pairs, tables):
        P, L = len(b.points), b.n_layers
        assert tuple(pair.shape) == (P, L)
        for p in range(P):
            for l in range(L):
                k = int(pair[p, l])
                seen.add(k)
                assert float(thick[k]) == b.points[p].thickness[l]
                assert int(idx[k]) == tab[p, l]
                # layer-major: pair k is layer l of sorted point s + p
                steps = [sum(1 for x in lays if x > j) for j in range(max(lays))]
                assert k == sum(steps[:l]) + s + p
    assert seen == set(range(B))
