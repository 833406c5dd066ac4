"""Pin the oracle's Laurent rule (``factorization='laurent'``) on the CPU.

* fmmax 1.7.1 ``Formulation.FFT`` goldens (``tests/golden/fmmax_laurent.npz``, extracted from
  the reference's ``tests/references/golden_fmmax_laurent.npz``): the same operator as Laurent's
  rule, computed with the midpoint rule at R samples per pixel. The gap to the exact pixel
  integral is a series in 1/R^2, so extrapolating over R = 32, 64, 128 removes it, and the
  distance to the estimate from R = 16, 32, 64 bounds what is left (reference
  ``tests/references/test_fmmax_laurent_golden.py:83-89``). Totals and every order of both
  polarizations are held to that bound; the mirrored order index must miss by far more.
* The legacy golden ``epz_i`` (Li's z slot, which is Laurent's too): ``epx @ epz_i = I``.
* Closed forms: a uniform layer is ``eps * I``; a y-invariant grating's TE reads ``epy``, where
  both rules build ``[[eps]]``, so TE equals Li's answer to round-off while TM does not.
"""
import functools
from pathlib import Path

import numpy as np
import pytest

from oracle import rcwa_oracle as orc

GOLDEN = Path(__file__).resolve().parent / 'golden'
ROUND_OFF = 1e-12

# tests/references/_fmmax_laurent_cases.py:39-71 (reference): the scene table.
WL, THETA_DEG, THICK, ORDER, PIXELS = 500.0, 11.0, 240.0, 7, 16


def _rod(eps_rod, eps_bg, rows, cols):
    a = np.full((PIXELS, PIXELS), eps_bg, dtype=complex)
    a[rows[0]:rows[1], cols[0]:cols[1]] = eps_rod
    return a


def _cross(eps_in, eps_bg):
    a = np.full((PIXELS, PIXELS), eps_bg, dtype=complex)
    a[3:13, 5:8] = eps_in          # a tall arm
    a[6:9, 2:14] = eps_in          # a wide arm, off the tall one's centre
    return a


CASES = {   # name -> (period (x, y) nm or a 2 x 2 lattice with the lattice vectors as columns, raster [v, u])
    'rod_eps12_rect': ((700.0, 560.0), _rod(12.25, 1.0, (4, 11), (2, 12))),
    'hole_metal_rect': ((600.0, 600.0), _rod(1.0, -16.49 + 1.06j, (5, 12), (3, 9))),
    'cross_eps6_hex': (np.array([[640.0, 320.0], [0.0, 320.0 * np.sqrt(3.0)]]), _cross(6.0, 1.0)),
}
RECT = [name for name, (cell, _) in CASES.items() if np.shape(cell) == (2,)]


def golden():
    return np.load(GOLDEN / 'fmmax_laurent.npz')


def _richardson(values):
    """Three values at samplings R, 2R, 4R -> the limit, the 1/R^2 and 1/R^4 terms removed."""
    a, b, c = values
    return (64.0 * c - 20.0 * b + a) / 45.0


def extrapolated(g, name, what):
    levels = [int(v) for v in g['__levels']]
    values = [g[f'{name}__phase64__R{r}__{what}'] for r in levels]
    fine, coarse = _richardson(values[1:]), _richardson(values[:3])
    return fine, float(np.abs(fine - coarse).max()) + ROUND_OFF


@functools.lru_cache(maxsize=None)
def solve(name, factorization='laurent'):
    period, raster = CASES[name]
    return orc.solve_raster(raster[None], [THICK], 1.0, 1.0, period, WL, np.deg2rad(THETA_DEG), 0.0,
                            's', (ORDER, ORDER), factorization)


def totals(res):
    return np.array([res['s_inc']['R'].sum(), res['s_inc']['T'].sum(),
                     res['p_inc']['R'].sum(), res['p_inc']['T'].sum()])


@pytest.mark.parametrize('name', list(CASES))
def test_the_gap_to_each_sampling_falls_fourfold(name):
    """The premise of the extrapolation: the raw gap shrinks 4x per doubling of R."""
    g = golden()
    got = totals(solve(name))
    gaps = [np.abs(g[f'{name}__phase64__R{int(r)}__totals'] - got).max() for r in g['__levels']]
    ratios = np.asarray(gaps[:-1]) / np.asarray(gaps[1:])
    assert np.all((ratios > 3.9) & (ratios < 4.1)), (gaps, ratios)


@pytest.mark.parametrize('name', list(CASES))
def test_laurent_totals_match_fmmax(name):
    want, tol = extrapolated(golden(), name, 'totals')
    assert np.abs(totals(solve(name)) - want).max() < tol


@pytest.mark.parametrize('name', list(CASES))
def test_laurent_orders_match_fmmax_and_the_mirror_misses(name):
    g = golden()
    want, tol = extrapolated(g, name, 'per_order')          # [pol, R/T, order]
    res = solve(name)
    worst = mirrored = 0.0
    for pol, sub in enumerate(('s_inc', 'p_inc')):
        for kind, grid in enumerate((res[sub]['R'], res[sub]['T'])):
            for i, (m, n) in enumerate(g[f'{name}__orders']):
                worst = max(worst, abs(grid[n + ORDER, m + ORDER] - want[pol, kind, i]))
                mirrored = max(mirrored, abs(grid[-n + ORDER, -m + ORDER] - want[pol, kind, i]))
    assert worst < tol, (worst, tol)
    assert mirrored > 1e3 * tol, (mirrored, tol)


@pytest.mark.parametrize('name', RECT)
def test_the_golden_tells_li_from_laurent(name):
    want, tol = extrapolated(golden(), name, 'totals')
    assert np.abs(totals(solve(name, 'li')) - want).max() > 1e6 * tol


@pytest.mark.parametrize('name', ['rod_eps12_rect', 'cross_eps6_hex'])
def test_laurent_conserves_energy_on_the_lossless_cells(name):
    got = totals(solve(name))
    assert abs(got[0] + got[1] - 1.0) < 1e-12 and abs(got[2] + got[3] - 1.0) < 1e-12


def _legacy_eps(seed, shape):
    # tests/rcwa/test_golden_legacy_parity.py:221-229 (reference)
    return np.random.default_rng(seed).uniform(1.0, 4.0, size=shape) ** 2


@pytest.mark.parametrize('name,seed,shape', [('2d_normal', 0, (1, 5, 5)), ('2d_oblique', 1, (1, 6, 6)),
                                             ('2d_2layer', 2, (2, 4, 4))])
def test_direct_rule_inverts_the_legacy_golden_epz_i(name, seed, shape):
    """The legacy golden holds Li's z slot, [[eps]]^-1, per layer; Laurent's in-plane slots are
    [[eps]] itself, so their product with the golden is the identity."""
    g = np.load(GOLDEN / 'legacy_subset.npz')
    eps = _legacy_eps(seed, shape)
    for layer in eps:
        epx, epy, epz_i = orc.laurent_conv_matrices(layer, 2, 2)
        assert epx is epy
        I = np.eye(epx.shape[0])
        best = min(np.abs(epx @ g[f'{name}__layer{j}__epz_i'] - I).max() for j in range(shape[0]))
        assert best < 1e-11
        np.testing.assert_array_equal(epz_i, orc.li_conv_matrices(layer, 2, 2)[2])


def test_uniform_layer_is_eps_times_identity():
    eps = 2.25 + 0.1j
    epx, epy, epz_i = orc.laurent_conv_matrices(np.full((5, 7), eps), 3, 2)
    I = np.eye(35)
    assert np.abs(epx - eps * I).max() < 1e-15
    assert np.abs(epz_i - I / eps).max() < 1e-15


@pytest.mark.parametrize('theta', [0.0, 0.2])
def test_y_invariant_te_is_li_and_tm_is_not(theta):
    """TE reads epy alone, where both rules build [[eps]]; TM reads epx, where Li takes the
    inverse rule -- the slowly converging case of L. Li, JOSA A 13, 1870 (1996)."""
    eps = np.array([[[9.0] * 5 + [1.0] * 5]])
    args = (eps, [0.15], 1.0, 2.25, (0.7, 1.0), 0.9, theta, None)
    te_l = orc.solve_raster(*args, 's', (10, 0), 'laurent')['incident']
    te_li = orc.solve_raster(*args, 's', (10, 0), 'li')['incident']
    tm_l = orc.solve_raster(*args, 'p', (10, 0), 'laurent')['incident']
    tm_li = orc.solve_raster(*args, 'p', (10, 0), 'li')['incident']
    for k in ('R', 'T'):
        assert np.abs(te_l[k] - te_li[k]).max() < 1e-12
    assert max(np.abs(tm_l[k] - tm_li[k]).max() for k in ('R', 'T')) > 1e-4
    assert abs(tm_l['R'].sum() + tm_l['T'].sum() - 1.0) < 1e-12
