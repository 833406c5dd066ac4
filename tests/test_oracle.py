"""Pin the CPU oracle against the reference's golden vectors and against closed forms.

These run on the CPU only. The oracle is the checker for every GPU parity test, so each of its
functions is pinned here by data the reference ships or by mathematics a mistake would break:
exact pixel integrals, uniform layers, Fresnel/Airy closed forms, energy conservation,
mode residuals.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import rcwa_oracle as orc

GOLDEN = Path(__file__).resolve().parent / 'golden'

# tests/references/test_reticolo_golden.py:57-93 (reference): inputs of the RETICOLO 1D cases.
_HALF = np.array([[[9.0] * 5 + [1.0] * 5]])
_QUARTER = np.array([[[12.25] * 3 + [1.0] * 9]])
_TWO_LAYER = np.array([[[9.0] * 6 + [1.0] * 6], [[4.840000000000001] * 8 + [1.0] * 4]])
RETICOLO = {
    'te_normal': dict(tetm='s', theta=0.0, top=1.0, bot=2.25, M=10, period=0.7, wl=0.9, eps=_HALF, d=[0.15]),
    'te_oblique': dict(tetm='s', theta=0.2, top=1.0, bot=2.25, M=10, period=0.7, wl=0.9, eps=_HALF, d=[0.15]),
    'tm_normal': dict(tetm='p', theta=0.0, top=1.0, bot=2.25, M=10, period=0.7, wl=0.9, eps=_HALF, d=[0.15]),
    'tm_oblique': dict(tetm='p', theta=0.2, top=1.0, bot=2.25, M=10, period=0.7, wl=0.9, eps=_HALF, d=[0.15]),
    'te_quarter': dict(tetm='s', theta=0.15, top=1.0, bot=2.25, M=12, period=0.8, wl=0.7, eps=_QUARTER, d=[0.25]),
    'tm_quarter': dict(tetm='p', theta=0.15, top=1.0, bot=2.25, M=12, period=0.8, wl=0.7, eps=_QUARTER, d=[0.25]),
    'te_two_layer': dict(tetm='s', theta=0.25, top=1.0, bot=4.0, M=10, period=0.7, wl=0.85, eps=_TWO_LAYER[::-1], d=[0.22, 0.15]),
    'tm_two_layer': dict(tetm='p', theta=0.25, top=1.0, bot=4.0, M=10, period=0.7, wl=0.85, eps=_TWO_LAYER[::-1], d=[0.22, 0.15]),
    'te_high_order': dict(tetm='s', theta=0.2, top=1.0, bot=2.25, M=30, period=0.7, wl=0.9, eps=_HALF, d=[0.15]),
    'tm_high_order': dict(tetm='p', theta=0.2, top=1.0, bot=2.25, M=30, period=0.7, wl=0.9, eps=_HALF, d=[0.15]),
}


def _rand(shape, seed):
    # tests/rcwa/test_golden_legacy_parity.py:221-229 (reference)
    return np.random.default_rng(seed).uniform(1.0, 4.0, size=shape) ** 2


# tests/rcwa/test_golden_legacy_parity.py:235-249 (reference); layers listed TOP to BOTTOM.
LEGACY = {
    'te_1layer': dict(pol='s', phi=None, eps=[[[9., 9., 9., 1., 1., 1.]]], d=[0.15], top=4., bot=1., theta=0., M=(6, 0), period=(0.77, 1.), wl=0.777),
    'tm_1layer': dict(pol='p', phi=None, eps=[[[9., 9., 9., 1., 1., 1.]]], d=[0.15], top=4., bot=1., theta=0., M=(6, 0), period=(0.77, 1.), wl=0.777),
    'te_oblique': dict(pol='s', phi=None, eps=[[[9., 9., 9., 1., 1., 1.]]], d=[0.2], top=1., bot=2.25, theta=0.3, M=(6, 0), period=(0.77, 1.), wl=0.777),
    'tm_oblique': dict(pol='p', phi=None, eps=[[[9., 9., 9., 1., 1., 1.]]], d=[0.2], top=1., bot=2.25, theta=0.3, M=(6, 0), period=(0.77, 1.), wl=0.777),
    'te_2layer': dict(pol='s', phi=None, eps=[[[9., 9., 1., 1., 9., 1.]], [[4., 4., 4., 1., 1., 1.]]], d=[0.15, 0.22], top=1., bot=1., theta=0., M=(5, 0), period=(0.7, 1.), wl=0.9),
    'te_same_mat': dict(pol='s', phi=None, eps=[[[9., 9., 1., 1., 9., 1.]], [[1., 1., 1., 1., 1., 1.]]], d=[0.2, 0.1], top=1., bot=1., theta=0., M=(5, 0), period=(0.7, 1.), wl=0.9),
    '2d_normal': dict(psi=0., phi=0., eps=_rand((1, 5, 5), 0), d=[0.2], top=1., bot=1., theta=0., M=(2, 2), period=(1., 1.), wl=0.9),
    '2d_oblique': dict(psi=0.9, phi=0.4, eps=_rand((1, 6, 6), 1), d=[0.3], top=1., bot=2.25, theta=0.3, M=(2, 2), period=(1., 1.2), wl=0.8),
    '2d_2layer': dict(psi=1.1, phi=0.5, eps=_rand((2, 4, 4), 2), d=[0.15, 0.22], top=1.44, bot=1., theta=0.2, M=(2, 2), period=(0.9, 0.9), wl=0.7),
    '2d_cfs': dict(psi=0.7, phi=0.3, eps=_rand((1, 5, 5), 4), d=[0.2], top=2.25, bot=1., theta=0.2, M=(2, 2), period=(1., 1.), wl=0.9, coef='integrated'),
}


def legacy_case(name):
    """Inputs of a legacy case, restated bottom-to-top, and the solve's pol argument."""
    c = LEGACY[name]
    eps = np.asarray(c['eps'], dtype=float)[::-1]
    d = list(c['d'])[::-1]
    pol = c['pol'] if 'pol' in c else (np.sin(c['psi']), np.cos(c['psi']))
    return c, eps, d, pol


def legacy_expectations(name, res_sub, kx, sub):
    """Golden efficiencies and converted amplitudes for one incidence variant.

    Conversions restate tests/rcwa/test_golden_legacy_parity.py:166-218 (reference): the
    amplitude factors of ``_amp_map``, the frame sign ``sigma(kx)*sigma_inc`` and the centre
    parity ``(-1)**(m+n)``.
    """
    g = np.load(GOLDEN / 'legacy_subset.npz')
    c = LEGACY[name]
    one_d = int(g[f'{name}__grating_type']) == 0
    n_top, n_bot = np.sqrt(c['top']), np.sqrt(c['bot'])
    if one_d:
        amap = (('R_s', 'r_s', 1.0), ('R_p', 'r_p', 1.0), ('T_s', 't_s', 1.0), ('T_p', 't_p', n_top / n_bot))
    else:
        amap = (('R_s', 'r_s', 1.0), ('R_p', 'r_p', 1 / (-1j * n_top)), ('T_s', 't_s', 1.0), ('T_p', 't_p', 1 / (-1j * n_bot)))
    kxr = np.real(kx)
    sigma = np.where(kxr < 0, -1.0, 1.0)
    frame = (sigma * sigma[kxr.size // 2]).reshape(1, -1)
    hu, hv = c['M']
    m, n = np.arange(-hu, hu + 1), np.arange(-hv, hv + 1)
    centre = np.where((n[:, None] + m[None, :]) % 2 == 1, -1.0, 1.0)
    effs = {k: g[f'{name}__{sub}__{gk}'] for gk, k in (('de_ri', 'R'), ('de_ri_s', 'R_s'), ('de_ri_p', 'R_p'),
                                                       ('de_ti', 'T'), ('de_ti_s', 'T_s'), ('de_ti_p', 'T_p'))}
    amps = {attr: g[f'{name}__{sub}__{key}'] * f * frame * centre for key, attr, f in amap}
    return effs, amps


def test_series_is_exact_pixel_integral():
    """Sampled coefficients equal the exact integral of the staircase (closed form), any order,
    including |k| > n where the DFT has wrapped (factorization.py:289-311)."""
    rng = np.random.default_rng(3)
    a = rng.uniform(1, 5, 7) + 1j * rng.uniform(0, 1, 7)
    n = a.size
    half = 6
    got = orc.raster_series(a, half)
    for i, k in enumerate(range(-2 * half, 2 * half + 1)):
        if k == 0:
            want = a.mean()
        else:
            j = np.arange(n)
            want = np.sum(a * (np.exp(-2j * np.pi * k * j / n) - np.exp(-2j * np.pi * k * (j + 1) / n))
                          / (2j * np.pi * k))
        assert abs(got[i] - want) < 1e-14, (k, got[i], want)


def test_uniform_layer_is_exactly_eps_identity():
    eps = np.full((7, 9), 3.3 + 0.2j)
    epx, epy, epz_i = orc.li_conv_matrices(eps, 3, 2)
    N = 35
    assert np.abs(epx - eps[0, 0] * np.eye(N)).max() < 1e-14
    assert np.abs(epy - eps[0, 0] * np.eye(N)).max() < 1e-14
    assert np.abs(epz_i - np.eye(N) / eps[0, 0]).max() < 1e-14


def test_y_invariant_layer_reduces_to_1d_li():
    """On a y-invariant raster epx is the 1D inverse rule per v order and epy the direct rule
    (factorization.py:610-617)."""
    row = np.array([2.0, 2.0, 7.0, 7.0, 7.0, 1.5])
    eps = np.tile(row, (4, 1))
    hu, hv = 3, 1
    epx, epy, _ = orc.li_conv_matrices(eps, hu, hv)
    tp = orc.toeplitz_positions(hu)
    nu = 2 * hu + 1
    inv_rule = np.linalg.inv(orc.raster_series(1 / row, hu)[tp].reshape(nu, nu))
    direct = orc.raster_series(row, hu)[tp].reshape(nu, nu)
    for b in range(2 * hv + 1):
        sl = slice(b * nu, (b + 1) * nu)
        assert np.abs(epx[sl, sl] - inv_rule).max() < 1e-13
        assert np.abs(epy[sl, sl] - direct).max() < 1e-13


@pytest.mark.parametrize('name', list(RETICOLO))
def test_oracle_matches_reticolo_1d(name):
    c = RETICOLO[name]
    g = np.load(GOLDEN / 'reticolo_1d.npz')
    res = orc.solve_raster(c['eps'], c['d'], c['top'], c['bot'], (c['period'], 1.0), c['wl'],
                           c['theta'], None, c['tetm'], (c['M'], 0))
    for side, arr in (('ri', res['incident']['R']), ('ti', res['incident']['T'])):
        idx = g[f'{name}__idx_{side}']
        want = g[f'{name}__de_{side}']
        got = np.real(arr).ravel()[idx]
        assert np.abs(got - want).max() < 1e-11, (name, side, np.abs(got - want).max())
        rest = np.delete(np.real(arr).ravel(), idx)
        assert np.abs(rest).max() < 1e-12


@pytest.mark.parametrize('name', list(LEGACY))
def test_oracle_matches_legacy_golden(name):
    c, eps, d, pol = legacy_case(name)
    res = orc.solve_raster(eps, d, c['top'], c['bot'], c['period'], c['wl'], c['theta'], c['phi'],
                           pol, c['M'], coefficients=c.get('coef', 'sampled'))
    hu, hv = c['M']
    kx = (orc.incident_kxy(c['top'], c['theta'], c['phi'])[0]
          + np.arange(-hu, hu + 1) * c['wl'] / c['period'][0])
    subs = (('res', 'incident'),) if res['path'].startswith('1d') else (
        ('res', 'incident'), ('res_te_inc', 's_inc'), ('res_tm_inc', 'p_inc'))
    for sub, attr in subs:
        effs, amps = legacy_expectations(name, None, kx, sub)
        got = res[attr]
        for k, want in effs.items():
            assert np.abs(np.reshape(want, got[k].shape) - got[k]).max() < 1e-12, (sub, k)
        for k, want in amps.items():
            assert np.abs(np.reshape(want, got[k].shape) - got[k]).max() < 1e-12, (sub, k)


@pytest.mark.parametrize('name', ['2d_normal', '2d_oblique', '2d_2layer', '2d_cfs'])
def test_oracle_epz_i_and_q_match_legacy(name):
    g = np.load(GOLDEN / 'legacy_subset.npz')
    c, eps, d, pol = legacy_case(name)
    hu, hv = c['M']
    coef = c.get('coef', 'sampled')
    res = orc.solve_raster(eps, d, c['top'], c['bot'], c['period'], c['wl'], c['theta'], c['phi'],
                           pol, c['M'], coefficients=coef)
    conv = [orc.conv_rule('li', coef)(e, hu, hv) for e in eps]
    kx, ky = orc.order_kxy_2d(c['wl'], c['period'], c['top'], c['theta'], c['phi'], hu, hv)
    kxf = np.broadcast_to(kx[None, :], (2 * hv + 1, 2 * hu + 1)).reshape(-1)
    kyf = np.broadcast_to(ky[:, None], (2 * hv + 1, 2 * hu + 1)).reshape(-1)
    L = len(d)
    for i in range(L):
        # the golden's layer i is our layer i: legacy_case reverses the fixture's top-to-bottom
        # structure list into our bottom-to-top order, and the fixture indexes its per-layer
        # records in the order the reference's kernel received them, which is that same order
        # (on 2d_2layer the other layer's epz_i differs by 7e-2, so a swap fails here)
        assert np.abs(g[f'{name}__layer{i}__epz_i'] - conv[i][2]).max() < 1e-12
        W, V, q, _, _ = orc.layer_modes_2d(kxf, kyf, *conv[i])
        want_q = np.sort(np.abs(g[f'{name}__layer{i}__q']))
        assert np.abs(want_q - np.sort(np.abs(q))).max() < 1e-10
    assert res['path'] == '2d'


def test_oracle_matches_reticolo_li_nesting_totals():
    """2D crossed grating vs RETICOLO (fixed nesting), tests/references/test_reticolo_li_nesting.py."""
    spec = json.loads((GOLDEN / 'li_nesting_fixed.json').read_text())
    n, period = 650, 325.0
    pixel = period / n
    x0, x1 = round((period - 200.0) / 2 / pixel), round((period + 200.0) / 2 / pixel)
    y0, y1 = round((period - 150.0) / 2 / pixel), round((period + 150.0) / 2 / pixel)
    eps = np.ones((1, n, n))
    eps[0, y0:y1, x0:x1] = 6.0
    for M, (te, tm) in spec['total_T_te_tm'].items():
        M = int(M)
        for psi, want in ((np.pi / 2, te), (0.0, tm)):
            res = orc.solve_raster(eps, [600.0], 1.0, 1.46 ** 2, (period, period), 532.0, 0.0, 0.0,
                                   (np.sin(psi), np.cos(psi)), (M, M))
            got = float(np.sum(res['incident']['T']))
            assert abs(got - want) < spec['tolerance'], (M, psi, got, want)


def test_uniform_slab_matches_airy():
    """Uniform lossy slab on a substrate: zeroth-order R/T equal the Airy closed form."""
    eps_l, eps_sub, d, wl, theta = 4.0 + 0.3j, 2.25, 0.37, 0.8, 0.4
    eps = np.full((1, 3, 3), eps_l)
    for pol, name in (((1.0, 0.0), 's'), ((0.0, 1.0), 'p')):
        res = orc.solve_raster(eps, [d], 1.0, eps_sub, (0.5, 0.5), wl, theta, 0.0, pol, (1, 1))
        k0 = 2 * np.pi / wl
        kx = np.sin(theta)
        kz = [np.sqrt(e - kx ** 2 + 0j) for e in (1.0, eps_l, eps_sub)]
        kz = [k if (k.real + k.imag) >= 0 else -k for k in kz]
        eps_list = (1.0, eps_l, eps_sub)
        adm = kz if name == 's' else [k / e for k, e in zip(kz, eps_list)]
        r01 = (adm[0] - adm[1]) / (adm[0] + adm[1])
        r12 = (adm[1] - adm[2]) / (adm[1] + adm[2])
        t01 = 2 * adm[0] / (adm[0] + adm[1])
        t12 = 2 * adm[1] / (adm[1] + adm[2])
        ph = np.exp(1j * k0 * kz[1] * d)
        r = (r01 + r12 * ph ** 2) / (1 + r01 * r12 * ph ** 2)
        t = t01 * t12 * ph / (1 + r01 * r12 * ph ** 2)
        R = abs(r) ** 2
        # s: t is the E ratio, power ~ |E|^2 Re(kz); p: t is the H ratio, power ~ |H|^2 Re(kz/eps)
        T = abs(t) ** 2 * adm[2].real / adm[0].real
        got = res['incident']
        c = got['R'].shape[0] // 2, got['R'].shape[1] // 2
        assert abs(got['R'][c] - R) < 1e-12, (name, got['R'][c], R)
        assert abs(got['T'][c] - T) < 1e-12, (name, got['T'][c], T)


def test_lossless_2d_stack_conserves_energy():
    rng = np.random.default_rng(11)
    eps = rng.choice([1.0, 12.25], size=(2, 16, 16))
    res = orc.solve_raster(eps, [0.2, 0.31], 1.0, 2.25, (1.0, 1.0), 0.83, 0.3, 1.1, 'p', (3, 3))
    for k in ('incident', 's_inc', 'p_inc'):
        tot = res[k]['R'].sum() + res[k]['T'].sum()
        assert abs(tot - 1) < 1e-10, (k, tot)


def test_modes_satisfy_both_maxwell_factors():
    """Omega2 W = W q^2 and Omega_R W = V q (tests/rcwa/test_solve_2d_core.py:195 idea)."""
    rng = np.random.default_rng(5)
    eps = rng.uniform(1, 6, size=(8, 8)) + 1j * rng.uniform(0, 0.3, size=(8, 8))
    epx, epy, epz_i = orc.li_conv_matrices(eps, 2, 2)
    kx, ky = orc.order_kxy_2d(0.7, (1.0, 1.0), 1.0, 0.2, 0.3, 2, 2)
    kxf = np.broadcast_to(kx[None, :], (5, 5)).reshape(-1)
    kyf = np.broadcast_to(ky[:, None], (5, 5)).reshape(-1)
    W, V, q, O2, OR = orc.layer_modes_2d(kxf, kyf, epx, epy, epz_i)
    assert np.abs(O2 @ W - W * q ** 2).max() < 1e-11 * np.abs(O2).max()
    assert np.abs(OR @ W - V * q).max() < 1e-11 * np.abs(OR).max()
    assert np.all(q.real >= 0)


# ------------------------------------------------------------- integrated (segment) coefficients

def test_segment_series_is_the_closed_form_integral():
    """Three segments on [0, 0.2), [0.2, 0.7), [0.7, 1): c_k against the integral written out."""
    a = np.array([2.0, 5.0 - 1j, 1.5])
    edges = np.array([0.2, 0.7, 1.0])
    got = orc.segment_series(a, edges, 3)
    left = np.array([0.0, 0.2, 0.7])
    for i, k in enumerate(range(-6, 7)):
        if k == 0:
            want = np.sum(a * (edges - left))
        else:
            want = np.sum(a * (np.exp(-2j * np.pi * k * left) - np.exp(-2j * np.pi * k * edges))
                          / (2j * np.pi * k))
        assert abs(got[i] - want) < 1e-15 * 10


def test_segment_series_on_pixel_edges_is_the_pixel_series():
    rng = np.random.default_rng(9)
    a = rng.uniform(1, 9, (3, 12)) + 1j * rng.uniform(0, 1, (3, 12))
    got = orc.segment_series(a, np.arange(1, 13) / 12.0, 7)
    assert np.abs(got - orc.raster_series(a, 7)).max() < 1e-13


@pytest.mark.parametrize('rule', ['li', 'laurent'])
@pytest.mark.parametrize('shape,hu,hv', [((16, 16), 4, 3), ((1, 40), 8, 0), ((9, 7), 2, 4)])
def test_integrated_rules_equal_the_sampled_ones_on_a_raster(rule, shape, hu, hv):
    """A raster is a staircase; both series integrate it exactly, the integrated one after
    compressing runs of equal columns and rows, so the two builds agree to round-off."""
    rng = np.random.default_rng(shape[0] * 31 + hu)
    eps = np.where(rng.uniform(size=shape) < 0.5, 2.25, 12.25 + 0.3j).astype(np.complex128)
    eps[:, : shape[1] // 3] = 1.0                                   # wide runs for the compression
    row_idx, col_idx, x, y = orc.compression_indices(eps)
    assert len(col_idx) < shape[1]
    for got, want in zip(orc.conv_rule(rule, 'integrated')(eps, hu, hv), orc.conv_rule(rule)(eps, hu, hv)):
        assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


def test_integrated_uniform_layer_is_exactly_eps_times_identity():
    eps = np.full((6, 5), 3.5 + 0.25j)
    for m in orc.li_conv_matrices_integrated(eps, 2, 2)[:2]:
        np.testing.assert_array_equal(m, (3.5 + 0.25j) * np.eye(25))


# ------------------------------------------------------------- the stack's scattering matrix

def _star(Sa, Sb):
    return dict(zip(('S11', 'S12', 'S21', 'S22'),
                    orc.redheffer_star(np.eye(Sa['S11'].shape[0]), [Sa[k] for k in ('S11', 'S12', 'S21', 'S22')],
                                       [Sb[k] for k in ('S11', 'S12', 'S21', 'S22')])))


@pytest.mark.parametrize('name', ['2d_normal', '2d_oblique', '2d_2layer', '2d_cfs'])
def test_scattering_specular_columns_are_the_pinned_amplitudes(name):
    """The published blocks read on the specular column are the amplitudes the legacy golden pins:
    S11 e_s = r_s, S11 e_p = -r_p (p port variable), S21 e_s = t_s, S21 e_p = t_p n_top / n_bot
    (reference scattering.py:524-536, in the stated units) -- the cross terms as well."""
    c, eps, d, pol = legacy_case(name)
    res = orc.solve_raster(eps, d, c['top'], c['bot'], c['period'], c['wl'], c['theta'], c['phi'], pol,
                           c['M'], coefficients=c.get('coef', 'sampled'))
    S = res['scattering']
    N = S['S11'].shape[0] // 2
    c0 = N // 2
    n_ref, n_trn = np.sqrt(complex(c['top'])), np.sqrt(complex(c['bot']))
    mi = -1j * n_ref
    s, p = res['s_inc'], res['p_inc']
    pairs = [(S['S11'][:N, c0], s['r_s']), (-S['S11'][N:, c0] / mi, s['r_p']),
             (S['S21'][:N, c0], s['t_s']), (S['S21'][N:, c0] / (-1j * n_trn), s['t_p']),
             (S['S11'][:N, N + c0] * mi, p['r_s']), (-S['S11'][N:, N + c0], p['r_p']),
             (S['S21'][:N, N + c0] * mi, p['t_s']), (S['S21'][N:, N + c0] * n_ref / n_trn, p['t_p'])]
    for got, want in pairs:
        assert np.abs(got - np.asarray(want).reshape(-1)).max() < 1e-12


@pytest.mark.parametrize('name', ['te_1layer', 'tm_1layer', 'te_oblique', 'tm_oblique'])
def test_scattering_1d_specular_column_is_the_pinned_amplitude(name):
    c, eps, d, pol = legacy_case(name)
    res = orc.solve_raster(eps, d, c['top'], c['bot'], c['period'], c['wl'], c['theta'], c['phi'], pol, c['M'])
    S, inc = res['scattering'], res['incident']
    c0 = S['S11'].shape[0] // 2
    n_ref, n_trn = np.sqrt(complex(c['top'])), np.sqrt(complex(c['bot']))
    if c['pol'] == 's':
        assert np.abs(S['S11'][:, c0] - inc['r_s'].reshape(-1)).max() < 1e-12
        assert np.abs(S['S21'][:, c0] - inc['t_s'].reshape(-1)).max() < 1e-12
    else:
        assert np.abs(-S['S11'][:, c0] - inc['r_p'].reshape(-1)).max() < 1e-12
        assert np.abs(S['S21'][:, c0] * n_ref / n_trn - inc['t_p'].reshape(-1)).max() < 1e-12


@pytest.mark.parametrize('pol,M,phi,shape', [((0.6, 0.8j), (2, 2), 0.4, (3, 5, 6)), ('s', (5, 0), None, (3, 1, 8)),
                                             ('p', (5, 0), None, (3, 1, 8))])
def test_scattering_of_a_stack_is_the_cascade_of_its_pieces(pol, M, phi, shape):
    """Cut a 3-layer stack after its top layer: the whole stack's matrix is the top piece's
    followed by the bottom piece's, the joint half-space being the top medium (reference
    ScatteringMatrix.cascade, scattering.py:550-604). Every block and column is compared."""
    rng = np.random.default_rng(41)
    eps = rng.uniform(1.5, 9.0, shape) + 1j * rng.uniform(0.0, 0.3, shape)
    d = [0.12, 0.2, 0.15]               # bottom-to-top
    top, bot, theta, wl, period = 1.0, 2.25, 0.25, 0.83, (0.7, 0.65)
    whole = orc.solve_raster(eps, d, top, bot, period, wl, theta, phi, pol, M)['scattering']
    upper = orc.solve_raster(eps[2:], d[2:], top, top, period, wl, theta, phi, pol, M)['scattering']
    lower = orc.solve_raster(eps[:2], d[:2], top, bot, period, wl, theta, phi, pol, M)['scattering']
    joined = _star(upper, lower)
    scale = max(np.abs(whole[k]).max() for k in joined)
    for k in joined:
        assert np.abs(joined[k] - whole[k]).max() < 1e-12 * scale, k


# ------------------------------------------------------------------------------ bottom incidence

@pytest.mark.parametrize('pol,M,phi,shape', [((0.6, 0.8j), (2, 2), 0.4, (2, 5, 6)), ('s', (5, 0), None, (2, 1, 8)),
                                             ('p', (5, 0), None, (2, 1, 8))])
def test_bottom_incidence_is_top_incidence_on_the_mirrored_stack(pol, M, phi, shape):
    """The front-side read of the kernel cascade (reference scattering.py:943-949, backside off)
    equals the backside read of the stack mirrored in z -- half-spaces swapped, layers reversed:
    every order's efficiencies and amplitudes, and the published scattering blocks."""
    rng = np.random.default_rng(23)
    eps = rng.uniform(1.5, 9.0, shape) + 1j * rng.uniform(0.0, 0.3, shape)
    d = [0.12, 0.2]
    top, bot, theta, wl, period = 1.44, 2.25 + 0.05j, 0.25, 0.83, (0.7, 0.65)
    up = orc.solve_raster(eps, d, top, bot, period, wl, theta, phi, pol, M, side='bottom')
    mirror = orc.solve_raster(eps[::-1], d[::-1], bot, top, period, wl, theta, phi, pol, M)
    for k in ('R', 'T', 'r_s', 'r_p', 't_s', 't_p'):
        assert np.abs(np.asarray(up['incident'][k]) - np.asarray(mirror['incident'][k])).max() < 1e-12, k
    for k in ('S11', 'S12', 'S21', 'S22'):
        assert np.abs(up['scattering'][k] - mirror['scattering'][k]).max() < 1e-12, k


def test_bottom_incidence_conserves_energy_on_a_lossless_stack():
    rng = np.random.default_rng(29)
    eps = rng.uniform(1.5, 9.0, (2, 6, 6))
    res = orc.solve_raster(eps, [0.15, 0.2], 1.0, 2.25, (0.8, 0.75), 0.9, 0.2, 0.5, (0.8, 0.6j), (2, 2),
                           side='bottom')['incident']
    assert abs(res['R'].sum() + res['T'].sum() - 1.0) < 1e-12


@pytest.mark.parametrize('pol,M,phi,shape', [((0.6, 0.8j), (2, 2), 0.4, (2, 5, 6)), ('p', (5, 0), None, (2, 1, 8))])
def test_direct_wavenumbers_are_the_angles_they_spell(pol, M, phi, shape):
    """PlaneWave(kx=, ky=) in radians per length unit is k0 times the angles' kx/k0, ky/k0
    (reference engine.py:2755-2758): the same solve to round-off."""
    rng = np.random.default_rng(37)
    eps = rng.uniform(1.5, 9.0, shape) + 1j * rng.uniform(0.0, 0.3, shape)
    top, theta, wl = 1.44, 0.3, 0.83
    n = np.sqrt(top)
    k0 = 2 * np.pi / wl
    p = 0.0 if phi is None else phi
    k = (n * np.sin(theta) * np.cos(p) * k0, n * np.sin(theta) * np.sin(p) * k0)
    a = orc.solve_raster(eps, [0.12, 0.2], top, 2.25, (0.7, 0.65), wl, theta, phi, pol, M)['incident']
    b = orc.solve_raster(eps, [0.12, 0.2], top, 2.25, (0.7, 0.65), wl, None, None if phi is None else None, pol, M,
                         k_direct=k)['incident']
    for key in ('R', 'T'):
        assert np.abs(np.asarray(a[key]) - np.asarray(b[key])).max() < 1e-12, key
