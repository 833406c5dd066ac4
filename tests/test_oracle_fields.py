"""Pin the oracle's field reconstruction (``oracle/rcwa_fields.py``) on the CPU.

Three kinds of evidence, each able to fail a plausible mistake:

* the reference's own frozen in-layer field cells (``tests/fixtures/legacy_rcwa/golden_field.npz``,
  extracted to ``tests/golden/legacy_field.npz``), every path and incidence variant, at the
  reference's own 1e-12 (``tests/rcwa/test_field_core.py:117-140``) -- a swapped component, a
  wrong sign, a transposed raster or a wrong layer order is O(1) there;
* fmmax 1.7.1's in-layer fields (``tests/references/golden_fmmax_field.npz``): exact (1e-13,
  1e-12 per component) on uniform layers, 1e-4 on the x grating, 5e-2 on the crossed patch,
  with the fitted amplitude ratio equal to its closed form (``-cos(theta)`` for s, ``-1`` for p,
  times the corner-to-centre phase; ``tests/references/test_fmmax_field_golden.py:95-115``);
* physics: Faraday's law per order, tangential continuity across internal faces, and the
  Poynting flux through every height of a lossless stack equal to the solve's transmission.
"""
from pathlib import Path

import numpy as np
import pytest

from oracle import rcwa_fields as orf
from oracle import rcwa_oracle as orc

GOLDEN = Path(__file__).resolve().parent / 'golden'


def _rand_eps(shape, seed):
    # tests/fixtures/legacy_rcwa/gen_goldens_field.py:56-70 (reference): index draws, squared
    return np.random.default_rng(seed).uniform(1, 4, shape) ** 2


# tests/fixtures/legacy_rcwa/gen_goldens_field.py:84-108 (reference); layers listed TOP to
# BOTTOM there, and a one-entry period is a square lattice of that side.
LEGACY = {
    'te_1d': dict(pol='s', phi=None, eps=[[[9., 9., 9., 1., 1., 1.]]], d=[0.15], top=4., bot=1.,
                  theta=0.2, M=(4, 0), period=(0.77, 0.77), wl=0.777),
    'tm_1d': dict(pol='p', phi=None, eps=[[[9., 9., 9., 1., 1., 1.]]], d=[0.15], top=4., bot=1.,
                  theta=0.2, M=(4, 0), period=(0.77, 0.77), wl=0.777),
    '1d_2layer': dict(pol='s', phi=None, eps=[[[9., 9., 1., 1., 9., 1.]], [[4., 4., 4., 1., 1., 1.]]],
                      d=[0.15, 0.22], top=1., bot=1., theta=0.1, M=(3, 0), period=(0.7, 0.7), wl=0.9),
    'conical': dict(psi=0.4, phi=0.3, eps=[[[9., 9., 9., 1., 1., 1.]]], d=[0.15], top=4., bot=1.,
                    theta=0.2, M=(4, 0), period=(0.77, 0.77), wl=0.777),
    '2d': dict(psi=0.9, phi=0.4, eps=_rand_eps((1, 5, 5), 0), d=[0.2], top=1., bot=2.25, theta=0.3,
               M=(2, 2), period=(1., 1.), wl=0.9),
    '2d_2layer': dict(psi=1.1, phi=0.5, eps=_rand_eps((2, 4, 4), 1), d=[0.15, 0.22], top=1.44, bot=1.,
                      theta=0.2, M=(2, 2), period=(0.9, 0.9), wl=0.7),
    # the reference solves this one with its first-order formulation; the field is the same
    '2d_first_order': dict(psi=0.7, phi=0.35, eps=_rand_eps((2, 4, 4), 2), d=[0.18, 0.13], top=1.,
                           bot=1.9599999999999997, theta=0.25, M=(2, 1), period=(0.85, 0.8), wl=0.66),
}
RES = (5, 4, 6)      # sample_x, sample_y, sample_z (gen_goldens_field.py:117)
PARITY = {6: np.array([1, 1, -1, -1, -1, 1]), 'te': np.array([1, -1, 1]), 'tm': np.array([-1, 1, -1])}


def legacy_state(name):
    c = LEGACY[name]
    eps = np.asarray(c['eps'], dtype=float)[::-1]
    pol = c['pol'] if 'pol' in c else (np.sin(c['psi']), np.cos(c['psi']))
    return orf.solve_state(eps, list(c['d'])[::-1], c['top'], c['bot'], c['period'], c['wl'],
                           c['theta'], c['phi'], pol, c['M'])


def legacy_expected(name, value):
    """The archive restated in the public frame with the public incidence: reversed along z,
    the mirror parity (E_z, H_x, H_y flip), times the centre phase ``exp(-i k_inc . o)`` and,
    on the pure TM path, the incidence medium's ``n`` (``tests/rcwa/test_field_core.py:47-100``)."""
    c = LEGACY[name]
    coupled = value.shape[-1] == 6
    par = PARITY[6] if coupled else PARITY['te' if c['pol'] == 's' else 'tm']
    out = np.flip(value, axis=1 if coupled else 0) * par
    phi = c['phi'] or 0.0
    k_inc = (2 * np.pi / c['wl'] * np.sqrt(c['top']) * np.sin(c['theta'])
             * np.array([np.cos(phi), np.sin(phi)]))
    out = out * np.exp(-1j * k_inc @ (np.asarray(c['period']) / 2))
    if not coupled and c['pol'] == 'p':
        out = out * np.sqrt(c['top'])
    return out


@pytest.mark.parametrize('name', list(LEGACY))
def test_fields_match_the_references_frozen_cells(name):
    g = np.load(GOLDEN / 'legacy_field.npz')
    state = legacy_state(name)
    coupled = state['kind'] == '2d'
    # the reference's route (0 pure 1D, 1 conical, 2 crossed); the conical one takes the 2D path here
    assert (int(g[f'{name}__grating_type']) == 0) == (not coupled)
    if coupled:
        want = legacy_expected(name, g[f'{name}__all'])
        for v in range(3):
            got = orf.sample(state, *RES, variant=v)
            assert got.shape == want[v].shape
            assert np.abs(got - want[v]).max() < 1e-12, (name, v, np.abs(got - want[v]).max())
        # the default request is the asked-for incidence, the 'all' stack's first slab
        assert np.array_equal(g[f'{name}__default'][0], g[f'{name}__all'][0])
    else:
        want = legacy_expected(name, g[f'{name}__default'])
        cols = [1, 3, 5] if state['kind'] == 'te' else [4, 0, 2]
        got = orf.sample(state, RES[0], 1, RES[2])[..., cols]
        assert got.shape == want.shape
        assert np.abs(got - want).max() < 1e-12, (name, np.abs(got - want).max())


# ---------------------------------------------------------------------------------- fmmax
# tests/references/_fmmax_field_cases.py (reference): scene constants.
WL, PERIOD, THICK = 620.0, 400.0, 260.0
NX, NY, RES_X, RES_Z = 400, 8, 9, 7
WL_2D, PERIOD_2D, THICK_2D, THETA_2D, NRAS_2D = 630.0, 500.0, 50.0, 41.0, 100
FMMAX_TOL = {'uniform': 1e-13, 'grating': 1e-4, 'uniform2d': 1e-13, 'patch': 5e-2}
FMMAX_1D = [('uniform', 0.0, 's'), ('uniform', 0.0, 'p'), ('uniform', 24.0, 's'), ('uniform', 24.0, 'p'),
            ('grating', 0.0, 's'), ('grating', 0.0, 'p'), ('grating', 24.0, 's'), ('grating', 24.0, 'p'),
            ('grating', 41.0, 's'), ('grating', 41.0, 'p')]
FMMAX_2D = [('uniform2d', 3, 's'), ('uniform2d', 3, 'p'), ('patch', 3, 's'), ('patch', 3, 'p')]


def _raster_1d(name):
    # _fmmax_field_cases.py:56-76: eps [y, x], a 0.4-fill bar of 4.0 in 1.0, or uniform 2.25
    if name == 'uniform':
        return np.full((NY, NX), 2.25)
    row = np.ones(NX)
    lo = int((0.5 - 0.4 / 2) * NX)
    row[lo:lo + int(0.4 * NX)] = 4.0
    return np.tile(row[None, :], (NY, 1))


def _raster_2d(name):
    # _fmmax_field_cases.py:128-139: a 300 x 120 patch of 4.0 centred at (-70, 60) in vacuum
    if name == 'uniform2d':
        return np.full((NRAS_2D, NRAS_2D), 2.25, dtype=complex)
    ax = (np.arange(NRAS_2D) + 0.5) / NRAS_2D * PERIOD_2D - PERIOD_2D / 2
    X, Y = np.meshgrid(ax, ax, indexing='xy')
    cell = np.ones((NRAS_2D, NRAS_2D), dtype=complex)
    cell[(np.abs(X + 70.0) <= 150.0) & (np.abs(Y - 60.0) <= 60.0)] = 4.0
    return cell


def _expected_scale(theta, pol, wavelength, period):
    # test_fmmax_field_golden.py:95-115: -cos(theta) for s, -1 for p, times the centre phase
    ratio = -np.cos(np.deg2rad(theta)) if pol == 's' else -1.0
    return ratio * np.exp(-1j * 2 * np.pi / wavelength * np.sin(np.deg2rad(theta)) * period / 2)


def _fit(ref, got):
    scale = np.vdot(ref.ravel(), got.ravel()) / np.vdot(ref.ravel(), ref.ravel())
    return scale, np.max(np.abs(got - scale * ref)) / np.max(np.abs(ref))


def _fmmax_frame(cell):
    """fmmax measures z down from the layer's top face: reverse z, mirror parity."""
    return cell[::-1] * PARITY[6]


@pytest.mark.parametrize('name,theta,pol', FMMAX_1D)
def test_fields_match_fmmax_on_the_x_grating(name, theta, pol):
    g = np.load(GOLDEN / 'fmmax_field.npz')
    ref = g[f'{name}__{theta:g}__{pol}']
    psi = np.pi / 2 if pol == 's' else 0.0
    state = orf.solve_state(_raster_1d(name)[None], [THICK], 1.0, 1.0, (PERIOD, PERIOD), WL,
                            np.deg2rad(theta), 0.0, (np.sin(psi), np.cos(psi)), (2, 2))
    # 2*RES_X+1 inclusive samples put every odd one on fmmax's cell-centred grid
    got = _fmmax_frame(orf.sample(state, 2 * RES_X + 1, 1, RES_Z)[:, 0, 1::2, :])
    assert got.shape == ref.shape
    scale, resid = _fit(ref, got)
    assert resid < FMMAX_TOL[name], (name, theta, pol, resid)
    assert abs(scale - _expected_scale(theta, pol, WL, PERIOD)) < FMMAX_TOL[name]
    if name == 'uniform':
        top = np.max(np.abs(ref))
        for i in range(6):
            if np.max(np.abs(ref[..., i])) >= 1e-9 * top:
                gap = np.max(np.abs(got[..., i] - scale * ref[..., i])) / np.max(np.abs(ref[..., i]))
                assert gap < 1e-12, (i, gap)


@pytest.mark.parametrize('name,order,pol', FMMAX_2D)
def test_fields_match_fmmax_on_the_crossed_patch(name, order, pol):
    g = np.load(GOLDEN / 'fmmax_field.npz')
    ref = g[f'2d__{name}__{order}__{pol}']
    psi = np.pi / 2 if pol == 's' else 0.0
    state = orf.solve_state(_raster_2d(name)[None], [THICK_2D], 1.0, 1.0, (PERIOD_2D, PERIOD_2D),
                            WL_2D, np.deg2rad(THETA_2D), 0.0, (np.sin(psi), np.cos(psi)), (order, order))
    xs = g['2d_x'] - PERIOD_2D / 2
    ys = g['2d_y'] - PERIOD_2D / 2
    got = _fmmax_frame(orf.sample(state, xs, ys, 5))
    assert got.shape == ref.shape
    scale, resid = _fit(ref, got)
    assert resid < FMMAX_TOL[name], (name, order, pol, resid)
    assert abs(scale - _expected_scale(THETA_2D, pol, WL_2D, PERIOD_2D)) < FMMAX_TOL[name]
    cell = np.max(np.abs(ref))
    for i in range(6):
        assert np.max(np.abs(got[..., i] - scale * ref[..., i])) / cell < FMMAX_TOL[name]


def test_reversed_raster_rows_move_the_patch_field():
    """Row 0 is the -y end: the rows reversed fit fmmax more than 5x worse (the reference's own
    A/B, test_fmmax_field_golden.py:294-320)."""
    g = np.load(GOLDEN / 'fmmax_field.npz')
    ref = g['2d__patch__3__s']
    xs, ys = g['2d_x'] - PERIOD_2D / 2, g['2d_y'] - PERIOD_2D / 2
    fits = []
    for raster in (_raster_2d('patch'), _raster_2d('patch')[::-1]):
        state = orf.solve_state(raster[None], [THICK_2D], 1.0, 1.0, (PERIOD_2D, PERIOD_2D), WL_2D,
                                np.deg2rad(THETA_2D), 0.0, (1.0, 0.0), (3, 3))
        fits.append(_fit(ref, _fmmax_frame(orf.sample(state, xs, ys, 5)))[1])
    assert fits[1] > 5 * fits[0], fits


# -------------------------------------------------------------------------------- physics

def _stack():
    rng = np.random.default_rng(11)
    eps = np.where(rng.random((2, 6, 7)) < 0.5, 6.25, 1.69).astype(np.complex128)
    return eps, [0.21, 0.34]


def _poynting(state, z, variant=0):
    """Cell-averaged z flux at heights z: per-order orthogonality, 1/2 Re(E x H*)_z summed."""
    out = []
    for zz in np.atleast_1d(z):
        zz_, owner, edges = orf.heights([lay['d'] for lay in state['layers']], [zz])
        l = owner[0]
        c = orf._coefficients(state, state['layers'][l], np.array([zz_[0] - edges[l]]), variant)[0]
        out.append(0.5 * np.sum(np.real(c[:, 0] * np.conj(c[:, 4]) - c[:, 1] * np.conj(c[:, 3]))))
    return np.array(out)


@pytest.mark.parametrize('pol', ['s', 'p', (0.6, 0.8j)])
def test_flux_through_a_lossless_stack_is_the_transmission(pol):
    eps, d = _stack()
    args = (eps, d, 2.25, 1.44, (0.9, 0.8), 0.7, 0.35, 0.6, pol, (3, 2))
    state = orf.solve_state(*args)
    sol = orc.solve_raster(*args)['incident']
    n_top, theta = np.sqrt(2.25), 0.35
    incident = 0.5 * n_top * np.cos(theta)
    z = np.concatenate([np.linspace(0, d[0], 4), d[0] + np.linspace(0, d[1], 4)])
    flux = _poynting(state, z)
    assert np.abs(flux + sol['T'].sum() * incident).max() < 1e-10 * incident, (flux, sol['T'].sum())


def test_tangential_fields_are_continuous_across_faces_and_faraday_holds():
    eps, d = _stack()
    state = orf.solve_state(eps, d, 1.0, 2.25, (0.9, 0.8), 0.8, 0.2, 0.3, (0.3, 0.7), (3, 2))
    cell = orf.sample(state, 6, 5, 3)         # each layer: bottom, middle, top
    lower_top, upper_bottom = cell[2], cell[3]
    scale = np.abs(cell).max()
    for i in (0, 1, 3, 4, 5):                  # Ex, Ey, Hx, Hy, Hz continuous (mu = 1)
        assert np.abs(lower_top[..., i] - upper_bottom[..., i]).max() < 1e-11 * scale, i
    assert np.abs(lower_top[..., 2] - upper_bottom[..., 2]).max() > 1e-3 * scale   # E_z jumps
    # Faraday, x row per order: i k0 ky E_z - dE_y/dz = i k0 H_x (H in units of E)
    layer = state['layers'][1]
    k0, h, z0 = state['k0'], 1e-4 * layer['d'], 0.5 * layer['d']
    zs = z0 + h * np.array([-2, -1, 1, 2])
    c = orf._coefficients(state, layer, np.concatenate([[z0], zs]), 0)
    dEy = (c[1, :, 1] - 8 * c[2, :, 1] + 8 * c[3, :, 1] - c[4, :, 1]) / (12 * h)
    lhs = 1j * k0 * state['ky'] * c[0, :, 2] - dEy
    assert np.abs(lhs - 1j * k0 * c[0, :, 3]).max() < 1e-7 * np.abs(k0 * c[0, :, 3]).max()


def test_explicit_heights_take_the_upper_layer_and_refuse_outside():
    zz, owner, edges = orf.heights([0.2, 0.3], [0.0, 0.2, 0.2 + 1e-17, 0.5])
    assert list(owner) == [0, 1, 1, 1] and zz[2] == 0.2
    with pytest.raises(ValueError):
        orf.heights([0.2, 0.3], [0.51])
