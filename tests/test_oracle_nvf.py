"""Pin the CPU oracle's NVF factorization (oracle/rcwa_nvf.py) by data and mathematics.

The reference ships no NVF golden values for rasters, so the pins are what the construction's
definition and the physics fix: the distance transform against brute force, covariance of the
normal field under lattice translation and transposition, NVF reducing to Li's 1D rules on a
one-row raster (where the RETICOLO goldens then hold through it), a uniform layer, the projector
identity, energy conservation, mirror symmetry, lattice-translation invariance of the
efficiencies, and the closing with a constant 45-degree normal against the same lamellar
grating rotated onto the x axis and solved on the independently pinned axis-aligned path.
"""
import numpy as np
import pytest

from oracle import rcwa_nvf as nvf
from oracle import rcwa_oracle as orc

from test_oracle import GOLDEN, RETICOLO


def _brute_edt(mask, sampling=None):
    sv, su = (1.0, 1.0) if sampling is None else sampling
    ny, nx = mask.shape
    js, is_ = np.nonzero(mask)
    out = np.empty(mask.shape)
    for j in range(ny):
        for i in range(nx):
            best = np.inf
            for tj in (-ny, 0, ny):
                for ti in (-nx, 0, nx):
                    dj = (j - js - tj) * sv
                    di = (i - is_ - ti) * su
                    best = min(best, float(np.min(dj * dj + di * di)))
            out[j, i] = np.sqrt(best)
    return out


def _blobs(shape, seed, values=(1.0, 6.25)):
    rng = np.random.default_rng(seed)
    ny, nx = shape
    y, x = np.mgrid[:ny, :nx]
    eps = np.full(shape, values[0], dtype=np.complex128)
    for _ in range(3):
        cy, cx = rng.uniform(0, ny), rng.uniform(0, nx)
        r = rng.uniform(0.15, 0.3) * min(shape)
        dy = (y - cy + ny / 2) % ny - ny / 2
        dx = (x - cx + nx / 2) % nx - nx / 2
        eps[dx ** 2 + dy ** 2 <= r ** 2] = values[1]
    return eps


@pytest.mark.parametrize('sampling', [None, (1.7, 1.0), (0.6, 1.0)])
@pytest.mark.parametrize('seed', [0, 1])
def test_periodic_distance_is_the_brute_force_nearest_image(seed, sampling):
    mask = nvf.interface_mask(_blobs((11, 14), seed))
    assert mask.any() and not mask.all()
    got = nvf.periodic_edt(mask, sampling)
    want = _brute_edt(mask, sampling)
    assert np.abs(got - want).max() < 1e-12


def test_interface_mask_wraps_and_marks_both_sides():
    eps = np.ones((6, 8))
    eps[:, 0] = 4.0                      # a stripe on the wall: its neighbour across the wall is column 7
    m = nvf.interface_mask(eps)
    assert m[:, [0, 1, 7]].all() and not m[:, 2:7].any()


def test_normal_field_moves_with_a_lattice_translation():
    eps = _blobs((18, 22), 3)
    nx0, ny0 = nvf.normal_vector_field(eps, nvf.lattice_of((1.0, 0.8)))
    shifted = np.roll(eps, (5, -7), axis=(0, 1))
    nx1, ny1 = nvf.normal_vector_field(shifted, nvf.lattice_of((1.0, 0.8)))
    assert np.abs(np.roll(nx0, (5, -7), axis=(0, 1)) - nx1).max() < 1e-14
    assert np.abs(np.roll(ny0, (5, -7), axis=(0, 1)) - ny1).max() < 1e-14


def test_normal_field_transposes_with_the_raster():
    eps = _blobs((16, 16), 4)
    nx0, ny0 = nvf.normal_vector_field(eps, np.eye(2))
    nx1, ny1 = nvf.normal_vector_field(eps.T.copy(), np.eye(2))
    assert np.abs(nx0.T - ny1).max() < 1e-15 and np.abs(ny0.T - nx1).max() < 1e-15


def test_stripes_have_an_axis_normal_and_unit_length_off_flat_pixels():
    eps = np.ones((10, 24))
    eps[:, 3:11] = 9.0
    nx_, ny_ = nvf.normal_vector_field(eps, np.eye(2))
    assert np.all(ny_ == 0.0)
    length = np.hypot(nx_, ny_)
    assert np.all((length == 0.0) | (np.abs(length - 1.0) < 1e-15))
    # flat pixels sit midway between the interfaces only (columns 6-7 inside, 16-17 outside)
    flat = np.flatnonzero(length[0] == 0.0)
    assert set(flat) <= {6, 7, 16, 17, 18}


def test_disc_normal_is_radial_on_its_axes():
    n = 33
    y, x = np.mgrid[:n, :n]
    eps = np.where((x - 16) ** 2 + (y - 16) ** 2 <= 10 ** 2, 6.25, 1.0).astype(np.complex128)
    nx_, ny_ = nvf.normal_vector_field(eps, np.eye(2))
    for (j, i), want in {(16, 26): (1, 0), (16, 6): (-1, 0), (26, 16): (0, 1), (6, 16): (0, -1)}.items():
        got = np.array([nx_[j, i], ny_[j, i]])
        assert abs(abs(got @ np.array(want, float)) - 1.0) < 1e-12


def test_skewed_cell_metric_reduces_to_the_index_grid_for_square_pixels():
    sampling, JT = nvf.cell_metric(np.diag([2.0, 2.0]), 8, 8)
    assert sampling is None and np.array_equal(JT, np.eye(2))
    sampling, JT = nvf.cell_metric(np.diag([1.0, 2.0]), 10, 10)
    assert sampling == (2.0, 1.0) and abs(JT[1, 1] - 0.5) < 1e-16


def test_floor_lift_counts():
    assert nvf.normal_field_floor((12, 12), 3, 3) == (2, 2)
    assert nvf.normal_field_floor((13, 40), 3, 3) == (1, 1)
    assert nvf.normal_field_floor((1, 5), 3, 0) == (1, 3)


def test_point_series_of_samples_is_the_phase_shifted_dft():
    rng = np.random.default_rng(5)
    a = rng.normal(size=7) + 1j * rng.normal(size=7)
    half = 4                                          # orders up to 8 > n: the DFT wraps
    got = nvf.point_series(a, half)
    for idx, k in enumerate(range(-2 * half, 2 * half + 1)):
        j = np.arange(7)
        want = np.sum(a * np.exp(-2j * np.pi * k * (j + 0.5) / 7)) / 7
        assert abs(got[idx] - want) < 1e-14


def test_projector_identity_off_flat_pixels():
    eps = _blobs((24, 24), 6)
    nx_, ny_ = nvf.normal_vector_field(eps, np.eye(2))
    flat = np.hypot(nx_, ny_) < 0.5
    Cxx, Cxy, Cyy = nvf.projector_matrices(nx_, ny_, 3, 3)
    I_flat = orc.conv_matrix((~flat).astype(np.complex128), 3, 3, nvf.point_series, nvf.point_series)
    assert np.abs(Cxx + Cyy - I_flat).max() < 1e-14
    assert np.abs(Cxy - Cxy.conj().T).max() < 1e-15      # a real field: Hermitian projector


@pytest.mark.parametrize('hu', [3, 10])
def test_one_row_raster_is_li_in_1d(hu):
    e1 = np.array([[1, 1, 12.25, 12.25, 3 + 0.1j]], dtype=np.complex128)
    got = nvf.nvf_conv_matrices(e1, hu, 0, lattice=nvf.lattice_of((0.7, 1.0)))
    want = orc.li_conv_matrices(e1, hu, 0)
    for g, w in zip(got[:3], want):
        assert np.abs(g - w).max() < 1e-13 * max(1.0, np.abs(w).max())
    assert not np.any(got[3]) and got[4] is got[3]


@pytest.mark.parametrize('name', ['te_oblique', 'tm_oblique', 'te_quarter', 'tm_two_layer'])
def test_reticolo_1d_goldens_hold_under_nvf(name):
    c = RETICOLO[name]
    g = np.load(GOLDEN / 'reticolo_1d.npz')
    res = orc.solve_raster(c['eps'], c['d'], c['top'], c['bot'], (c['period'], 1.0), c['wl'],
                           c['theta'], None, c['tetm'], (c['M'], 0), factorization='nvf')
    for side, arr in (('ri', res['incident']['R']), ('ti', res['incident']['T'])):
        idx = g[f'{name}__idx_{side}']
        got = np.real(arr).ravel()[idx]
        assert np.abs(got - g[f'{name}__de_{side}']).max() < 1e-11


def test_uniform_layer_is_eps_identity():
    eps = np.full((9, 9), 2.5 + 0.1j)
    epx, epy, epz_i, epxy, _ = nvf.nvf_conv_matrices(eps, 2, 2, lattice=np.eye(2))
    I = np.eye(25)
    for m, want in ((epx, (2.5 + 0.1j) * I), (epy, (2.5 + 0.1j) * I), (epz_i, I / (2.5 + 0.1j))):
        assert np.abs(m - want).max() < 1e-14
    assert np.abs(epxy).max() < 1e-14


@pytest.mark.parametrize('period', [(1.0, 1.0), (1.0, 0.75), [[1.0, 0.4], [0.0, 0.9]]])
def test_lossless_layers_conserve_energy(period):
    eps = np.stack([_blobs((20, 20), 7), _blobs((20, 20), 8, (2.25, 9.0))])
    out = orc.solve_raster(eps, [0.21, 0.13], 1.0, 2.25, np.asarray(period), 0.83, 0.25, 0.6,
                           (0.6, 0.8j), (3, 3), factorization='nvf')
    R, T = out['incident']['R'].sum(), out['incident']['T'].sum()
    assert abs(R + T - 1.0) < 1e-12


def test_lattice_translation_leaves_the_efficiencies():
    eps = _blobs((18, 18), 9)[None]
    args = ([0.3], 1.0, 2.25, (1.0, 1.0), 0.9, 0.2, 0.4, 'p', (3, 3))
    a = orc.solve_raster(eps, *args, factorization='nvf')['incident']
    b = orc.solve_raster(np.roll(eps, (4, -6), axis=(1, 2)), *args, factorization='nvf')['incident']
    for k in ('R', 'T'):
        assert np.abs(a[k] - b[k]).max() < 1e-12


def test_mirror_symmetric_raster_diffracts_symmetrically():
    # a raster symmetric under x -> -x about the cell centre, lit at normal incidence with x
    # polarization: R(n, m) = R(n, -m). A projector off by half a pixel from the permittivity
    # breaks this (factorization.py:322-338).
    n = 20
    y, x = np.mgrid[:n, :n]
    # a triangle: symmetric about the cell's vertical centre line, not under y -> -y
    eps = np.where((np.abs(x + 0.5 - n / 2) <= 0.6 * (y - 3)) & (y <= 16), 9.0, 1.0)
    out = orc.solve_raster(eps[None].astype(np.complex128), [0.25], 1.0, 2.25, (1.0, 1.0), 0.7,
                           0.0, 0.0, 'p', (3, 3), factorization='nvf')['incident']
    for k in ('R', 'T'):
        assert np.abs(out[k] - out[k][:, ::-1]).max() < 1e-12
        assert np.abs(out[k] - out[k][::-1, :]).max() > 1e-6       # not symmetric in y


def _rotated_stripes(M, theta, phi, pol, flip=1.0):
    """Efficiencies of a lamellar grating along x - y (square cell, orders (M, M)) under the NVF
    closing with the constant normal (1, -1)/sqrt(2), and of the same grating rotated onto x
    (period / sqrt(2), orders (M, 0), azimuth phi + pi/4) under the normal (1, 0)."""
    g = np.array([1.0, 1.0, 6.25, 6.25, 6.25, 1.0, 2.25 + 0.0j, 1.0])
    a = orc.raster_series(g, M)                # coefficients of eps along x', orders -2M..2M
    b = orc.raster_series(1.0 / g, M)
    nu = 2 * M + 1
    N = nu * nu
    E2 = np.zeros((N, N), complex)
    A2 = np.zeros((N, N), complex)
    for r in range(N):
        for c in range(N):
            dv, du = r // nu - c // nu, r % nu - c % nu
            if dv == -du:
                E2[r, c], A2[r, c] = a[du + 2 * M], b[du + 2 * M]
    I2 = np.eye(N)
    slots2 = nvf.nvf_assemble(E2, A2, 0.5 * I2, -0.5 * flip * I2, 0.5 * I2)
    T = lambda s: np.array([[s[i - j + 2 * M] for j in range(nu)] for i in range(nu)])
    I1 = np.eye(nu)
    slots1 = nvf.nvf_assemble(T(a), T(b), I1, 0 * I1, 0 * I1)
    L = 1.0
    kw = dict(thickness=[0.37], eps_top=1.0, eps_bot=2.25, theta=theta, pol=pol)
    two = orc.smm_2d(0.83, (L, L), [slots2], kw['thickness'], 1.0, 2.25, theta, phi, pol, M, M)
    one = orc.smm_2d(0.83, (L / np.sqrt(2), 1.0), [slots1], kw['thickness'], 1.0, 2.25, theta,
                     phi + np.pi / 4, pol, M, 0)
    return two, one


@pytest.mark.parametrize('pol', [(1.0, 0.0), (0.0, 1.0), (0.6, 0.8j)])
def test_constant_diagonal_normal_is_the_rotated_lamellar_grating(pol):
    M = 3
    two, one = _rotated_stripes(M, 0.31, 0.2, pol)
    nu = 2 * M + 1
    for key in ('R_s', 'R_p', 'T_s', 'T_p'):
        diag2 = np.array([two['incident'][key][M - m, M + m] for m in range(-M, M + 1)])
        assert np.abs(diag2 - one['incident'][key][0]).max() < 1e-11, key
        off = two['incident'][key].copy()
        for m in range(-M, M + 1):
            off[M - m, M + m] = 0.0
        assert np.abs(off).max() < 1e-12
    # the off-diagonal slot's sign is what tells the two diagonals apart
    wrong, one = _rotated_stripes(M, 0.31, 0.2, pol, flip=-1.0)
    diag2 = np.array([wrong['incident']['R'][M - m, M + m] for m in range(-M, M + 1)])
    assert np.abs(diag2 - one['incident']['R'][0]).max() > 1e-4


def test_floor_is_transparent_where_it_does_not_bind_and_clips_where_it_does():
    rng = np.random.default_rng(11)
    U, _ = np.linalg.qr(rng.normal(size=(6, 6)) + 1j * rng.normal(size=(6, 6)))
    V, _ = np.linalg.qr(rng.normal(size=(6, 6)) + 1j * rng.normal(size=(6, 6)))
    s = np.array([3.0, 2.0, 1.0, 0.5, 0.05, 0.01])
    B = (U * s) @ V.conj().T
    assert nvf.floor_singular_values(B, 0.009) is B
    got = nvf.floor_singular_values(B, 0.1)
    want = (U * np.maximum(s, 0.1)) @ V.conj().T
    assert np.abs(got - want).max() < 1e-14


def test_metal_layer_takes_the_floor_and_passive_layers_do_not():
    eps = _blobs((16, 16), 12, (1.0, -25.0 + 1.5j))
    assert not nvf.passive(eps) and nvf.passive(_blobs((16, 16), 12))
    slots = nvf.nvf_conv_matrices(eps, 3, 3, lattice=np.eye(2))
    assert all(np.all(np.isfinite(m)) for m in slots)
    out = orc.solve_raster(eps[None], [0.05], 1.0, 2.25, (0.6, 0.6), 0.807, 0.1, 0.2, 's', (3, 3),
                           factorization='nvf')['incident']
    assert 0.0 < out['R'].sum() + out['T'].sum() < 1.0


def test_integrated_coefficients_are_the_sampled_ones_on_a_raster():
    eps = _blobs((14, 14), 13, (1.0, 4.0 + 0.2j))
    for got, want in zip(nvf.nvf_conv_matrices(eps, 3, 3, np.eye(2), coefficients='integrated'),
                         nvf.nvf_conv_matrices(eps, 3, 3, np.eye(2))):
        assert np.abs(got - want).max() < 1e-12
