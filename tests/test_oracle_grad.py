"""Pins the gradient oracle (CPU): its forward equals the oracle's 2D solve, and its reverse
pass equals central differences of the oracle in every input it differentiates."""
import numpy as np
import pytest

from oracle import rcwa_oracle as orc
from oracle import rcwa_oracle_grad as og

H = 1e-6        # central-difference step; absolute floors sit at its rounding noise (eps |f| / H)


def _case(seed=0, hu=2, hv=1, L=2, lossy=True):
    rng = np.random.default_rng(seed)
    eps = rng.uniform(1, 6, (L, 8, 10)) + (1j * rng.uniform(0, 0.3, (L, 8, 10)) if lossy else 0)
    return eps.astype(np.complex128), rng.uniform(0.15, 0.35, L)


@pytest.mark.parametrize('theta,phi,pol', [(0.2, 0.3, (0.6, 0.8j)), (0.0, 0.0, (0.0, 1.0)),
                                           (0.35, 1.9, (1.0, 0.0))])
def test_forward_equals_the_oracle_solve(theta, phi, pol):
    eps, thick = _case()
    conv = [orc.li_conv_matrices(e, 2, 1) for e in eps]
    s = og.setup_point(0.9, (0.77, 0.7), 1.0, 2.25, theta, phi, pol, 2, 1)
    eff, _ = og.forward_2d(s, conv, thick)
    ref = orc.smm_2d(0.9, (0.77, 0.7), conv, thick, 1.0, 2.25, theta, phi, pol, 2, 1)
    for v, name in enumerate(('incident', 's_inc', 'p_inc')):
        for k, key in enumerate(('R_s', 'R_p', 'T_s', 'T_p')):
            assert np.abs(eff[v, k].reshape(3, 5) - ref[name][key]).max() < 1e-13


def test_reverse_pass_equals_central_differences_of_the_oracle():
    eps, thick = _case(1)
    hu, hv = 2, 1
    conv = [list(orc.li_conv_matrices(e, hu, hv)) for e in eps]
    wl, period, theta, phi, pol = 0.85, (0.77, 0.7), 0.25, 0.4, (0.6, 0.8j)
    s = og.setup_point(wl, period, 1.0, 2.25, theta, phi, pol, hu, hv)
    rng = np.random.default_rng(5)
    eb = rng.normal(size=(3, 4, s['N']))
    _, st = og.forward_2d(s, conv, thick)
    grads = og.backward_2d(s, st, eb)

    def objective(cv, t):
        ref = orc.smm_2d(wl, period, cv, t, 1.0, 2.25, theta, phi, pol, hu, hv)
        return sum(float(np.sum(ref[name][key].reshape(-1) * eb[v, k]))
                   for v, name in enumerate(('incident', 's_inc', 'p_inc'))
                   for k, key in enumerate(('R_s', 'R_p', 'T_s', 'T_p')))

    for which in range(3):
        for _ in range(3):
            l = int(rng.integers(0, 2))
            i, j = (int(v) for v in rng.integers(0, s['N'], 2))
            for dz in (1.0, 1j):
                plus = [[m.copy() for m in c] for c in conv]
                minus = [[m.copy() for m in c] for c in conv]
                plus[l][which][i, j] += H * dz
                minus[l][which][i, j] -= H * dz
                fd = (objective(plus, thick) - objective(minus, thick)) / (2 * H)
                g = grads[l][which][i, j]
                an = g.real if dz == 1.0 else g.imag
                assert abs(an - fd) <= 1e-6 * abs(fd) + 1e-8, (which, l, i, j, dz, an, fd)
    for l in range(2):
        tp, tm = thick.copy(), thick.copy()
        tp[l] += H
        tm[l] -= H
        fd = (objective(conv, tp) - objective(conv, tm)) / (2 * H)
        assert abs(grads[l][3] - fd) <= 1e-6 * abs(fd) + 1e-8


def test_assembly_adjoint_equals_central_differences():
    rng = np.random.default_rng(2)
    hu, hv = 3, 2
    eps = (rng.uniform(1, 9, (12, 9)) + 1j * rng.uniform(0, 0.5, (12, 9))).astype(np.complex128)
    N = (2 * hu + 1) * (2 * hv + 1)
    G = [rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N)) for _ in range(3)]
    d = og.li_conv_adjoint(eps, hu, hv, *G)

    def objective(e):
        return float(sum(np.real(np.sum(np.conj(g) * m)) for g, m in zip(G, orc.li_conv_matrices(e, hu, hv))))

    for _ in range(6):
        y, x = int(rng.integers(0, 12)), int(rng.integers(0, 9))
        for dz in (1.0, 1j):
            ep, em = eps.copy(), eps.copy()
            ep[y, x] += H * dz
            em[y, x] -= H * dz
            fd = (objective(ep) - objective(em)) / (2 * H)
            an = d[y, x].real if dz == 1.0 else d[y, x].imag
            assert abs(an - fd) <= 1e-6 * abs(fd) + 1e-8, (y, x, dz, an, fd)


def test_raster_gradient_equals_central_differences_of_solve_raster():
    rng = np.random.default_rng(4)
    L, hu, hv = 2, 3, 3
    eps = rng.uniform(1, 9, (L, 8, 8)).astype(np.complex128)
    thick = np.array([0.2, 0.13])
    wls = [0.777, 0.9]
    N = (2 * hu + 1) * (2 * hv + 1)
    eb = np.zeros((len(wls), 3, 4, N))
    eb[:, 0, 2] = eb[:, 0, 3] = 1.0               # sum of T of the incidence as given
    total, d_eps, d_t = og.raster_gradient(eps, thick, 1.0, 2.25, (0.77, 0.77), wls, 0.2, 0.3,
                                           'p', (hu, hv), eb)

    def objective(e, t):
        return sum(float(orc.solve_raster(e, t, 1.0, 2.25, (0.77, 0.77), wl, 0.2, 0.3, 'p',
                                          (hu, hv))['incident']['T'].sum()) for wl in wls)

    assert abs(total - objective(eps, thick)) < 1e-12

    def check(f_of, an, h3):
        # central difference at the stated step (its noise, ~eps |f| / H, sets the floor) and a
        # fourth-order five-point difference at h3, tight enough for 1e-6 relative (h3 is
        # smaller for thickness: evanescent orders decay at k0 |q| ~ 1e2 per micron, so the
        # third derivative in thickness is ~1e6)
        fd1 = (f_of(H) - f_of(-H)) / (2 * H)
        fd4 = (-f_of(2 * h3) + 8 * f_of(h3) - 8 * f_of(-h3) + f_of(-2 * h3)) / (12 * h3)
        assert abs(an - fd1) <= 1e-6 * abs(fd1) + 1e-8, (an, fd1)
        assert abs(an - fd4) <= 1e-6 * abs(fd4) + 1e-11, (an, fd4)

    for _ in range(4):
        l, y, x = int(rng.integers(0, L)), int(rng.integers(0, 8)), int(rng.integers(0, 8))

        def f_pix(dv, l=l, y=y, x=x):
            e = eps.copy()
            e[l, y, x] += dv
            return objective(e, thick)
        check(f_pix, d_eps[l, y, x].real, 1e-3)
    for l in range(L):
        def f_thk(dv, l=l):
            t = thick.copy()
            t[l] += dv
            return objective(eps, t)
        check(f_thk, d_t[l], 1e-6)
