"""The oracle's Rayleigh / exceptional-point handling, pinned on the CPU.

Scenes (lengths in nm, layers bottom-to-top, normal incidence, Jones pair (sin pi/2, cos pi/2),
orders (5, 1), period 1000 x 1000, air above a 2.25 substrate, thicknesses 300 and 200):
* the reference's own graze stacks (tests/rcwa/test_rayleigh_guard.py:221-256): a 2-row
  pillar (12.25 in air) over a uniform spacer of index 2.0 or 1.5; at 500 nm the order m = 4
  (resp. 3) sits on the spacer's light line;
* the pillar itself grazing: wavelengths where one of its mode eigenvalues crosses zero (found
  by bisection on the count of negative eigenvalues; at them min|w| / max|w| < 1e-16).

What pins the rebuild: the reference's correct-or-loud contract (the exact graze raises, or
matches the 1e-8-detuned solve within 1e-4 and conserves energy within 1e-4), continuity across
the graze at 1e-11 (the physics is continuous; the unstable product eigenvectors were the
singularity), energy conservation, and the two first-order relations the rebuilt modes satisfy
to round-off where the product construction misses one of them by orders of magnitude.
"""
import math

import numpy as np
import pytest

from oracle import rcwa_oracle as orc

PERIOD = (1000.0, 1000.0)
POL = (math.sin(math.pi / 2), math.cos(math.pi / 2))
PILLAR_GRAZES = (616.5157327418264, 793.192493734249, 538.6449010842978)


def _pillar():
    p = np.ones((1, 2, 20))
    p[0, :, 6:14] = 12.25
    return p


def _stack(n_spacer):
    return np.concatenate([np.full((1, 2, 20), n_spacer ** 2), _pillar()], axis=0)


def _solve(stack, wl, **kw):
    with np.errstate(divide='ignore', invalid='ignore'):
        return orc.solve_raster(stack, [300.0, 200.0], 1.0, 2.25, PERIOD, wl, 0.0, 0.0, POL, (5, 1),
                                **kw)


def _totals(res):
    inc = res['incident']
    return float(inc['R'].sum()), float(inc['T'].sum())


@pytest.mark.parametrize('n_spacer', [2.0, 1.5])
def test_reference_graze_stacks_are_correct_or_loud(n_spacer):
    """test_rayleigh_guard.py:221-237 of the reference, on the oracle."""
    stack = _stack(n_spacer)
    R_far, _ = _totals(_solve(stack, 500.0 * (1 + 1e-8)))
    try:
        res = _solve(stack, 500.0)
    except orc.DefectiveGrazeError:
        return
    R, T = _totals(res)
    assert abs(R - R_far) < 1e-4 and abs(R + T - 1.0) < 1e-4


@pytest.mark.parametrize('n_spacer', [2.0, 1.5])
def test_one_ulp_off_the_reference_graze_is_rebuilt(n_spacer):
    """One ULP of wavelength above the exact point the spacer still grazes (ratio ~1e-16): the
    cluster is representable, rebuilt, equal to the 1e-8-detuned solve within the reference's
    1e-4 and lossless to 1e-9 (its coupled rates are ~1e-8, set by round-off, so they carry an
    O(1) relative error that reaches R + T at the 1e-10 level)."""
    stack = _stack(n_spacer)
    res = _solve(stack, np.nextafter(500.0, 501.0))
    assert res['layer_status'][0] == 'rebuilt'
    R, T = _totals(res)
    R_far, _ = _totals(_solve(stack, 500.0 * (1 + 1e-8)))
    assert abs(R - R_far) < 1e-4
    assert abs(R + T - 1.0) < 1e-9


@pytest.mark.parametrize('wl', PILLAR_GRAZES)
def test_pillar_graze_is_rebuilt_and_continuous(wl):
    """The rebuilt exact graze equals its neighbours 1e-13 either side to 1e-11 and conserves
    energy: the physics is continuous across the condition."""
    stack = _stack(1.5)
    res = _solve(stack, wl)
    assert res['layer_status'][1] == 'rebuilt'
    R, T = _totals(res)
    assert abs(R + T - 1.0) < 1e-11
    for d in (-1e-13, 1e-13):
        Rn, Tn = _totals(_solve(stack, wl * (1 + d)))
        assert abs(R - Rn) < 1e-11 and abs(T - Tn) < 1e-11


@pytest.mark.parametrize('wl', PILLAR_GRAZES + (PILLAR_GRAZES[0] * (1 + 1e-10),))
def test_rebuilt_cluster_satisfies_both_first_order_relations(wl):
    """Omega_R W = V q and Omega_L V = W q on the rebuilt cluster columns to round-off; the
    product construction V = Omega_R W / q satisfies the first by definition and misses the
    second by orders of magnitude more (the reason the reference rebuilds)."""
    hu, hv = 5, 1
    conv = orc.li_conv_matrices(_pillar()[0].astype(complex), hu, hv)
    kx, ky = orc.order_kxy_2d(wl, PERIOD, 1.0, 0.0, 0.0, hu, hv)
    kxf = np.broadcast_to(kx[None, :], (3, 11)).reshape(-1).copy()
    kyf = np.broadcast_to(ky[:, None], (3, 11)).reshape(-1).copy()
    with np.errstate(divide='ignore', invalid='ignore'):
        W, V, q, _, OR = orc.layer_modes_2d(kxf, kyf, *conv)
    OL = orc.omega_left(kxf, kyf, conv[2])
    w = q ** 2
    J = np.flatnonzero(np.abs(w) < orc.GRAZE_REL * np.abs(w).max())
    assert J.size == 2
    Wr, Vr, qr = orc.rebuild_grazing_cluster(w, W, V, q, OL, OR)
    scale = np.abs(OR).max()

    def residuals(W_, V_, q_):
        return (np.abs(OR @ W_[:, J] - V_[:, J] * q_[J]).max() / scale,
                np.abs(OL @ V_[:, J] - W_[:, J] * q_[J]).max() / scale)

    r_right, r_left = residuals(Wr, Vr, qr)
    assert r_right < 1e-13 and r_left < 1e-13
    _, kept_left = residuals(W, V, q)
    assert kept_left > 1e3 * max(r_left, 1e-16)
    # every other column is untouched
    rest = np.setdiff1d(np.arange(W.shape[1]), J)
    assert np.array_equal(Wr[:, rest], W[:, rest]) and np.array_equal(qr[rest], q[rest])


def test_pair_sines_match_the_closed_form_of_a_jordan_perturbation():
    """A = [[l, 1], [delta, l]] has eigenvalues l +- sqrt(delta) and unit eigenvectors
    (1, +-sqrt(delta)) / sqrt(1 + delta), which meet at the sine 2 sqrt(delta) / (1 + delta)."""
    for delta in (1e-4, 1e-8, 1e-12):
        blocks = [np.array([[2.0, 1.0], [delta, 2.0]]), np.diag([5.0, -3.0, 9.0])]
        A = np.zeros((5, 5), dtype=np.complex128)
        A[:2, :2], A[2:, 2:] = blocks
        w, W = np.linalg.eig(A)
        s = orc.pair_sines(w, W)
        want = 2 * math.sqrt(delta) / (1 + delta)
        # eig's eigenvectors of a pair at sine s carry errors of about u / s
        assert s.size == 1 and abs(s[0] - want) <= 1e-6 * want + 1e-9, (delta, s, want)
    w, W = np.linalg.eig(np.diag([1.0, 2.0, 4.0]).astype(complex))
    assert np.allclose(orc.pair_sines(w, W), 1.0)


def test_loud_pair_sine_is_the_stated_bound():
    assert abs(orc.EP_LOUD_SINE ** 2 - 2.0 ** -53 / 1e-10) < 1e-20
    assert 1.0e-3 < orc.EP_LOUD_SINE < 1.1e-3
