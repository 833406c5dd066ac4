"""CPU oracle for the RCWA hot path: a NumPy complex128 restatement of the reference's
raster / sampled-Li / SMM solve, written from the reference's arithmetic, not copied.

TEST INFRASTRUCTURE ONLY. Nothing in the product package imports this module: only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may use it,
and there only as the checker (or the timed CPU baseline), never as the thing measured or
shipped. The HIP engine fails loudly when its extension is missing instead of falling back
here.

Each function cites the reference file:line it restates (paths relative to the reference
repository root, ``src/wavebit/...``). Parity status of every function is recorded in the
module-level ``PINNED`` table and checked by ``tests/test_oracle.py`` against:

* the reference's RETICOLO 1D goldens (``tests/references/golden_reticolo.npz``) -- per-order
  efficiencies, atol 1e-11;
* the reference's RETICOLO 2D Li-nesting totals (``tests/references/test_reticolo_li_nesting.py:34-39``)
  -- 1e-9;
* the reference's legacy per-order goldens (``tests/fixtures/legacy_rcwa/golden_legacy.npz``) --
  efficiencies, amplitudes and ``epz_i`` at 1e-12;
* the reference's fmmax ``Formulation.FFT`` goldens for Laurent's rule
  (``tests/references/golden_fmmax_laurent.npz``, the two rectangular cells), extrapolated over
  the sampling as ``tests/references/test_fmmax_laurent_golden.py:83-89`` does -- per order,
  within the extrapolation's own bound;
* closed forms: exact pixel integrals of a staircase, uniform layer = eps*I, Fresnel/Airy for a
  uniform slab, R+T=1 for lossless stacks, mode residuals.
"""
import math

import numpy as np

_PI = 3.141592653589793
_HALF_PI = 1.5707963267948966
_FLAT = 1e-12                  # conventions.py:70 -- mode_q tie-break threshold
_HALF_PI_F32 = float(np.float32(_HALF_PI))
_HALF_PI_F32_BELOW = float(np.nextafter(np.float32(_HALF_PI), np.float32(0)))

PINNED = {
    'raster_series': 'closed-form pixel integrals (tests/test_oracle.py::test_series_is_exact_pixel_integral)',
    'li_conv_matrices': 'legacy golden epz_i 1e-12; uniform -> eps*I; y-invariant -> 1D Li rule',
    'segment_series': 'closed-form integrals of a staircase; equals raster_series on pixel edges',
    'compression_indices': 'the compressed layer on its edges has the raster\'s own series',
    'li_conv_matrices_integrated': 'legacy golden 2d_cfs (integrated) 1e-12; equals the sampled rule',
    'laurent_conv_matrices_integrated': 'equals the sampled Laurent rule; fmmax goldens through it',
    'laurent_conv_matrices': 'fmmax FFT goldens per order (2 cells); inv(legacy golden epz_i) 1e-12; '
                             'uniform -> eps*I; y-invariant TE equals Li',
    'boundary_kz': 'Fresnel slab closed form; RETICOLO goldens',
    'mode_q': 'uniform-layer closed form q**2 = kx**2+ky**2-eps; mode residuals',
    'order_azimuth': 'legacy golden amplitudes (signed) 1e-12',
    'smm_1d': 'RETICOLO 1D goldens 1e-11; legacy 1D goldens 1e-12',
    'smm_2d': 'legacy 2D goldens 1e-12; RETICOLO 2D Li-nesting totals 1e-9; Airy slab',
}


# --------------------------------------------------------------------------- Fourier series

def raster_series(a, half):
    """Fourier coefficients of a staircase along the last axis, orders -2*half..2*half.

    Restates ``factorization.py:289-311`` (``_pixel_correction``) and ``:398-440``
    (``_raster_series``): the n-point DFT of ``(a - a0)/n`` (1/n applied before the transform,
    ``:369-395`` ``_row_dft``), bin ``k mod n``, times ``sinc(k/n) exp(-i pi k/n)``, plus ``a0``
    at order 0 only.
    """
    a = np.asarray(a, dtype=np.complex128)
    n = a.shape[-1]
    orders = np.arange(-2 * half, 2 * half + 1)
    base = a[..., :1]
    if n == 1:
        spectrum = (a - base)
    else:
        spectrum = np.fft.fft((a - base) / n, axis=-1)
    picked = spectrum[..., orders % n]
    t = orders / n
    with np.errstate(invalid='ignore', divide='ignore'):
        sinc = np.where(t == 0, 1.0, np.sin(_PI * t) / np.where(t == 0, 1.0, _PI * t))
    corr = sinc * np.exp(-1j * _PI * orders / n)
    out = picked * corr
    out = out + np.where(orders == 0, base, np.zeros_like(base))
    return out


def segment_series(a, edges, half):
    """Fourier coefficients along the last axis of a profile constant on segments whose
    right-hand edges are ``edges`` (increasing, the last closing the period), orders
    ``-2*half..2*half``: the closed-form segment integral in its jump form (``factorization.py:
    443-500``, ``_segment_bounds`` / ``_segment_series``):
    ``c_k = sum_j (a_j - a_{j-1}) exp(-2 pi i k left_j) / (2 pi i k)``, ``c_0 = sum_j a_j width_j``.
    """
    a = np.asarray(a, dtype=np.complex128)
    right = np.asarray(edges, dtype=np.float64).reshape(-1)
    right = right / right[-1]
    left = np.concatenate([[0.0], right[:-1]])
    width = right - left
    orders = np.arange(-2 * half, 2 * half + 1)
    with np.errstate(divide='ignore'):
        scale = np.where(orders == 0, 0.0, 1.0 / (2j * _PI * np.where(orders == 0, 1, orders)))
    weights = np.exp(-2j * _PI * (left[:, None] * orders[None, :]))
    jumps = a - np.roll(a, 1, axis=-1)
    ac = jumps @ (weights * scale[None, :])
    dc = (a @ width)[..., None]
    return np.where(orders == 0, dc, ac)


def compression_indices(cell):
    """Where a raster's runs of equal columns, then of equal rows (on the column-reduced raster),
    end, and their right-hand edges on the unit cell (``factorization.py:3005-3047``)."""
    cell = np.asarray(cell)
    n_row, n_col = cell.shape
    col_idx = np.flatnonzero(np.any(cell[:, :-1] != cell[:, 1:], axis=0)).tolist() + [n_col - 1]
    reduced = cell[:, col_idx]
    row_idx = np.flatnonzero(np.any(reduced[:-1, :] != reduced[1:, :], axis=1)).tolist() + [n_row - 1]
    x = (np.asarray(col_idx, dtype=np.float64) + 1.0) / n_col
    y = (np.asarray(row_idx, dtype=np.float64) + 1.0) / n_row
    return row_idx, col_idx, x, y


def toeplitz_positions(half):
    """Gather index of order ``r - c`` in a ``4*half+1`` vector (``factorization.py:276-286,
    358-366``): row-major over an ``(2*half+1)**2`` matrix."""
    size = 2 * half + 1
    r = np.arange(size)
    return (r[:, None] - r[None, :] + 2 * half).reshape(-1)


def conv_matrix(grid, hu, hv, series_u=raster_series, series_v=raster_series):
    """Direct rule both ways, BTTB layout ``M[(n,m),(n',m')] = c(n-n', m-m')``, v first
    (``factorization.py:504-520``). ``series_u`` / ``series_v``: ``f(a, half)`` along the last
    axis (the pixel series by default)."""
    nu, nv = 2 * hu + 1, 2 * hv + 1
    by_column = series_v(grid.T, hv)                  # [j_u, k_v]
    coeffs = series_u(by_column.T, hu)                # [k_v, k_u]
    blocks = coeffs[toeplitz_positions(hv)][:, toeplitz_positions(hu)]
    out = blocks.reshape(nv, nv, nu, nu).transpose(0, 2, 1, 3)
    return out.reshape(nv * nu, nv * nu)


def nested_rule_v(grid, hu, hv, invert_outer, series_u=raster_series, series_v=raster_series):
    """Li's nested operator with ``v`` inside (``factorization.py:523-567``, ``inner='v'``):
    ``T_u([T_v(g)]^-1)``, optionally inverted."""
    nu, nv = 2 * hu + 1, 2 * hv + 1
    lines = series_v(grid.T, hv)                      # [j_u, k_v]
    inner = np.linalg.inv(lines[:, toeplitz_positions(hv)].reshape(lines.shape[0], nv, nv))
    entries = inner.reshape(lines.shape[0], nv * nv).T
    outer = series_u(entries, hu)[:, toeplitz_positions(hu)]
    out = outer.reshape(nv, nv, nu, nu).transpose(0, 2, 1, 3).reshape(nv * nu, nv * nu)
    return np.linalg.inv(out) if invert_outer else out


def li_conv_matrices(eps, hu, hv):
    """``(epx, epy, epz_i)`` for one raster layer, sampled coefficients, fixed nesting
    (``factorization.py:600-642``, ``:658-671``)."""
    eps = np.asarray(eps, dtype=np.complex128)
    epx = nested_rule_v(eps, hu, hv, invert_outer=True)
    epy = nested_rule_v(1.0 / eps, hu, hv, invert_outer=False)
    epz_i = np.linalg.inv(conv_matrix(eps, hu, hv))
    return epx, epy, epz_i


def laurent_conv_matrices(eps, hu, hv):
    """``(epx, epy, epz_i)`` for one raster layer under Laurent's rule, sampled coefficients
    (``factorization.py:944-1000``): the direct rule ``[[eps]]`` -- one matrix -- in both in-plane
    slots and its inverse in the z slot, which is Li's z slot too."""
    direct = conv_matrix(np.asarray(eps, dtype=np.complex128), hu, hv)
    return direct, direct, np.linalg.inv(direct)


def _integrated_parts(eps):
    """The compressed layer and its segment series pair (``factorization.py:651-656, 3135-3156``)."""
    eps = np.asarray(eps, dtype=np.complex128)
    row_idx, col_idx, x, y = compression_indices(eps)
    comp = eps[row_idx][:, col_idx]
    return comp, (lambda a, h: segment_series(a, x, h)), (lambda a, h: segment_series(a, y, h))


def li_conv_matrices_integrated(eps, hu, hv):
    """Li's fixed nesting with integrated (closed-form segment) coefficients on the compressed
    raster (``factorization.py:674-686, 3181-3206``)."""
    comp, su, sv = _integrated_parts(eps)
    epx = nested_rule_v(comp, hu, hv, True, su, sv)
    epy = nested_rule_v(1.0 / comp, hu, hv, False, su, sv)
    epz_i = np.linalg.inv(conv_matrix(comp, hu, hv, su, sv))
    return epx, epy, epz_i


def laurent_conv_matrices_integrated(eps, hu, hv):
    """Laurent's rule with integrated coefficients (``factorization.py:1003-1013, 3209-3232``)."""
    comp, su, sv = _integrated_parts(eps)
    direct = conv_matrix(comp, hu, hv, su, sv)
    return direct, direct, np.linalg.inv(direct)


CONV_RULES = {'li': li_conv_matrices, 'laurent': laurent_conv_matrices,
              ('li', 'integrated'): li_conv_matrices_integrated,
              ('laurent', 'integrated'): laurent_conv_matrices_integrated}


def conv_rule(factorization='li', coefficients='sampled', lattice=None):
    """The oracle's builder for one (factorization, fourier_coefficients) pair. NVF
    (``oracle/rcwa_nvf.py``) also reads the cell basis ``lattice`` (``None``: square pixels)
    and returns five slots ``(epx, epy, epz_i, epxy, epyx)``; the others return three."""
    if factorization == 'nvf':
        from oracle import rcwa_nvf
        return lambda e, hu, hv: rcwa_nvf.nvf_conv_matrices(e, hu, hv, lattice=lattice,
                                                            coefficients=coefficients)
    return CONV_RULES[factorization if coefficients in (None, 'sampled') else (factorization, coefficients)]


# --------------------------------------------------------------------------- conventions

def boundary_kz(eps, k_par_sq):
    """Maystre's cut, ``Re(kz)+Im(kz) > 0`` (``_common/conventions.py:655-778``, non-magnetic)."""
    v = np.sqrt(np.asarray(eps - k_par_sq, dtype=np.complex128))
    return np.where(v.real + v.imag < 0, -v, v)


def mode_q(w):
    """``q = sqrt(w)`` with ``Re(q) >= 0``, ``Im(q) <= 0`` on the flat tie (``conventions.py:186-250``)."""
    v = np.sqrt(np.asarray(w, dtype=np.complex128))
    flip = np.where(np.abs(v.real) > _FLAT, v.real < 0, v.imag > 0)
    return np.where(flip, -v, v)


def order_azimuth(ky, kx, phi):
    """Per-order azimuth ``atan2``-style with the ``pi`` half-turn and the origin rule
    (``conventions.py:1316-1430``). ``kx``, ``ky`` broadcast to the order grid; flat result."""
    kx = np.asarray(kx, dtype=np.complex128)
    ky = np.asarray(ky, dtype=np.complex128)
    kx, ky = np.broadcast_arrays(kx, ky)
    zero_kx = kx == 0
    safe_kx = np.where(zero_kx, 1.0 + 0j, kx)
    varphi = np.arctan(ky / safe_kx)
    varphi = varphi + _PI * (kx.real < 0)
    right = np.where(ky.real < 0, -_HALF_PI, _HALF_PI).astype(np.complex128)
    safe_ky = np.where(ky == 0, 1.0 + 0j, ky)
    right = right - kx / safe_ky
    varphi = np.where(zero_kx, right, varphi)
    p = np.complex128(0.0 if phi is None else phi)
    incident = np.arctan(np.sin(p) / np.cos(p)) + _PI * (np.cos(p).real < 0)
    return np.where(zero_kx & (ky == 0), incident, varphi).reshape(-1)


def order_frame_sign(kx, phi=None):
    """``-1`` where ``Re(kx) < 0`` on the pure-1D row (``conventions.py:1433-1464``)."""
    kx = np.asarray(kx, dtype=np.complex128)
    sign = np.where(kx.real < 0, -1.0, 1.0)
    if phi is None:
        return sign
    turned = np.cos(np.complex128(phi)).real < 0
    return np.where((kx == 0) & turned, -1.0, sign)


def incident_phase_index(eps_mu, theta):
    """``|Re(k)|/k0`` for a real angle (``conventions.py:84-106``)."""
    a, b = eps_mu.real, eps_mu.imag
    loss = b / math.cos(theta)
    radius = math.hypot(a, loss)
    if a >= 0:
        return math.sqrt((radius + a) / 2)
    large = math.sqrt((radius - a) / 2)
    return abs(loss) / (2 * (large if large > 0 else 1.0))


def incident_kxy(eps_inc, theta, phi):
    """``(k_inc_x, k_inc_y)`` normalized by k0, real angles (``engine.py:2706-2795``)."""
    theta = 0.0 if theta is None else float(theta)
    phi = 0.0 if phi is None else float(phi)
    eps_inc = complex(eps_inc)
    if eps_inc.imag == 0 and eps_inc.real <= 0:
        raise ValueError('real-angle incidence requires a propagating incident wave')
    if eps_inc.imag != 0 and abs(theta) >= _PI / 2:
        raise ValueError('real-angle incidence requires a propagating incident wave')
    index = incident_phase_index(eps_inc, theta)
    if theta >= _HALF_PI_F32:
        theta = _HALF_PI_F32_BELOW
    elif theta <= -_HALF_PI_F32:
        theta = -_HALF_PI_F32_BELOW
    s = math.sin(theta)
    return complex(index * s * math.cos(phi)), complex(index * s * math.sin(phi))


# --------------------------------------------------------------------------- S-matrix algebra

def ab_matrices(W_to, V_to, W_from, V_from):
    """``scattering.py:116-125``."""
    P = np.linalg.inv(W_to) @ W_from
    Q = np.linalg.inv(V_to) @ V_from
    return P + Q, P - Q


def gap_waves(kx, ky=None):
    """``sqrt(1 + |k_par|**2)`` (``scattering.py:128-160``)."""
    abs_sq = (kx * np.conj(kx)).real
    if ky is not None:
        abs_sq = abs_sq + (ky * np.conj(ky)).real
    y = np.sqrt(1 + abs_sq).astype(np.complex128)
    return y, y


def s_layer(A, B, X):
    """Gap-embedded layer (``scattering.py:163-174``); X is the diagonal as a vector."""
    A_i = np.linalg.inv(A)
    XB = X[:, None] * B
    common = np.linalg.inv(A - XB @ A_i @ (X[:, None] * B))
    S11 = common @ (XB @ A_i @ (X[:, None] * A) - B)
    S12 = common @ (X[:, None] * (A - B @ A_i @ B))
    return S11, S12, S12, S11


def s_reflection(W_top, V_top, W_gap, V_gap):
    """``scattering.py:177-181``."""
    A, B = ab_matrices(W_gap, V_gap, W_top, V_top)
    A_i = np.linalg.inv(A)
    return -A_i @ B, 2 * A_i, 0.5 * (A - B @ A_i @ B), B @ A_i


def s_transmission(W_bot, V_bot, W_gap, V_gap):
    """``scattering.py:184-188``."""
    A, B = ab_matrices(W_gap, V_gap, W_bot, V_bot)
    A_i = np.linalg.inv(A)
    return B @ A_i, 0.5 * (A - B @ A_i @ B), 2 * A_i, -A_i @ B


def published_scattering(S, sign, flip=(1.0, 1.0, 1.0, 1.0), reverse=True):
    """The stack's scattering matrix as the reference publishes it for top incidence: the kernel's
    blocks (start port = scene bottom) flipped per block (``flip``: the 1D TM reflections,
    ``scattering.py:627-656``), reversed into incidence-relative order (``:607-624``) and signed
    per port amplitude by ``sign`` (frame sign x cell-centre sign, ``engine.py:2995-3009``)."""
    blocks = [f * (sign[:, None] * b * sign[None, :]) for f, b in zip(flip, S)]
    blocks = blocks[::-1] if reverse else blocks     # bottom incidence: the kernel's own order
    return dict(zip(('S11', 'S12', 'S21', 'S22'), blocks))


def redheffer_star(I, S_a, S_b):
    """``scattering.py:191-200``."""
    a11, a12, a21, a22 = S_a
    b11, b12, b21, b22 = S_b
    D_i = np.linalg.inv(I - b11 @ a22)
    F_i = np.linalg.inv(I - a22 @ b11)
    return (a11 + a12 @ D_i @ b11 @ a21, a12 @ D_i @ b12,
            b21 @ F_i @ a21, b22 + b21 @ F_i @ a22 @ b12)


# --------------------------------------------------------------------------- layer modes

def layer_modes_2d(kx, ky, epx, epy, epz_i, epxy=None, epyx=None):
    """``(W, V, q)`` of one 2D layer, non-magnetic, no exceptional-point rebuild
    (``solve_2d.py:2303-2448``). ``epxy`` / ``epyx``: the off-diagonal slots of an NVF layer
    (``None``: zero), entering ``Omega^2`` and ``Omega_R`` as ``solve_2d.py:2393-2411`` writes
    them."""
    N = kx.shape[0]
    I = np.eye(N, dtype=np.complex128)
    Kx, Ky = np.diag(kx), np.diag(ky)
    B = (kx[:, None] * epz_i) * kx[None, :] - I
    D = (ky[:, None] * epz_i) * ky[None, :] - I
    Cc = (kx[:, None] * epz_i) * ky[None, :]
    E = (ky[:, None] * epz_i) * kx[None, :]
    zero = np.zeros((N, N), dtype=np.complex128)
    exy = zero if epxy is None else epxy
    eyx = zero if epyx is None else epyx
    O2 = np.block([[Ky @ Ky + B @ epx - Cc @ eyx, Cc @ epy - Kx @ Ky - B @ exy],
                   [E @ epx - Ky @ Kx - D @ eyx, Kx @ Kx + D @ epy - E @ exy]])
    w, W = np.linalg.eig(O2)
    OR = np.block([[-(Kx @ Ky) + eyx, Kx @ Kx - epy], [epx - Ky @ Ky, Ky @ Kx - exy]])
    q = mode_q(w)
    V = (OR @ W) / q
    return W, V, q, O2, OR


# --------------------------------------------------------------------------- Rayleigh / EP set

GRAZE_REL = 1e-9        # solve_2d.py:44-56 (_COUPLED_REL): a grazing layer, min|w| < 1e-9 max|w|
DEFECTIVE_REL = 1e-10   # solve_2d.py:58-69 (_DEFECTIVE_REL): cluster rates below 1e-10 sqrt(max|w|)
EP_SINE = 1e-2          # solve_2d.py:222-231 (_EP_SINE): the near-defective pair gate
EP_ROOT_FLOOR = 1e-9    # solve_2d.py:233-235 (_EP_ROOT_FLOOR)
# The sine below which the unrebuilt pair may be off by more than 1e-10: the reference measures
# the product construction's error as K u / s**2 with K <= 1e-2 (solve_2d.py:222-229); this
# bound takes K = 1, so u / s**2 > 1e-10 -> s < sqrt(2**-53 / 1e-10).
EP_LOUD_SINE = math.sqrt(2.0 ** -53 / 1e-10)


class DefectiveGrazeError(FloatingPointError):
    """A grazing cluster whose coupled rates split below working precision: no exponential mode
    basis represents the layer (solve_2d.py:190-201). Move off the condition."""


class DefectivePairError(FloatingPointError):
    """A near-defective mode pair whose unrebuilt construction may be off by more than 1e-10."""


def omega_left(kx, ky, epz_i):
    """``Omega_L = [[-Cc, B], [-D, E]]`` (solve_2d.py:2377-2392, 2474-2476), non-magnetic."""
    N = kx.shape[0]
    I = np.eye(N, dtype=np.complex128)
    B = (kx[:, None] * epz_i) * kx[None, :] - I
    D = (ky[:, None] * epz_i) * ky[None, :] - I
    Cc = (kx[:, None] * epz_i) * ky[None, :]
    E = (ky[:, None] * epz_i) * kx[None, :]
    return np.block([[-Cc, B], [-D, E]])


def _small(w, rel):
    a = np.abs(w)
    return np.flatnonzero(a < rel * a.max())


def rebuild_grazing_cluster(w, W, V, q, OL, OR):
    """Re-derive the grazing cluster of a layer from the coupled first-order problem
    (solve_2d.py:118-228, ``_rebuild_grazing_cluster``; non-magnetic SMM root).

    The cluster columns J (|w| below GRAZE_REL of the largest) span S_E; the same cluster of the
    swapped product ``OR @ OL`` spans S_H. With orthonormal bases Q_E, Q_H the factors restricted
    to them, ``R_r = Q_H^H OR Q_E`` and ``L_r = Q_E^H OL Q_H``, give the k x k first-order problem
    ``[[0, L_r], [R_r, 0]] g = mu g``; k of its 2k eigenvectors (the +-mu twins share a field
    pattern) are picked greedily so that both their E parts and their H parts stay independent,
    each branch-fixed to ``mode_q(mu**2)``. Returns ``(W, V, q)`` with the cluster replaced,
    ``None`` when the two products disagree on the cluster size (the caller keeps the
    construction), or raises DefectiveGrazeError when a picked rate is below DEFECTIVE_REL of
    ``sqrt(max|w|)``.
    """
    J = _small(w, GRAZE_REL)
    wh, Vh = np.linalg.eig(OR @ OL)
    Jh = _small(wh, GRAZE_REL)
    if J.size != Jh.size:
        return None
    k = J.size
    Q_E = np.linalg.qr(W[:, J])[0]
    Q_H = np.linalg.qr(Vh[:, Jh])[0]
    R_r = Q_H.conj().T @ OR @ Q_E
    L_r = Q_E.conj().T @ OL @ Q_H
    zero = np.zeros((k, k), dtype=np.complex128)
    mu, g = np.linalg.eig(np.block([[zero, L_r], [R_r, zero]]))
    g = g / np.linalg.norm(g, axis=0)
    # greedy pick: the candidate with the most E and H left after projecting out the picked ones
    rest_e, rest_h = g[:k].copy(), g[k:].copy()
    picked = []
    for _ in range(k):
        score = np.sum(np.abs(rest_e) ** 2, axis=0) + np.sum(np.abs(rest_h) ** 2, axis=0)
        score[picked] = -1.0
        j = int(np.argmax(score))
        picked.append(j)
        for rest in (rest_e, rest_h):
            u = rest[:, j]
            nu = np.linalg.norm(u)
            if nu > 1e-150:
                u = u / nu
                rest -= np.outer(u, u.conj() @ rest)
    mu_p = mu[picked]
    if np.abs(mu_p).min() < DEFECTIVE_REL * np.sqrt(np.abs(w).max()):
        raise DefectiveGrazeError('a grazing mode cluster is numerically defective (its coupled '
                                  'rates split below 1e-10 of the layer scale); move off the '
                                  'condition by a relative step of the wavelength or angle')
    q_p = mode_q(mu_p ** 2)
    sign = np.where(np.abs(q_p - mu_p) > np.abs(q_p + mu_p), -1.0, 1.0)   # took the -q twin
    W, V, q = W.copy(), V.copy(), q.copy()
    W[:, J] = Q_E @ g[:k, picked]
    V[:, J] = Q_H @ (g[k:, picked] * sign[None, :])
    q[J] = q_p
    return W, V, q


def pair_sines(w, W):
    """Sines between the unit eigenvectors of the near-defective pair candidates of a layer:
    modes that are each other's nearest eigenvalue with both roots above EP_ROOT_FLOOR
    (solve_2d.py:583-627, ``_defective_pairs``, the gate without its separation rule). One
    entry per pair; empty when there is none."""
    d = np.abs(w[:, None] - w[None, :])
    np.fill_diagonal(d, np.inf)
    nb = np.argmin(d, axis=1)
    idx = np.arange(w.size)
    roots = np.sqrt(np.abs(w))
    keep = (nb[nb] == idx) & (idx < nb) & (np.minimum(roots, roots[nb]) > EP_ROOT_FLOOR)
    Wn = W / np.linalg.norm(W, axis=0)
    a, b = idx[keep], nb[keep]
    cos = np.abs(np.sum(np.conj(Wn[:, a]) * Wn[:, b], axis=0))
    return np.sqrt(np.maximum(0.0, 1.0 - cos ** 2))


def layer_modes_2d_screened(kx, ky, epx, epy, epz_i, exceptional_points='rebuild', epxy=None,
                            epyx=None):
    """``layer_modes_2d`` with the reference's Rayleigh/EP handling (solve_2d.py:2450-2581):
    a grazing layer's cluster is rebuilt (``rebuild_grazing_cluster``: ``'rebuilt'``, or
    ``'kept'`` when the products disagree on the cluster); otherwise, under
    ``exceptional_points='rebuild'``, a pair whose sine is below EP_LOUD_SINE raises
    DefectivePairError (the reference rebuilds it; this restatement keeps pairs above that sine,
    whose construction is within 1e-10). Returns ``(W, V, q, status)``, status one of
    ``'plain'``, ``'rebuilt'``, ``'kept'``, ``'pair'`` (a gated pair above the loud sine)."""
    with np.errstate(divide='ignore', invalid='ignore'):
        W, V, q, _, OR = layer_modes_2d(kx, ky, epx, epy, epz_i, epxy, epyx)
    w = q ** 2
    if _small(w, GRAZE_REL).size:
        rebuilt = rebuild_grazing_cluster(w, W, V, q, omega_left(kx, ky, epz_i), OR)
        if rebuilt is not None:
            return (*rebuilt, 'rebuilt')
        status = 'kept'
    else:
        status = 'plain'
    if exceptional_points == 'rebuild':
        s = pair_sines(w, W)
        if s.size and s.min() < EP_LOUD_SINE:
            raise DefectivePairError(f'a near-defective mode pair (eigenvector sine {s.min():.2e}) '
                                     'whose unrebuilt construction may be off by more than 1e-10')
        if status == 'plain' and s.size and s.min() < EP_SINE:
            status = 'pair'
    return W, V, q, status


def sp_modes_2d(W, V, Fc, Fs, N):
    """Rotation into the s/p boundary basis (``solve_2d.py:2584-2616``)."""
    c, s = Fc[:, None], Fs[:, None]
    W11, W12, W21, W22 = W[:N, :N], W[:N, N:], W[N:, :N], W[N:, N:]
    V11, V12, V21, V22 = V[:N, :N], V[:N, N:], V[N:, :N], V[N:, N:]
    big_W = np.block([[c * W21 - s * W11, c * W22 - s * W12], [c * W11 + s * W21, c * W12 + s * W22]])
    big_V = np.block([[c * V11 + s * V21, c * V12 + s * V22], [c * V21 - s * V11, c * V22 - s * V12]])
    return big_W, big_V


def layer_modes_1d(tetm, kx, epx, epy, epz_i):
    """TE/TM modes (``solve_1d.py:151-253``, non-magnetic)."""
    N = kx.shape[0]
    I = np.eye(N, dtype=np.complex128)
    if tetm == 0:
        Kx = np.diag(kx)
        A = Kx @ Kx - epy
        w, W = np.linalg.eig(A)
        q = mode_q(w)
        return W, W * q[None, :], q
    Bm = kx[:, None] * epz_i * kx[None, :] - I
    op = epx @ Bm
    w, W = np.linalg.eig(op)
    base = np.linalg.inv(epx) @ W
    q = mode_q(w)
    return W, base * q[None, :], q


# --------------------------------------------------------------------------- solves

def _centre_parity(hu, hv):
    """``(-1)**(m+n)`` sign mask on the order grid (``conventions.py:1232-1263``)."""
    m = np.arange(-hu, hu + 1)
    n = np.arange(-hv, hv + 1)
    return np.where((n[:, None] + m[None, :]) % 2 == 1, -1.0, 1.0)


def smm_1d(tetm, wavelength, period_x, eps_layers, thickness, eps_top, eps_bot, theta, hu,
           phi=None, factorization='li', coefficients='sampled', side='top', k_direct=None):
    """Pure 1D TE/TM solve of a one-row raster stack, top incidence.

    Restates ``scattering.py:704-902`` as the engine calls it (``engine.py:3283-3316``): kernel
    ports swapped (start port = scene bottom), layers bottom-to-top, backside read S22/S12.
    ``eps_layers`` is ``(L, 1, Nx)`` bottom-to-top; returns per-order arrays of shape
    ``(2*hu+1,)``.
    """
    k_inc_x, _ = incident_k(eps_top if side == 'top' else eps_bot, theta, phi, wavelength, k_direct)
    kx = (k_inc_x + np.arange(-hu, hu + 1) * (wavelength / period_x)).astype(np.complex128)
    N = kx.shape[0]
    I = np.eye(N, dtype=np.complex128)
    k0 = 2 * _PI / wavelength
    sigma = order_frame_sign(kx, phi)
    # kernel 'top' is the scene bottom
    n_kt, n_kb = np.sqrt(np.complex128(eps_bot)), np.sqrt(np.complex128(eps_top))
    e_kt, e_kb = np.complex128(eps_bot), np.complex128(eps_top)
    kz_kt = boundary_kz(n_kt ** 2, kx ** 2)
    kz_kb = boundary_kz(n_kb ** 2, kx ** 2)
    y_s, z_p = gap_waves(kx)
    if tetm == 0:
        V_kt, V_kb, V_gap = -1j * np.diag(kz_kt), -1j * np.diag(kz_kb), -1j * np.diag(y_s)
    else:
        V_kt, V_kb, V_gap = -1j * np.diag(kz_kt / e_kt), -1j * np.diag(kz_kb / e_kb), -1j * np.diag(z_p)
    S_ref = s_reflection(I, V_kt, I, V_gap)
    S_trn = s_transmission(I, V_kb, I, V_gap)
    S = S_ref
    for li in range(len(thickness)):
        epx, epy, epz_i = conv_rule(factorization, coefficients)(eps_layers[li], hu, 0)[:3]
        W, V, q = layer_modes_1d(tetm, kx, epx, epy, epz_i)
        X = np.exp(-k0 * thickness[li] * q)
        A, B = ab_matrices(W, V, I, V_gap)
        S = redheffer_star(I, S, s_layer(A, B, X))
    S = redheffer_star(I, S, S_trn)
    top = side == 'top'
    scattering = published_scattering(S, sigma * _centre_parity(hu, 0).reshape(-1),
                                      (-1.0, 1.0, 1.0, -1.0) if tetm == 1 else (1.0, 1.0, 1.0, 1.0), top)
    n_in, n_out = np.sqrt(np.complex128(eps_top)), np.sqrt(np.complex128(eps_bot))
    scattering.update(kx=kx, ky=None, n_ref=n_in if top else n_out, n_trn=n_out if top else n_in,
                      channel='tm' if tetm == 1 else 'te')
    delta = np.zeros(N, dtype=np.complex128)
    delta[N // 2] = 1.0
    delta = delta * sigma[N // 2]
    # top incidence reads the kernel's backside (S22 / S12, ref = kernel bottom = scene top);
    # bottom incidence its front side (S11 / S21, ref = kernel top = scene bottom)
    # (scattering.py:943-949, the backside flag)
    R = (S[3] if top else S[0]) @ delta
    Tt = (S[1] if top else S[2]) @ delta
    kz_ref, kz_trn = (kz_kb, kz_kt) if top else (kz_kt, kz_kb)
    n_ref, n_trn = (n_kb, n_kt) if top else (n_kt, n_kb)
    e_ref, e_trn = (e_kb, e_kt) if top else (e_kt, e_kb)
    adm_ref = kz_ref if tetm == 0 else kz_ref / e_ref
    adm_trn = kz_trn if tetm == 0 else kz_trn / e_trn
    incoming = adm_ref[N // 2].real
    de_ri = (R * np.conj(R)).real * adm_ref.real / incoming
    de_ti = (Tt * np.conj(Tt)).real * adm_trn.real / incoming
    parity = _centre_parity(hu, 0).reshape(-1)
    zeros = np.zeros(N, dtype=np.complex128)
    out = dict(kx=kx)
    if tetm == 0:
        out.update(r_s=R * sigma * parity, r_p=zeros, t_s=Tt * sigma * parity, t_p=zeros,
                   R_s=de_ri, R_p=np.zeros(N), T_s=de_ti, T_p=np.zeros(N))
    else:
        out.update(r_s=zeros, r_p=R * sigma * parity, t_s=zeros,
                   t_p=Tt * (n_ref / n_trn) * sigma * parity,
                   R_s=np.zeros(N), R_p=de_ri, T_s=np.zeros(N), T_p=de_ti)
    out['R'] = out['R_s'] + out['R_p']
    out['T'] = out['T_s'] + out['T_p']
    # power entering the lit face (observations/power.py:37-48 net_power, summed / incoming)
    out['power_entering'] = float(np.sum(((delta + R) * np.conj(adm_ref * (delta - R))).real) / incoming)
    out['scattering'] = scattering
    return out


def incident_k(eps_inc, theta, phi, wavelength, k_direct=None):
    """``incident_kxy``, or direct physical wavenumbers over k0 (``engine.py:2755-2758``)."""
    if k_direct is not None:
        return k_direct[0] * wavelength / (2 * _PI), k_direct[1] * wavelength / (2 * _PI)
    return incident_kxy(eps_inc, theta, phi)


def order_kxy_2d(wavelength, period, eps_top, theta, phi, hu, hv, k_direct=None):
    """Separable ``kx``, ``ky`` (``engine.py:953-993``); on a skewed cell, ``period`` a 2 x 2
    lattice whose columns are a1 and a2, the two (2hv+1, 2hu+1) grids of ``lattice.py:899-944``."""
    k_inc_x, k_inc_y = incident_k(eps_top, theta, phi, wavelength, k_direct)
    A = np.asarray(period, dtype=np.float64)
    if A.shape == (2, 2):
        a1x, a2x, a1y, a2y = A[0, 0], A[0, 1], A[1, 0], A[1, 1]
        lam = wavelength / (a1x * a2y - a2x * a1y)
        m = np.arange(-hu, hu + 1, dtype=np.float64)[None, :]
        n = np.arange(-hv, hv + 1, dtype=np.float64)[:, None]
        kx = k_inc_x + lam * (a2y * m - a1y * n)
        ky = k_inc_y + lam * (-a2x * m + a1x * n)
        return kx.astype(np.complex128), ky.astype(np.complex128)
    kx = (k_inc_x + np.arange(-hu, hu + 1) * (wavelength / period[0])).astype(np.complex128)
    ky = (k_inc_y + np.arange(-hv, hv + 1) * (wavelength / period[1])).astype(np.complex128)
    return kx, ky


def smm_2d(wavelength, period, conv, thickness, eps_top, eps_bot, theta, phi, pol, hu, hv,
           return_modes=False, side='top', k_direct=None, exceptional_points='rebuild'):
    """Coupled s/p 2D solve, top incidence (``scattering.py:1185-1456``, ``:907-994``, as called
    by ``engine.py:3402-3420``). ``conv`` is a list of ``(epx, epy, epz_i)`` bottom-to-top;
    ``pol`` the unit pair ``(a_s, a_p)``. Returns the three incidence variants
    (``incident``, ``s_inc``, ``p_inc``), each with per-order arrays ``(2hv+1, 2hu+1)``."""
    top = side == 'top'
    kx, ky = order_kxy_2d(wavelength, period, eps_top if top else eps_bot, theta, phi, hu, hv, k_direct)
    nu, nv = 2 * hu + 1, 2 * hv + 1
    N = nu * nv
    if kx.ndim == 2:                    # skewed cell: the grids themselves
        kx_g, ky_g = kx, ky
    else:
        kx_g = np.broadcast_to(kx[None, :], (nv, nu))
        ky_g = np.broadcast_to(ky[:, None], (nv, nu))
    kxf, kyf = kx_g.reshape(-1).copy(), ky_g.reshape(-1).copy()
    k0 = 2 * _PI / wavelength
    varphi = order_azimuth(ky_g, kx_g, phi)
    Fc, Fs = np.cos(varphi), np.sin(varphi)
    I = np.eye(N, dtype=np.complex128)
    O = np.zeros((N, N), dtype=np.complex128)
    bI = np.eye(2 * N, dtype=np.complex128)
    k_par_sq = kxf ** 2 + kyf ** 2
    e_kt, e_kb = np.complex128(eps_bot), np.complex128(eps_top)   # kernel ports
    n_kt, n_kb = np.sqrt(e_kt), np.sqrt(e_kb)
    kz_kt = boundary_kz(n_kt ** 2, k_par_sq)
    kz_kb = boundary_kz(n_kb ** 2, k_par_sq)
    y_s, z_p = gap_waves(kxf, kyf)
    W_kt = np.block([[I, O], [O, -1j * np.diag(kz_kt) / e_kt]])
    V_kt = np.block([[-1j * np.diag(kz_kt), O], [O, I]])
    W_kb = np.block([[I, O], [O, -1j * np.diag(kz_kb) / e_kb]])
    V_kb = np.block([[-1j * np.diag(kz_kb), O], [O, I]])
    W_gap = np.block([[I, O], [O, -1j * np.diag(z_p)]])
    V_gap = np.block([[-1j * np.diag(y_s), O], [O, I]])
    S = s_reflection(W_kt, V_kt, W_gap, V_gap)
    S_trn = s_transmission(W_kb, V_kb, W_gap, V_gap)
    modes = []
    status = []
    for li, slots in enumerate(conv):
        epx, epy, epz_i = slots[:3]
        epxy, epyx = slots[3:5] if len(slots) == 5 else (None, None)
        W, V, q, st = layer_modes_2d_screened(kxf, kyf, epx, epy, epz_i, exceptional_points,
                                              epxy, epyx)
        status.append(st)
        big_W, big_V = sp_modes_2d(W, V, Fc, Fs, N)
        X = np.exp(-k0 * thickness[li] * q)
        A, B = ab_matrices(big_W, big_V, W_gap, V_gap)
        S = redheffer_star(bI, S, s_layer(A, B, X))
        if return_modes:
            modes.append((W, V, q))
    S = redheffer_star(bI, S, S_trn)
    par = _centre_parity(hu, hv).reshape(-1)
    scattering = published_scattering(S, np.concatenate([par, par]), reverse=top)
    n_in, n_out = np.sqrt(np.complex128(eps_top)), np.sqrt(np.complex128(eps_bot))
    scattering.update(kx=kx, ky=ky, n_ref=n_in if top else n_out, n_trn=n_out if top else n_in,
                      channel='sp')
    # top incidence: the backside read (reflection S22, transmission S12); bottom: the front
    # side (S11, S21) -- scattering.py:943-949
    kz_ref, kz_trn = (kz_kb, kz_kt) if top else (kz_kt, kz_kb)
    n_ref, n_trn = (n_kb, n_kt) if top else (n_kt, n_kb)
    e_ref, e_trn = (e_kb, e_kt) if top else (e_kt, e_kb)
    delta = np.zeros(N, dtype=np.complex128)
    delta[N // 2] = 1.0

    def incident(a_s, a_p):
        return np.concatenate([a_s * delta, -1j * n_ref * a_p * delta])

    a_s, a_p = np.complex128(pol[0]), np.complex128(pol[1])
    cols = [incident(a_s, a_p), incident(np.sin(np.complex128(_PI / 2)), np.cos(np.complex128(_PI / 2))),
            incident(np.sin(np.complex128(0.0)), np.cos(np.complex128(0.0)))]
    c_inc = np.stack(cols, axis=1)
    c_ref = (S[3] if top else S[0]) @ c_inc
    c_trn = (S[1] if top else S[2]) @ c_inc
    s_power = kz_ref[N // 2].real
    p_power = abs(n_ref) ** 2 * (kz_ref[N // 2] / e_ref).real
    powers = [abs(a_s) ** 2 * s_power + abs(a_p) ** 2 * p_power, s_power, p_power]
    parity = _centre_parity(hu, hv)
    out = {}
    for i, name in enumerate(('incident', 's_inc', 'p_inc')):
        Rs = c_ref[:N, i]
        Rp = -c_ref[N:, i]
        Ts = c_trn[:N, i]
        Tp = c_trn[N:, i]
        den = powers[i]
        eff = dict(R_s=(Rs * np.conj(Rs)).real * kz_ref.real / den,
                   R_p=(Rp * np.conj(Rp)).real * (kz_ref / e_ref).real / den,
                   T_s=(Ts * np.conj(Ts)).real * kz_trn.real / den,
                   T_p=(Tp * np.conj(Tp)).real * (kz_trn / e_trn).real / den)
        flux = (((c_inc[:N, i] + Rs) * np.conj(kz_ref * (c_inc[:N, i] - Rs))).real
                + ((c_inc[N:, i] + Rp) * np.conj((kz_ref / e_ref) * (c_inc[N:, i] - Rp))).real)
        sub = {k: v.reshape(nv, nu) for k, v in eff.items()}
        sub['R'] = sub['R_s'] + sub['R_p']
        sub['T'] = sub['T_s'] + sub['T_p']
        sub['r_s'] = Rs.reshape(nv, nu) * parity
        sub['r_p'] = (Rp / (-1j * n_ref)).reshape(nv, nu) * parity
        sub['t_s'] = Ts.reshape(nv, nu) * parity
        sub['t_p'] = (Tp / (-1j * n_trn)).reshape(nv, nu) * parity
        sub['power_entering'] = flux.sum() / den
        out[name] = sub
    out['kx'], out['ky'] = kx, ky
    out['scattering'] = scattering
    out['layer_status'] = status
    if return_modes:
        out['modes'] = modes
    return out


def solve_raster(eps_layers, thickness, eps_top, eps_bot, period, wavelength, theta, phi,
                 polarization, orders, factorization='li', coefficients='sampled', side='top',
                 k_direct=None, exceptional_points='rebuild'):
    """One plane-wave point through a raster stack, routed the way the reference routes it.

    ``eps_layers``: ``(L, Ny, Nx)`` complex, bottom-to-top (``problem.py:1-5``). ``polarization``
    is ``'s'``/``'p'`` or a Jones pair. Routing (``engine.py:3092-3181``): a one-row raster with
    a polarization word and ``phi is None`` takes the pure 1D path; everything else 2D (the 1D
    conical layout is out of scope). ``period``: ``(Lx, Ly)``, or a 2 x 2 lattice whose columns
    are the lattice vectors (a skewed cell; Laurent's rule). ``factorization``: ``'li'`` (the
    default) or ``'laurent'``; ``coefficients``: ``'sampled'`` (the default) or ``'integrated'``.
    """
    lattice = None
    if factorization == 'nvf':
        from oracle.rcwa_nvf import lattice_of
        lattice = lattice_of(period)
    rule = conv_rule(factorization, coefficients, lattice)
    eps_layers = np.asarray(eps_layers, dtype=np.complex128)
    hu, hv = orders
    one_row = eps_layers.shape[1] == 1
    if isinstance(polarization, str):
        if one_row and phi is None and hv == 0 and (k_direct is None or k_direct[1] == 0):
            tetm = 0 if polarization == 's' else 1
            out = smm_1d(tetm, wavelength, period[0], eps_layers, thickness, eps_top, eps_bot,
                         theta, hu, phi, factorization, coefficients, side, k_direct)
            inc = {k: np.asarray(v).reshape(1, -1) for k, v in out.items()
                   if k not in ('kx', 'power_entering', 'scattering')}
            inc['power_entering'] = out['power_entering']
            return {'path': '1d_te' if tetm == 0 else '1d_tm', 'incident': inc,
                    'scattering': out['scattering']}
        pol = (1.0, 0.0) if polarization == 's' else (0.0, 1.0)
    else:
        pol = np.asarray(polarization, dtype=np.complex128)
        pol = pol / np.sqrt(np.sum(np.abs(pol) ** 2))
    if one_row and hv != 0:
        raise ValueError('a one-row raster has no y orders')
    conv = [rule(e, hu, hv) for e in eps_layers]
    out = smm_2d(wavelength, period, conv, thickness, eps_top, eps_bot, theta, phi, pol, hu, hv,
                 side=side, k_direct=k_direct, exceptional_points=exceptional_points)
    out['path'] = '2d'
    return out
