"""CPU oracle for in-layer field reconstruction: the per-layer mode amplitudes a solve leaves
in its cascade, and the six Cartesian field components they give at any point of a layer.

TEST INFRASTRUCTURE ONLY, like ``rcwa_oracle``: only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg may use it, and only as the checker.

Restated from (paths under the reference's ``src/wavebit/``):

* ``rcwa/_engine/scattering.py:221-333`` (``_below_reflections``, ``_field_state``): the
  reflection looking down at every split plane carried up from the substrate, the down-going
  gap wave walked down from the superstrate, and each layer's amplitudes read off the gap waves
  at its faces, ``(A d + B u)/2`` entering and ``(B d + A u)/2`` leaving; the up-going mode is
  read off the top face and carried by ``exp(k0 q d)`` where ``|k0 q d| < 1``
  (``modal_gradient.py:1115-1175`` ``two_face``, threshold ``scattering.py:233``), else off
  the bottom face;
* ``rcwa/_engine/fields.py:503-630`` (``field_1d``) and ``:630-814`` (``_field_coupled``):
  transverse fields ``W (a_up + a_down)`` and ``V (a_down - a_up)``, the longitudinal ones
  from Ampere's and Faraday's z rows, the components in the public frame;
* ``rcwa/_engine/field_result.py:85-114`` (``xy_grid``, ``expand_spectrum``), ``:125-160``
  (face snapping), ``:330-363`` (``_heights``), ``:673-679`` (centre parity of the
  coefficients);
* ``rcwa/results.py:149-227`` (``Result.field``: ``H`` divided by ``Z0``, times the source
  amplitude).

The arithmetic is carried in the frame the cascade runs in: the start port is the scene
bottom, ``+z`` points up, layers are listed bottom to top. In that frame a layer's up-going
mode amplitude is anchored at its bottom face and its down-going one at its top face, and the
public components follow without a mirror (``H`` in units of ``E``, i.e. ``Z0 H``):

    E_t = W (a_up + a_down),  U_t = V (a_down - a_up),  H_t = i U_t,
    E_z = -i epz_i (Kx U_y - Ky U_x),  H_z = Kx E_y - Ky E_x                       (2D)
    TE:  E_y = M,  H_x = i U,  H_z = Kx M;   TM:  H_y = M,  E_x = -i U,  E_z = -epz_i Kx M

with ``M = W (a_up + a_down)``, ``U = V (a_down - a_up)``: Faraday's and Ampere's z rows
(``H_z = Kx E_y - Ky E_x``, ``eps E_z = -(Kx H_y - Ky H_x)``), TM the dual of TE under
``(E, H) -> (H, -E)``, and, in a uniform layer, a down-going wave with ``E = y`` carrying
``H = +x``. The pure TM solve is lit by a unit port
variable (``H_y``); its public field is lit by a unit ``E``, ``n`` of the incidence medium
times more (``tests/rcwa/test_field_core.py:63-86`` of the reference states the factor).

Parity: pinned against the reference's own frozen field cells
(``tests/fixtures/legacy_rcwa/golden_field.npz``) at 1e-12 and against fmmax 1.7.1's in-layer
fields (``tests/references/golden_fmmax_field.npz``: 1e-13 on uniform layers, 1e-4 on the
grating, 5e-2 on the crossed patch) by ``tests/test_oracle_fields.py``, which also checks the
curl equations, the flux through every height of a lossless stack against the efficiencies,
and the continuity of the tangential fields across internal faces.
"""
import numpy as np

from . import rcwa_oracle as orc

Z0 = 1.0 / (8.8541878188e-12 * 299792458.0)       # units.py:7-9
TOP_FACE = 1.0                                      # scattering.py:233
FACE_ULPS_PER_LAYER = 4                             # field_result.py:122


# --------------------------------------------------------------------------- amplitudes

def layer_amplitudes(layers, S_bottom, S_top, c_inc, k0):
    """Each layer's ``(c_down, c_up)`` for the incident columns ``c_inc`` (top incidence).

    ``layers`` bottom to top, each a dict holding its gap-embedded ``A``, ``B`` and symmetric
    S-matrix blocks ``s11``, ``s12``, its roots ``q`` and thickness ``d``; ``S_bottom`` the
    substrate interface (port 2 the gap), ``S_top`` the superstrate one (port 1 the gap).
    ``c_down`` is anchored at the layer's top face, ``c_up`` at its bottom face. Returns the
    reflected columns as well, which equal the solve's."""
    n = c_inc.shape[0]
    I = np.eye(n, dtype=np.complex128)
    gam = [S_bottom[3]]                      # reflection looking down at plane 0
    for lay in layers:
        G = gam[-1]
        gam.append(lay['s11'] + lay['s12'] @ (G @ np.linalg.solve(I - lay['s11'] @ G, lay['s12'])))
    L = len(layers)
    d = [None] * (L + 1)
    d[L] = np.linalg.solve(I - S_top[0] @ gam[L], S_top[1] @ c_inc)
    for l in range(L - 1, -1, -1):
        d[l] = np.linalg.solve(I - layers[l]['s11'] @ gam[l], layers[l]['s12'] @ d[l + 1])
    for l, lay in enumerate(layers):
        A, B = lay['A'], lay['B']
        u_top, u_bot = gam[l + 1] @ d[l + 1], gam[l] @ d[l]
        lay['c_down'] = 0.5 * (A @ d[l + 1] + B @ u_top)
        up_bottom = 0.5 * (B @ d[l] + A @ u_bot)
        up_top = 0.5 * (B @ d[l + 1] + A @ u_top)
        kqd = k0 * lay['d'] * lay['q']
        near = np.abs(kqd) < TOP_FACE
        lay['c_up'] = np.where(near[:, None], np.exp(np.where(near, kqd, 0))[:, None] * up_top, up_bottom)
    reflected = S_top[3] @ c_inc + S_top[2] @ (gam[L] @ d[L])
    return reflected


def state_2d(wavelength, period, conv, thickness, eps_top, eps_bot, theta, phi, pol, hu, hv,
             k_direct=None):
    """Coupled 2D field state, top incidence; the arithmetic of ``rcwa_oracle.smm_2d`` with the
    layers' modes, S-matrices and amplitudes kept. ``pol`` the unit pair ``(a_s, a_p)``; the
    three incident columns are the asked-for one, pure s and pure p."""
    kx, ky = orc.order_kxy_2d(wavelength, period, eps_top, theta, phi, hu, hv, k_direct)
    nu, nv = 2 * hu + 1, 2 * hv + 1
    N = nu * nv
    if kx.ndim == 2:
        kx_g, ky_g = kx, ky
    else:
        kx_g = np.broadcast_to(kx[None, :], (nv, nu))
        ky_g = np.broadcast_to(ky[:, None], (nv, nu))
    kxf, kyf = kx_g.reshape(-1).copy(), ky_g.reshape(-1).copy()
    k0 = 2 * orc._PI / wavelength
    varphi = orc.order_azimuth(ky_g, kx_g, phi)
    Fc, Fs = np.cos(varphi), np.sin(varphi)
    I = np.eye(N, dtype=np.complex128)
    O = np.zeros((N, N), dtype=np.complex128)
    k_par_sq = kxf ** 2 + kyf ** 2
    e_kt, e_kb = np.complex128(eps_bot), np.complex128(eps_top)
    kz_kt = orc.boundary_kz(e_kt, k_par_sq)
    kz_kb = orc.boundary_kz(e_kb, k_par_sq)
    y_s, z_p = orc.gap_waves(kxf, kyf)
    W_kt = np.block([[I, O], [O, -1j * np.diag(kz_kt) / e_kt]])
    V_kt = np.block([[-1j * np.diag(kz_kt), O], [O, I]])
    W_kb = np.block([[I, O], [O, -1j * np.diag(kz_kb) / e_kb]])
    V_kb = np.block([[-1j * np.diag(kz_kb), O], [O, I]])
    W_gap = np.block([[I, O], [O, -1j * np.diag(z_p)]])
    V_gap = np.block([[-1j * np.diag(y_s), O], [O, I]])
    S_bottom = orc.s_reflection(W_kt, V_kt, W_gap, V_gap)
    S_top = orc.s_transmission(W_kb, V_kb, W_gap, V_gap)
    layers = []
    for li, slots in enumerate(conv):
        epx, epy, epz_i = slots[:3]
        epxy, epyx = slots[3:5] if len(slots) == 5 else (None, None)
        W, V, q, status = orc.layer_modes_2d_screened(kxf, kyf, epx, epy, epz_i, 'rebuild', epxy,
                                                      epyx)
        big_W, big_V = orc.sp_modes_2d(W, V, Fc, Fs, N)
        A, B = orc.ab_matrices(big_W, big_V, W_gap, V_gap)
        s11, s12, _, _ = orc.s_layer(A, B, np.exp(-k0 * thickness[li] * q))
        layers.append(dict(W=W, V=V, q=q, d=float(thickness[li]), epz_i=epz_i, A=A, B=B,
                           s11=s11, s12=s12, status=status))
    n_ref = np.sqrt(e_kb)
    delta = np.zeros(N, dtype=np.complex128)
    delta[N // 2] = 1.0

    def incident(a_s, a_p):
        return np.concatenate([a_s * delta, -1j * n_ref * a_p * delta])

    a_s, a_p = np.complex128(pol[0]), np.complex128(pol[1])
    c_inc = np.stack([incident(a_s, a_p), incident(np.sin(np.complex128(orc._PI / 2)),
                                                   np.cos(np.complex128(orc._PI / 2))),
                      incident(np.sin(np.complex128(0.0)), np.cos(np.complex128(0.0)))], axis=1)
    reflected = layer_amplitudes(layers, S_bottom, S_top, c_inc, k0)
    k_inc = orc.incident_k(eps_top, theta, phi, wavelength, k_direct)
    return dict(kind='2d', k0=k0, kx=kxf, ky=kyf, hu=hu, hv=hv, period=period, layers=layers,
                k_inc=k_inc, reflected=reflected, variants=('incident', 's', 'p'))


def state_1d(tetm, wavelength, period_x, eps_layers, thickness, eps_top, eps_bot, theta, hu,
             factorization='li', coefficients='sampled'):
    """Pure TE (``tetm=0``) / TM (``tetm=1``) field state, top incidence; the arithmetic of
    ``rcwa_oracle.smm_1d`` with the layers kept. The one incident column is the solve's own."""
    k_inc_x, _ = orc.incident_k(eps_top, theta, None, wavelength)
    kx = (k_inc_x + np.arange(-hu, hu + 1) * (wavelength / period_x)).astype(np.complex128)
    N = kx.shape[0]
    I = np.eye(N, dtype=np.complex128)
    k0 = 2 * orc._PI / wavelength
    sigma = orc.order_frame_sign(kx, None)
    e_kt, e_kb = np.complex128(eps_bot), np.complex128(eps_top)
    kz_kt = orc.boundary_kz(e_kt, kx ** 2)
    kz_kb = orc.boundary_kz(e_kb, kx ** 2)
    y_s, z_p = orc.gap_waves(kx)
    if tetm == 0:
        V_kt, V_kb, V_gap = -1j * np.diag(kz_kt), -1j * np.diag(kz_kb), -1j * np.diag(y_s)
    else:
        V_kt, V_kb, V_gap = -1j * np.diag(kz_kt / e_kt), -1j * np.diag(kz_kb / e_kb), -1j * np.diag(z_p)
    S_bottom = orc.s_reflection(I, V_kt, I, V_gap)
    S_top = orc.s_transmission(I, V_kb, I, V_gap)
    layers = []
    rule = orc.conv_rule(factorization, coefficients)
    for li in range(len(thickness)):
        epx, epy, epz_i = rule(eps_layers[li], hu, 0)[:3]
        W, V, q = orc.layer_modes_1d(tetm, kx, epx, epy, epz_i)
        A, B = orc.ab_matrices(W, V, I, V_gap)
        s11, s12, _, _ = orc.s_layer(A, B, np.exp(-k0 * thickness[li] * q))
        layers.append(dict(W=W, V=V, q=q, d=float(thickness[li]), epz_i=epz_i, A=A, B=B,
                           s11=s11, s12=s12))
    delta = np.zeros((N, 1), dtype=np.complex128)
    delta[N // 2, 0] = sigma[N // 2]
    reflected = layer_amplitudes(layers, S_bottom, S_top, delta, k0)
    return dict(kind='te' if tetm == 0 else 'tm', k0=k0, kx=kx, ky=np.zeros_like(kx), hu=hu, hv=0,
                period=(period_x, 1.0), layers=layers, k_inc=(k_inc_x, 0j), reflected=reflected,
                variants=('incident',), lit=1.0 if tetm == 0 else np.sqrt(e_kb))


# --------------------------------------------------------------------------- sampling

def is_count(spec):
    """An integer (not a bool) is a number of samples; anything else is coordinates."""
    return isinstance(spec, (int, np.integer)) and not isinstance(spec, bool)


def heights(thickness, z):
    """``(z, owner, edges)``: the heights ``z`` asks for (from the bottom interface), the layer
    owning each and the stack's faces (``field_result.py:330-363``). A count gives each layer
    its own ``linspace(0, 1, count)`` fractions, both faces included; explicit heights within
    ``4 L`` roundings of a face are moved onto it, and a face belongs to the layer above it
    (the top face to the top layer)."""
    ds = np.asarray(thickness, dtype=np.float64)
    edges = np.concatenate([[0.0], np.cumsum(ds)])
    L = ds.shape[0]
    if is_count(z):
        frac = np.linspace(0.0, 1.0, int(z))
        zz = np.concatenate([edges[j] + frac * ds[j] for j in range(L)])
        owner = np.repeat(np.arange(L), int(z))
        return zz, owner, edges
    zz = np.asarray(z, dtype=np.float64).reshape(-1).copy()
    tol = FACE_ULPS_PER_LAYER * max(L, 1) * np.finfo(np.float64).eps * np.max(np.abs(edges))
    if tol > 0 and zz.size:
        gap = edges[None, :] - zz[:, None]
        near = np.argmin(np.abs(gap), axis=1)
        snap = np.abs(gap[np.arange(zz.size), near]) <= tol
        zz = np.where(snap, edges[near], zz)
    if zz.size and (zz.min() < 0 or zz.max() > edges[-1]):
        raise ValueError(f'z must lie inside the stack, 0 to {edges[-1]:g}')
    owner = np.minimum(np.sum(zz[:, None] >= edges[None, 1:], axis=1), L - 1)
    return zz, owner, edges


def axis_points(spec, extent):
    """Centre-origin coordinates along one axis: a count spans ``[-extent/2, extent/2]``."""
    if is_count(spec):
        return np.linspace(-0.5, 0.5, int(spec)) * extent
    return np.asarray(spec, dtype=np.float64).reshape(-1)


def _coefficients(state, layer, zl, variant):
    """Six public components per order at local heights ``zl`` (from the layer's bottom face):
    ``(len(zl), N, 6)``."""
    k0, q, d = state['k0'], layer['q'], layer['d']
    c_up = layer['c_up'][:, variant]
    c_down = layer['c_down'][:, variant]
    a_up = np.exp(-k0 * q[None, :] * zl[:, None]) * c_up[None, :]
    a_down = np.exp(k0 * q[None, :] * (zl[:, None] - d)) * c_down[None, :]
    M = (a_up + a_down) @ layer['W'].T
    U = (a_down - a_up) @ layer['V'].T
    kx, ky = state['kx'], state['ky']
    N = kx.shape[0]
    out = np.zeros((zl.shape[0], N, 6), dtype=np.complex128)
    if state['kind'] == '2d':
        Sx, Sy, Ux, Uy = M[:, :N], M[:, N:], U[:, :N], U[:, N:]
        out[..., 0], out[..., 1] = Sx, Sy
        out[..., 2] = -1j * ((kx * Uy - ky * Ux) @ layer['epz_i'].T)
        out[..., 3], out[..., 4] = 1j * Ux, 1j * Uy
        out[..., 5] = kx * Sy - ky * Sx
    elif state['kind'] == 'te':
        out[..., 1] = M
        out[..., 3] = 1j * U
        out[..., 5] = kx * M
    else:
        out[..., 0] = -1j * U
        out[..., 2] = -((kx * M) @ layer['epz_i'].T)
        out[..., 4] = M
    return out * state.get('lit', 1.0)


def centre_parity(hu, hv):
    """``(-1)**(m+n)`` over the flattened order grid (v slowest)."""
    return orc._centre_parity(hu, hv).reshape(-1)


def sample(state, x, y, z, variant=0):
    """Fields at centre-origin ``x``, ``y`` (counts or 1-D coordinates) and heights ``z``:
    ``(nz, ny, nx, 6)`` = ``(Ex, Ey, Ez, Hx, Hy, Hz)``, ``H`` in units of ``E`` (times ``Z0``),
    the incident wave unit at the cell centre."""
    Px, Py = state['period'][0], state['period'][1]
    xs, ys = axis_points(x, Px), axis_points(y, Py)
    zz, owner, edges = heights([lay['d'] for lay in state['layers']], z)
    N = state['kx'].shape[0]
    coeff = np.zeros((zz.shape[0], N, 6), dtype=np.complex128)
    for l, layer in enumerate(state['layers']):
        sel = np.flatnonzero(owner == l)
        if sel.size:
            zl = np.clip(zz[sel] - edges[l], 0.0, layer['d'])
            coeff[sel] = _coefficients(state, layer, zl, variant)
    coeff = coeff * centre_parity(state['hu'], state['hv'])[None, :, None]
    k0 = state['k0']
    phase = np.exp(1j * k0 * (xs[None, :, None] * state['kx'][None, None, :]
                              + ys[:, None, None] * state['ky'][None, None, :]))   # (ny, nx, N)
    return np.einsum('yxo,zoc->zyxc', phase, coeff)


def public(values, amplitude=1.0):
    """``Result.field``'s units: E in V/m per unit incident amplitude, H in A/m."""
    return values * np.array([1, 1, 1, 1 / Z0, 1 / Z0, 1 / Z0]) * amplitude


def solve_state(eps_layers, thickness, eps_top, eps_bot, period, wavelength, theta, phi,
                polarization, orders, factorization='li', coefficients='sampled'):
    """The field state of one plane-wave point, routed as ``rcwa_oracle.solve_raster`` routes
    it (top incidence)."""
    eps_layers = np.asarray(eps_layers, dtype=np.complex128)
    hu, hv = orders
    if isinstance(polarization, str) and eps_layers.shape[1] == 1 and phi is None and hv == 0:
        return state_1d(0 if polarization == 's' else 1, wavelength, period[0], eps_layers,
                        thickness, eps_top, eps_bot, theta, hu, factorization, coefficients)
    if isinstance(polarization, str):
        pol = (1.0, 0.0) if polarization == 's' else (0.0, 1.0)
    else:
        pol = np.asarray(polarization, dtype=np.complex128)
        pol = pol / np.sqrt(np.sum(np.abs(pol) ** 2))
    lattice = None
    if factorization == 'nvf':
        from oracle.rcwa_nvf import lattice_of
        lattice = lattice_of(period)
    rule = orc.conv_rule(factorization, coefficients, lattice)
    conv = [rule(e, hu, hv) for e in eps_layers]
    return state_2d(wavelength, period, conv, thickness, eps_top, eps_bot, theta, phi, pol, hu, hv)
