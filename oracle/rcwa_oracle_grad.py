"""CPU oracle for the RCWA gradient: the derivative of the oracle's 2D raster solve with respect
to the Li convolution matrices, the layer thicknesses and the raster pixels, by the rules the
reference's PyTorch gradient uses -- written as an explicit reverse pass in NumPy complex128.

TEST INFRASTRUCTURE ONLY (like ``rcwa_oracle``): only ``tests/`` may use it, as the checker of
the device adjoint. Nothing in the product imports it.

Reference rules restated (paths relative to the reference's ``src/wavebit/``):

* the eigenbasis W of each layer is held fixed (``solve_2d.py:2427-2429`` runs ``eig`` on a
  stopped Omega^2) and q moves as a matrix Q in that basis, with tangent
  ``(W^-1 dOmega2 W) / (q_i + q_j)`` (``rcwa/_engine/modal_gradient.py:161-251``,
  ``generator``); its cotangent map is ``Omega2bar = W^-H (Qbar / conj(q_i + q_j)) W^H``;
* ``V = Omega_R W Q^-1`` (``modal_gradient.py:1239-1244``);
* ``X = exp(-k0 d Q)`` through divided differences of exp on (q_i, q_j)
  (``modal_gradient.py:446-470``);
* everything else (the Li assembly's FFTs and inverses, the S-matrix algebra, the
  efficiencies) by the ordinary product / inverse rules, as autograd takes them.

Cotangent convention: for a real objective f of a complex z, ``zbar = df/dRe z + i df/dIm z``
(PyTorch's), so ``df = Re(conj(zbar) dz)``.

The forward here is the same arithmetic as ``rcwa_oracle.smm_2d`` arranged as the device
cascade: the reflection looking back to the start port, ``G_{l+1} = s11 + s12 G (I - s11 G)^-1
s12``, and the transmitted columns swept back through the layers -- the S22 / S12 columns the
Redheffer star cascade gives (``scattering.py:191-200``). ``tests/test_oracle_grad.py`` checks
its efficiencies against ``smm_2d`` and every gradient here against central differences of
``rcwa_oracle`` (PINNED below).
"""
import numpy as np

from . import rcwa_oracle as orc

_PI = 3.141592653589793

PINNED = {
    'forward_2d': 'efficiencies equal rcwa_oracle.smm_2d (all three incidence variants) to 1e-13',
    'backward_2d': 'central differences of rcwa_oracle.smm_2d in epx, epy, epz_i (re and im) and '
                   'thickness, 1e-6 relative (tests/test_oracle_grad.py)',
    'li_conv_adjoint': 'central differences of rcwa_oracle.li_conv_matrices, 1e-6 relative',
    'raster_gradient': 'central differences of rcwa_oracle.solve_raster in raster pixels and '
                       'thickness, 1e-6 relative',
}


# --------------------------------------------------------------------------- one point

def setup_point(wavelength, period, eps_top, eps_bot, theta, phi, pol, hu, hv):
    """Per-order quantities of one 2D point (``engine.py:953-993``, ``scattering.py:128-188``,
    ``conventions.py:655-778, 1316-1430``), the half-space port blocks as diagonals and the
    incidence variants of ``scattering.py:943-946``."""
    kx, ky = orc.order_kxy_2d(wavelength, period, eps_top, theta, phi, hu, hv)
    nu, nv = 2 * hu + 1, 2 * hv + 1
    N = nu * nv
    kx_g = np.broadcast_to(kx[None, :], (nv, nu))
    ky_g = np.broadcast_to(ky[:, None], (nv, nu))
    varphi = orc.order_azimuth(ky_g, kx_g, phi)
    kxf, kyf = kx_g.reshape(-1).copy(), ky_g.reshape(-1).copy()
    kpar2 = kxf ** 2 + kyf ** 2
    e_bot, e_top = np.complex128(eps_bot), np.complex128(eps_top)
    kz_bot = orc.boundary_kz(np.sqrt(e_bot) ** 2, kpar2)
    kz_top = orc.boundary_kz(np.sqrt(e_top) ** 2, kpar2)
    y = orc.gap_waves(kxf, kyf)[0]

    def port(kz, e):
        # half-space -> gap interface, diagonal: P = Wg^-1 Wh, Q = Vg^-1 Vh
        Pd = np.concatenate([np.ones(N), kz / (e * y)])
        Qd = np.concatenate([kz / y, np.ones(N)])
        A, B = Pd + Qd, Pd - Qd
        return B / A, 2 / A, 0.5 * (A - B * (B / A))
    BA, twoA, half = port(kz_bot, e_bot)          # start port: the scene bottom
    dg = dict(r12=twoA, r22=BA)
    BA, twoA, half = port(kz_top, e_top)          # end port: the scene top (lit)
    dg.update(t11=BA, t12=half, t21=twoA, t22=-BA)
    n_ref = np.sqrt(e_top)
    mi_n = -1j * n_ref
    if isinstance(pol, str):                      # rcwa_oracle.solve_raster's 2D mapping
        pol = (1.0, 0.0) if pol == 's' else (0.0, 1.0)
    else:
        pol = np.asarray(pol, dtype=np.complex128)
        pol = pol / np.sqrt(np.sum(np.abs(pol) ** 2))
    a_s, a_p = complex(pol[0]), complex(pol[1])
    al = [a_s, np.sin(_PI / 2), 0.0]
    be = [mi_n * a_p, mi_n * np.cos(_PI / 2), mi_n * 1.0]
    s_power = kz_top[N // 2].real
    p_power = abs(n_ref) ** 2 * (kz_top[N // 2] / e_top).real
    den = [abs(a_s) ** 2 * s_power + abs(a_p) ** 2 * p_power, s_power, p_power]
    adm = np.stack([kz_top.real, (kz_top / e_top).real, kz_bot.real, (kz_bot / e_bot).real])
    return dict(kx=kxf, ky=kyf, fc=np.cos(varphi), fs=np.sin(varphi), N=N, n=2 * N, dg=dg,
                Dw=np.concatenate([np.ones(N), -1j * y]), Dv=np.concatenate([-1j * y, np.ones(N)]),
                k0=2 * _PI / wavelength, al=al, be=be, den=den, adm=adm)


def _rot_w(W, s):
    """R W with R = [[-Fs, Fc], [Fc, Fs]] by order rows (``solve_2d.py:2584-2616``)."""
    N = s['N']
    c, sn = s['fc'][:, None], s['fs'][:, None]
    return np.concatenate([c * W[N:] - sn * W[:N], c * W[:N] + sn * W[N:]])


def _rot_v(V, s):
    """R' V with R' = [[Fc, Fs], [-Fs, Fc]]."""
    N = s['N']
    c, sn = s['fc'][:, None], s['fs'][:, None]
    return np.concatenate([c * V[:N] + sn * V[N:], c * V[N:] - sn * V[:N]])


def _rot_v_t(U, s):
    """R'^T U."""
    N = s['N']
    c, sn = s['fc'][:, None], s['fs'][:, None]
    return np.concatenate([c * U[:N] - sn * U[N:], sn * U[:N] + c * U[N:]])


def _omega(s, epx, epy, epz_i):
    """Omega^2, Omega_R and the two n x N factors of Omega^2 (``solve_2d.py:2377-2418``)."""
    N = s['N']
    kx, ky = s['kx'], s['ky']
    I = np.eye(N)
    BE = np.concatenate([kx[:, None] * epz_i * kx[None, :] - I, ky[:, None] * epz_i * kx[None, :]])
    CD = np.concatenate([kx[:, None] * epz_i * ky[None, :], ky[:, None] * epz_i * ky[None, :] - I])
    O2 = np.zeros((2 * N, 2 * N), complex)
    O2[:N, :N] = np.diag(ky * ky)
    O2[:N, N:] = np.diag(-kx * ky)
    O2[N:, :N] = np.diag(-ky * kx)
    O2[N:, N:] = np.diag(kx * kx)
    O2[:, :N] += BE @ epx
    O2[:, N:] += CD @ epy
    OR = np.block([[np.diag(-kx * ky), np.diag(kx * kx) - epy], [epx - np.diag(ky * ky), np.diag(ky * kx)]])
    return O2, OR, BE, CD


def forward_2d(s, conv, thickness):
    """Efficiencies (3 variants x (R_s, R_p, T_s, T_p) x N) and the state of the reverse pass."""
    n = s['n']
    N = s['N']
    I = np.eye(n)
    layers = []
    for (epx, epy, epz_i), d in zip(conv, thickness):
        O2, OR, BE, CD = _omega(s, epx, epy, epz_i)
        w, W = np.linalg.eig(O2)
        q = orc.mode_q(w)
        U = _rot_v(OR @ W, s)                   # R' Omega_R W, so V_sp = U Q^-1
        Wi = np.linalg.inv(_rot_w(W, s))
        Vi = q[:, None] * np.linalg.inv(U)
        A = Wi * s['Dw'][None, :] + Vi * s['Dv'][None, :]
        B = Wi * s['Dw'][None, :] - Vi * s['Dv'][None, :]
        Y = B @ np.linalg.inv(A)
        x = np.exp(-s['k0'] * d * q)
        XYX = x[:, None] * Y * x[None, :]
        M = A - XYX @ B
        S11 = np.linalg.solve(M, XYX @ A - B)
        S12 = np.linalg.solve(M, x[:, None] * (A - Y @ B))
        layers.append(dict(W=W, q=q, Wi=Wi, Vi=Vi, S11=S11, S12=S12, BE=BE, CD=CD, epx=epx,
                           epy=epy, d=d))
    dg = s['dg']
    G = [np.diag(dg['r22'])]
    Mx, Xs, H = [], [], []
    for lay in layers:
        Mx.append(I - lay['S11'] @ G[-1])
        Xs.append(np.linalg.solve(Mx[-1], lay['S12']))
        H.append(G[-1] @ Xs[-1])
        G.append(lay['S11'] + lay['S12'] @ H[-1])
    E = np.zeros((n, 2), complex)
    E[N // 2, 0] = 1
    E[N + N // 2, 1] = 1
    Tf = I - dg['t11'][:, None] * G[-1]
    u = np.linalg.solve(Tf, dg['t12'][:, None] * E)
    cref = dg['t22'][:, None] * E + dg['t21'][:, None] * (G[-1] @ u)
    L = len(layers)
    dvec = [None] * (L + 1)
    dvec[L] = u
    for l in range(L - 1, -1, -1):
        dvec[l] = np.linalg.solve(Mx[l], layers[l]['S12'] @ dvec[l + 1])
    ctrn = dg['r12'][:, None] * dvec[0]
    eff = np.zeros((3, 4, N))
    for v in range(3):
        c_r = cref[:, 0] * s['al'][v] + cref[:, 1] * s['be'][v]
        c_t = ctrn[:, 0] * s['al'][v] + ctrn[:, 1] * s['be'][v]
        for k, a in enumerate((c_r[:N], -c_r[N:], c_t[:N], c_t[N:])):
            eff[v, k] = np.abs(a) ** 2 * s['adm'][k] / s['den'][v]
    return eff, dict(layers=layers, G=G, Mx=Mx, Xs=Xs, H=H, Tf=Tf, d=dvec, cref=cref, ctrn=ctrn)


def expm_divided_differences(q, sc):
    """(f(q_i) - f(q_j)) / (q_i - q_j) for f(z) = exp(sc z), with the series
    f(q_j) sc (1 + z/2 + z^2/6 + z^3/24 + z^4/120), z = sc (q_i - q_j), when |z| < 1e-3."""
    f = np.exp(sc * q)
    dq = q[:, None] - q[None, :]
    z = sc * dq
    with np.errstate(all='ignore'):
        full = (f[:, None] - f[None, :]) / dq
    series = f[None, :] * sc * (1 + z / 2 + z ** 2 / 6 + z ** 3 / 24 + z ** 4 / 120)
    return np.where(np.abs(z) < 1e-3, series, full)


def _h(a):
    return a.conj().T


def backward_2d(s, st, eff_bar):
    """Reverse pass: cotangent of the efficiencies (3, 4, N) -> per layer
    (epx_bar, epy_bar, epz_bar, thickness_bar)."""
    N, n = s['N'], s['n']
    dg = s['dg']
    layers, G, Mx, Xs, H, dvec = st['layers'], st['G'], st['Mx'], st['Xs'], st['H'], st['d']
    L = len(layers)
    crb = np.zeros((n, 2), complex)
    ctb = np.zeros((n, 2), complex)
    for v in range(3):
        al, be = s['al'][v], s['be'][v]
        c_r = st['cref'][:, 0] * al + st['cref'][:, 1] * be
        c_t = st['ctrn'][:, 0] * al + st['ctrn'][:, 1] * be
        w = 2.0 / s['den'][v]
        gr = np.concatenate([c_r[:N] * eff_bar[v, 0] * s['adm'][0], c_r[N:] * eff_bar[v, 1] * s['adm'][1]]) * w
        gt = np.concatenate([c_t[:N] * eff_bar[v, 2] * s['adm'][2], c_t[N:] * eff_bar[v, 3] * s['adm'][3]]) * w
        crb += gr[:, None] * np.conj([al, be])[None, :]
        ctb += gt[:, None] * np.conj([al, be])[None, :]
    u = dvec[L]
    gub = np.conj(dg['t21'])[:, None] * crb
    db = np.conj(dg['r12'])[:, None] * ctb
    Mxb_sweep, S12b_sweep = [None] * L, [None] * L
    for l in range(L):
        z = np.linalg.solve(_h(Mx[l]), db)
        Mxb_sweep[l] = -z @ _h(dvec[l])
        S12b_sweep[l] = z @ _h(dvec[l + 1])
        db = _h(layers[l]['S12']) @ z
    ub = _h(G[L]) @ gub + db
    zf = np.linalg.solve(_h(st['Tf']), ub)
    Gb = (gub + np.conj(dg['t11'])[:, None] * zf) @ _h(u)
    out = [None] * L
    for l in range(L - 1, -1, -1):
        lay = layers[l]
        S11b = Gb.copy()
        S12b = Gb @ _h(H[l]) + S12b_sweep[l]
        Hb = _h(lay['S12']) @ Gb
        Gn = Hb @ _h(Xs[l])
        Z = np.linalg.solve(_h(Mx[l]), _h(G[l]) @ Hb)
        S12b += Z
        Mxb = Mxb_sweep[l] - Z @ _h(Xs[l])
        S11b -= Mxb @ _h(G[l])
        Gb = Gn - _h(lay['S11']) @ Mxb
        out[l] = _layer_backward(s, lay, S11b, S12b)
    return out


def _layer_backward(s, lay, S11b, S12b):
    N = s['N']
    Dw, Dv = s['Dw'], s['Dv']
    Wi, Vi, q, W = lay['Wi'], lay['Vi'], lay['q'], lay['W']
    A = Wi * Dw[None, :] + Vi * Dv[None, :]
    B = Wi * Dw[None, :] - Vi * Dv[None, :]
    Ai = np.linalg.inv(A)
    Y = B @ Ai
    sc = -s['k0'] * lay['d']
    x = np.exp(sc * q)
    XYX = x[:, None] * Y * x[None, :]
    M = A - XYX @ B
    F = A - Y @ B
    E1b = np.linalg.solve(_h(M), S11b)
    E2b = np.linalg.solve(_h(M), S12b)
    Mb = -(E1b @ _h(lay['S11']) + E2b @ _h(lay['S12']))
    Xb = E2b @ _h(F)
    Fb = np.conj(x)[:, None] * E2b
    Ab = Fb + Mb + _h(XYX) @ E1b
    Bb = -_h(Y) @ Fb - E1b - _h(XYX) @ Mb
    Yb = -Fb @ _h(B)
    XYXb = E1b @ _h(A) - Mb @ _h(B)
    Xb += (XYXb * np.conj(x)[None, :]) @ _h(Y) + _h(Y) @ (np.conj(x)[:, None] * XYXb)
    Yb += np.conj(x)[:, None] * XYXb * np.conj(x)[None, :]
    T = Yb @ _h(Ai)
    Bb += T
    Ab -= _h(Y) @ T
    Vib = (Ab - Bb) * np.conj(Dv)[None, :]            # the eigenbasis side Wi is held fixed
    Gm = Vi / q[:, None]                             # G = U^-1, Vi = Q G
    Qb = Vib @ _h(Gm)
    Ub = -_h(Gm) @ (np.conj(q)[:, None] * Vib) @ _h(Gm)
    ORb = _rot_v_t(Ub, s) @ _h(W)
    Qb += Xb * np.conj(expm_divided_differences(q, sc))
    d_bar = float(np.real(np.sum(np.conj(sc / lay['d'] * q * x) * np.diag(Xb))))
    Qt = Qb / np.conj(q[:, None] + q[None, :])
    c, sn = s['fc'][None, :], s['fs'][None, :]
    Winv = np.concatenate([-Wi[:, :N] * sn + Wi[:, N:] * c, Wi[:, :N] * c + Wi[:, N:] * sn], axis=1)
    O2b = _h(Winv) @ Qt @ _h(W)                      # W^-1 = (R W)^-1 R
    epx_b = _h(lay['BE']) @ O2b[:, :N] + ORb[N:, :N]
    epy_b = _h(lay['CD']) @ O2b[:, N:] - ORb[:N, N:]
    BEb = O2b[:, :N] @ _h(lay['epx'])
    CDb = O2b[:, N:] @ _h(lay['epy'])
    kx, ky = np.conj(s['kx']), np.conj(s['ky'])
    epz_b = (kx[:, None] * BEb[:N] * kx[None, :] + ky[:, None] * BEb[N:] * kx[None, :]
             + kx[:, None] * CDb[:N] * ky[None, :] + ky[:, None] * CDb[N:] * ky[None, :])
    return epx_b, epy_b, epz_b, d_bar


# --------------------------------------------------------------------------- assembly

def series_matrix(n, half):
    """T (4 half + 1, n) with c_k = sum_j T[k, j] a_j: the pixel-corrected series of
    ``factorization.py:289-311, 369-440`` as the linear map it is."""
    K = 4 * half + 1
    T = np.zeros((K, n), complex)
    for kk in range(K):
        k = kk - 2 * half
        if n > 1:
            g = np.sinc(k / n) * np.exp(-1j * _PI * k / n)
            om = np.exp(-2j * _PI * k * np.arange(n) / n)
            T[kk, 1:] = g * om[1:] / n
            T[kk, 0] = -g * om[1:].sum() / n
        if k == 0:
            T[kk, 0] += 1.0
    return T


def li_conv_adjoint(eps, hu, hv, epx_bar, epy_bar, epz_bar):
    """Cotangent of one raster (Ny, Nx) from those of ``rcwa_oracle.li_conv_matrices``'s
    outputs (fixed nesting, ``factorization.py:523-671``)."""
    Ny, Nx = eps.shape
    nv, nu = 2 * hv + 1, 2 * hu + 1
    N = nu * nv
    Tv, Tu = series_matrix(Ny, hv), series_matrix(Nx, hu)
    lines0, lines1 = (Tv @ eps).T, (Tv @ (1 / eps)).T                 # (Nx, Kv)
    r = np.arange(nv)
    toe = r[:, None] - r[None, :] + 2 * hv
    inner0, inner1 = np.linalg.inv(lines0[:, toe]), np.linalg.inv(lines1[:, toe])
    m = np.arange(nu)
    dm = m[:, None] - m[None, :] + 2 * hu
    dn = r[:, None] - r[None, :] + 2 * hv

    def outer(inner):
        cu = np.einsum('kj,jab->kab', Tu, inner)
        return cu[dm].transpose(2, 0, 3, 1).reshape(N, N)
    epx = np.linalg.inv(outer(inner0))
    c2d = (Tu @ lines0).T
    epz = np.linalg.inv(c2d[dn[:, None, :, None], dm[None, :, None, :]].reshape(N, N))
    pre_b = -_h(epx) @ epx_bar @ _h(epx)
    zpre_b = -_h(epz) @ epz_bar @ _h(epz)
    c2d_b = np.zeros_like(c2d)
    np.add.at(c2d_b, (np.broadcast_to(dn[:, None, :, None], (nv, nu, nv, nu)),
                      np.broadcast_to(dm[None, :, None, :], (nv, nu, nv, nu))), zpre_b.reshape(nv, nu, nv, nu))

    def outer_adj(out_b):
        cu_b = np.zeros((4 * hu + 1, nv, nv), complex)
        np.add.at(cu_b, dm, out_b.reshape(nv, nu, nv, nu).transpose(1, 3, 0, 2))
        return np.einsum('kj,kab->jab', np.conj(Tu), cu_b)

    def inner_adj(inner, inner_b):
        toe_b = -np.conj(np.swapaxes(inner, 1, 2)) @ inner_b @ np.conj(np.swapaxes(inner, 1, 2))
        lb = np.zeros((Nx, 4 * hv + 1), complex)
        for a in range(nv):
            for b in range(nv):
                lb[:, toe[a, b]] += toe_b[:, a, b]
        return lb
    lines0_b = inner_adj(inner0, outer_adj(pre_b)) + (_h(Tu) @ c2d_b.T)
    lines1_b = inner_adj(inner1, outer_adj(epy_bar))
    return _h(Tv) @ lines0_b.T + (_h(Tv) @ lines1_b.T) * np.conj(-1 / eps ** 2)


# --------------------------------------------------------------------------- end to end

def raster_gradient(eps_layers, thickness, eps_top, eps_bot, period, wavelengths, theta, phi,
                    pol, orders, eff_bar):
    """Objective sum_w <eff_bar[w], eff(w)> over the given wavelengths (2D path) and its
    gradient: returns (objective, d_eps (L, Ny, Nx) complex, d_thickness (L,))."""
    hu, hv = orders
    conv = [orc.li_conv_matrices(e, hu, hv) for e in eps_layers]
    L = len(conv)
    N = (2 * hu + 1) * (2 * hv + 1)
    cbar = [[np.zeros((N, N), complex) for _ in range(3)] for _ in range(L)]
    t_bar = np.zeros(L)
    total = 0.0
    for w_i, wl in enumerate(wavelengths):
        s = setup_point(wl, period, eps_top, eps_bot, theta, phi, pol, hu, hv)
        eff, st = forward_2d(s, conv, thickness)
        total += float(np.sum(eff * eff_bar[w_i]))
        for l, g in enumerate(backward_2d(s, st, eff_bar[w_i])):
            for k in range(3):
                cbar[l][k] += g[k]
            t_bar[l] += g[3]
    d_eps = np.stack([li_conv_adjoint(np.asarray(eps_layers[l], complex), hu, hv, *cbar[l])
                      for l in range(L)])
    return total, d_eps, t_bar
