"""CPU oracle for the normal-vector-field (NVF) factorization of raster layers: a NumPy
restatement of the reference's ``factorization='nvf'`` raster path with its default normal-field
construction (``raster_normal_method=None`` = ``'distance'``) and projector
(``nvf_projector=None`` = ``'linear'``, ``P = n n^T``), sampled or integrated coefficients.

TEST INFRASTRUCTURE ONLY. Nothing in the product package imports this module: only ``tests/``
and ``bench.py``'s ``cpu_baseline`` leg may use it, and there only as the checker, never as the
thing measured or shipped.

The distance transform is SciPy's ``distance_transform_edt``, the routine the reference itself
calls (``factorization.py:1054-1100``, SciPy 1.15 here; the reference pins >= 1.18, whose
transform is the same exact Euclidean one). Everything else is restated from the reference's
arithmetic (paths relative to ``src/wavebit/rcwa/_engine/``):

* the floor lift to ``4 * order + 1`` samples per axis (``factorization.py:811-827, 854-878``);
* the periodic interface mask (``factorization.py:1467-1481``);
* the periodic distance map on a 3 x 3 tiling, its wrapping central difference and the pull-back
  of the index gradient through the cell's ``J^-T`` (``factorization.py:1054-1178``, the metric
  sampling of ``normal_methods.py:138-191``);
* the unit field with its flat-pixel floor, and the constant field of a one-cell axis
  (``factorization.py:1180-1404``);
* the point-sample series (half-pixel phase, no box transform; ``factorization.py:322-338,
  709-738``) of ``n_x^2``, ``n_x n_y``, ``n_y^2`` (``factorization.py:1934-1966``);
* the closing ``eps_xx = E - (Delta C_xx + C_xx Delta)/2`` etc. with ``Delta = E - <1/eps>^-1``
  and the slot order ``(eps_xx, eps_yy, inv(E), -eps_xy, -eps_yx)`` (``factorization.py:
  2461-2540, 2796-2838``), with the singular-value floor of ``<1/eps>`` on layers that are not
  passive dielectrics (``factorization.py:1968-2160``).

Pinned (``tests/test_oracle_nvf.py``): SciPy's transform against a brute-force periodic
distance on small masks; translation and transposition covariance of the field; the constant
field of a one-row raster reducing NVF to Li's 1D rules (so the RETICOLO 1D goldens hold through
it); a uniform layer giving ``eps * I``; the projector identity ``C_xx + C_yy = I`` off flat
pixels; energy conservation on lossless layers; the closing with a constant 45-degree normal
equal to the rotated lamellar problem solved on the axis-aligned path; the singular-value floor
against its definition (``U max(S, f) V^H``) and its transparency where it does not bind.
"""
import numpy as np
from scipy.ndimage import distance_transform_edt

from oracle import rcwa_oracle as orc

FLAT_PIXEL = 1e-9            # factorization.py:1462-1464 (_FLAT_PIXEL)
EQUAL_STEPS = 1e-12          # normal_methods.py:138 (_EQUAL_STEPS)
INVERSE_RULE_BOUND = 2.0     # factorization.py:1968-1992


def normal_field_floor(shape, hu, hv):
    """Block-repeat factors ``(repeat_v, repeat_u)`` that lift a raster to the ``4*order+1``
    samples the point-sample series needs (``factorization.py:811-827``)."""
    ny, nx = shape
    rv = max(1, (4 * hv + 1) // ny + 1) if ny < 4 * hv + 1 else 1
    ru = max(1, (4 * hu + 1) // nx + 1) if nx < 4 * hu + 1 else 1
    return rv, ru


def lift(eps, hu, hv):
    """The raster block-repeated to its floor (``factorization.py:854-878``)."""
    rv, ru = normal_field_floor(eps.shape, hu, hv)
    return np.repeat(np.repeat(eps, rv, axis=0), ru, axis=1)


def interface_mask(eps):
    """Pixels differing from a 4-neighbour, neighbours wrapping (``factorization.py:1467-1481``)."""
    border = np.zeros(eps.shape, dtype=bool)
    for axis in (0, 1):
        differs = eps != np.roll(eps, 1, axis=axis)
        border |= differs
        border |= np.roll(differs, -1, axis=axis)
    return border


def cell_metric(lattice, ny, nx):
    """``(step_v, J^-T)`` of the raster's index steps in a cell whose columns are ``a1``, ``a2``,
    scaled to one u-step (``factorization.py:1110-1133``, ``lattice.py:819-839``): ``step_v`` is
    the physical length of one v-step in u-steps (``None`` when equal to 1 within
    ``EQUAL_STEPS``, ``normal_methods.py:141-191``), ``J^-T`` the 2 x 2 map from
    ``(df/di, df/dj)`` to ``(df/dx, df/dy)``."""
    A = np.asarray(lattice, dtype=np.float64)
    J = np.column_stack((A[:, 0] / nx, A[:, 1] / ny))
    scale = float(np.hypot(J[0, 0], J[1, 0]))
    J = J / (scale or 1.0)
    step_j = float(np.hypot(J[0, 1], J[1, 1]))
    step_i = float(np.hypot(J[0, 0], J[1, 0]))
    same = abs(step_j - step_i) <= EQUAL_STEPS * max(abs(step_j), abs(step_i))
    return (None if same else (step_j, step_i)), np.linalg.inv(J).T


def periodic_edt(mask, sampling=None):
    """Distance of every pixel to the nearest interface pixel of the periodic cell: the exact
    Euclidean transform of the 3 x 3 tiling, cropped to the centre (``factorization.py:
    1076-1086``)."""
    ny, nx = mask.shape
    tiled = np.tile(~mask, (3, 3))
    full = (distance_transform_edt(tiled) if sampling is None
            else distance_transform_edt(tiled, sampling=sampling))
    return full[ny:2 * ny, nx:2 * nx]


def periodic_gradient(f):
    """Wrapping central differences along axis 0 (v) and axis 1 (u)
    (``factorization.py:1088-1095``)."""
    return ((np.roll(f, -1, axis=0) - np.roll(f, 1, axis=0)) / 2.0,
            (np.roll(f, -1, axis=1) - np.roll(f, 1, axis=1)) / 2.0)


def distance_gradient(mask, lattice=None):
    """Unnormalised Cartesian gradient ``(g_x, g_y)`` of the periodic distance map
    (``factorization.py:1054-1178``, the concrete-basis branches)."""
    if lattice is None:
        g_j, g_i = periodic_gradient(periodic_edt(mask))
        return g_i, g_j
    sampling, JT = cell_metric(lattice, *mask.shape)
    g_j, g_i = periodic_gradient(periodic_edt(mask, sampling))
    return JT[0, 0] * g_i + JT[0, 1] * g_j, JT[1, 0] * g_i + JT[1, 1] * g_j


def constant_normal(shape, lattice=None):
    """The normal of a raster with a one-cell axis: along ``b1`` for a single row, ``b2`` for a
    single column (``factorization.py:1337-1375``)."""
    if lattice is None:
        return (1.0, 0.0) if shape[0] == 1 else (0.0, 1.0)
    recip = np.linalg.inv(np.asarray(lattice, dtype=np.float64))
    b = recip[0] if shape[0] == 1 else recip[1]
    u = b / np.hypot(b[0], b[1])
    return float(u[0]), float(u[1])


def normal_vector_field(eps, lattice=None):
    """Unit normal field ``(n_x, n_y)`` of a (lifted) raster, the ``'distance'`` construction
    (``factorization.py:1180-1404``): the gradient of the periodic distance to the interface
    set, normalised where its length is at least ``FLAT_PIXEL`` and left as it is elsewhere."""
    shape = eps.shape
    if 1 in shape:
        ux, uy = constant_normal(shape, lattice)
        return np.full(shape, ux), np.full(shape, uy)
    gx, gy = distance_gradient(interface_mask(eps), lattice)
    norm = np.hypot(gx, gy)
    norm[norm < FLAT_PIXEL] = 1.0
    return gx / norm, gy / norm


def point_series(a, half):
    """Fourier coefficients of point samples along the last axis, orders ``-2 half..2 half``:
    the DFT of ``a / n`` at bin ``k mod n`` times the half-pixel phase ``exp(-i pi k / n)``, no
    box transform and no base split (``factorization.py:322-338, 369-395, 398-440``)."""
    a = np.asarray(a, dtype=np.complex128)
    n = a.shape[-1]
    orders = np.arange(-2 * half, 2 * half + 1)
    spectrum = a if n == 1 else np.fft.fft(a / n, axis=-1)
    return spectrum[..., orders % n] * np.exp(-1j * orders_pi(orders, n))


def orders_pi(orders, n):
    return orc._PI * orders / n


def projector_matrices(nx_, ny_, hu, hv):
    """``(C_xx, C_xy, C_yy)``: BTTB matrices of the point-sample series of ``n_x^2``,
    ``n_x n_y``, ``n_y^2`` (``factorization.py:1934-1966``, ``P = n n^T``)."""
    conv = lambda f: orc.conv_matrix(np.asarray(f, dtype=np.complex128), hu, hv,
                                     series_u=point_series, series_v=point_series)
    return conv(nx_ ** 2), conv(nx_ * ny_), conv(ny_ ** 2)


def passive(eps):
    """Every value a passive dielectric (``Re eps > 0``, ``Im eps >= 0``): the singular-value
    floor provably never binds there (``factorization.py:1995-2043``)."""
    eps = np.asarray(eps)
    return bool(np.all((eps.real > 0) & (eps.imag >= 0)))


def floor_singular_values(B, floor):
    """``B`` with every singular value below ``floor`` raised to it, singular vectors kept, and
    ``B`` itself where none is below (``factorization.py:2046-2160``)."""
    U, s, Vh = np.linalg.svd(B)
    if not np.any(s < floor):
        return B
    return (U * np.maximum(s, floor)) @ Vh


def nvf_assemble(E, Ainv, Cxx, Cxy, Cyy, eps_bound=None):
    """Close the factorization (``factorization.py:2461-2540``):
    ``(epx, epy, epz_i, epxy, epyx) = (eps_xx, eps_yy, E^-1, -eps_xy, -eps_xy)``."""
    if eps_bound is not None:
        Ainv = floor_singular_values(Ainv, 1.0 / (INVERSE_RULE_BOUND * eps_bound))
    delta = E - np.linalg.inv(Ainv)
    eps_xx = E - 0.5 * (delta @ Cxx + Cxx @ delta)
    eps_yy = E - 0.5 * (delta @ Cyy + Cyy @ delta)
    eps_xy = -0.5 * (delta @ Cxy + Cxy @ delta)
    slot = -eps_xy
    return eps_xx, eps_yy, np.linalg.inv(E), slot, slot


def nvf_conv_matrices(eps, hu, hv, lattice=None, normal=None, coefficients='sampled'):
    """``(epx, epy, epz_i, epxy, epyx)`` of one raster layer under NVF
    (``factorization.py:2796-2915, 3235-3300``). ``lattice``: the 2 x 2 cell basis (columns
    ``a1``, ``a2``); the engine passes ``diag(period)`` for an axis-aligned cell
    (``engine.py:1975-2006``). ``normal``: an explicit ``(n_x, n_y)`` on the lifted grid, in place
    of the constructed one. ``coefficients='integrated'`` takes ``E`` and ``<1/eps>`` from the
    closed-form segment series of the compressed raster; the projector is always read off the
    lifted raster's point samples."""
    eps = np.asarray(eps, dtype=np.complex128)
    grid = lift(eps, hu, hv)
    if normal is None:
        normal = normal_vector_field(grid, lattice)
    if coefficients in (None, 'sampled'):
        E = orc.conv_matrix(grid, hu, hv)
        Ainv = orc.conv_matrix(1.0 / grid, hu, hv)
    else:
        comp, su, sv = orc._integrated_parts(eps)
        E = orc.conv_matrix(comp, hu, hv, su, sv)
        Ainv = orc.conv_matrix(1.0 / comp, hu, hv, su, sv)
    Cxx, Cxy, Cyy = projector_matrices(normal[0], normal[1], hu, hv)
    bound = None if passive(eps) else float(np.abs(eps).max())
    return nvf_assemble(E, Ainv, Cxx, Cxy, Cyy, bound)


def lattice_of(period):
    """The cell basis the NVF normal field reads: ``diag(period)`` for ``(Lx, Ly)``, else the
    2 x 2 basis itself (columns ``a1``, ``a2``)."""
    p = np.asarray(period, dtype=np.float64)
    if p.shape == (2,):
        return np.diag(p)
    return p.reshape(2, 2)
