"""Seeded synthetic workloads: the benchmark configurations as raster problems.

Each generator returns plain NumPy data (rasters bottom-to-top, thicknesses, half-space
permittivities, period and the list of plane-wave points) so the same inputs feed the GPU
engine, the public lifecycle and the CPU oracle. Lengths are in nm.
"""
import math

import numpy as np

NM = 'nm'


def binary_grating_1d():
    """1D binary grating: period 700 nm, one 300 nm layer eps [[1, 1, 12.25, 12.25]], substrate
    2.25, air above, normal s incidence at 900 nm, orders (10, 0)."""
    return dict(rasters=np.array([[[1.0, 1.0, 12.25, 12.25]]], dtype=np.complex128),
                thickness=np.array([300.0]), eps_top=1.0, eps_bottom=2.25, period=(700.0, 700.0),
                orders=(10, 0), points=[dict(wavelength=900.0, theta=0.0, phi=None, polarization='s')])


def _shapes_raster(rng, n, fill_lo=0.4, fill_hi=0.6, eps=(1.0, 12.25)):
    """A binary n x n raster of random rectangles and discs with fill fraction in [lo, hi]."""
    target = rng.uniform(fill_lo, fill_hi)
    mask = np.zeros((n, n), dtype=bool)
    yy, xx = np.mgrid[0:n, 0:n]
    for _ in range(400):
        if mask.mean() >= target:
            break
        if rng.random() < 0.5:
            w, h = rng.integers(n // 16, n // 3, size=2)
            x0, y0 = rng.integers(0, n, size=2)
            xs = (np.arange(x0, x0 + w) % n)
            ys = (np.arange(y0, y0 + h) % n)
            mask[np.ix_(ys, xs)] = True
        else:
            r = rng.uniform(n / 20, n / 6)
            cx, cy = rng.uniform(0, n, size=2)
            dx = np.minimum(np.abs(xx - cx), n - np.abs(xx - cx))
            dy = np.minimum(np.abs(yy - cy), n - np.abs(yy - cy))
            mask |= dx ** 2 + dy ** 2 <= r ** 2
    return np.where(mask, eps[1], eps[0]).astype(np.complex128)


def crossed_grating_2d(points=512, seed=0, n=256, orders=(7, 7)):
    """3 patterned layers (thickness U[100,400] nm) of n x n binary rasters {1, 12.25} at 40-60%
    fill, square period 1000 nm; points: wavelength U[400,800] nm, theta U[0,30] deg,
    phi U[0,360) deg, s/p alternating."""
    rng = np.random.default_rng(seed)
    rasters = np.stack([_shapes_raster(rng, n) for _ in range(3)])
    thickness = rng.uniform(100.0, 400.0, 3)
    pts = []
    for i in range(points):
        pts.append(dict(wavelength=float(rng.uniform(400, 800)),
                        theta=float(np.deg2rad(rng.uniform(0, 30))),
                        phi=float(np.deg2rad(rng.uniform(0, 360))),
                        polarization='s' if i % 2 == 0 else 'p'))
    return dict(rasters=rasters, thickness=thickness, eps_top=1.0, eps_bottom=2.25,
                period=(1000.0, 1000.0), orders=orders, points=pts)


def continuous_stack(wavelengths=64, seed=0, n=128, orders=(9, 9), layers=4):
    """``layers`` layers of n x n continuous rasters eps U[1, 12], 200 nm each, square period
    1000 nm, air above a 2.25 substrate; normal-incidence p-polarized points at ``wavelengths``
    wavelengths spread over [500, 900] nm. Orders (9, 9): 361 harmonics, 722 x 722 modes."""
    rng = np.random.default_rng(seed)
    rasters = rng.uniform(1.0, 12.0, (layers, n, n)).astype(np.complex128)
    thickness = np.full(layers, 200.0)
    wl = np.linspace(500.0, 900.0, wavelengths)
    pts = [dict(wavelength=float(w), theta=0.0, phi=0.0, polarization='p') for w in wl]
    return dict(rasters=rasters, thickness=thickness, eps_top=1.0, eps_bottom=2.25,
                period=(1000.0, 1000.0), orders=orders, points=pts)


def uneven_problems(count=2048, seed=0, n=128, m_lo=3, m_hi=12, l_lo=1, l_hi=8):
    """``count`` problems: orders per axis U{m_lo..m_hi}, 1-8 layers of thickness U[50,500] nm,
    each an n x n two-material raster (eps' U[1,16], eps'' U[0,2]); substrate eps U[1.5, 4]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        M = int(rng.integers(m_lo, m_hi + 1))
        L = int(rng.integers(l_lo, l_hi + 1))
        mats = rng.uniform(1, 16, 2) + 1j * rng.uniform(0, 2, 2)
        rasters = np.stack([np.where(_shapes_raster(rng, n).real > 1, mats[1], mats[0])
                            for _ in range(L)]).astype(np.complex128)
        out.append(dict(rasters=rasters, thickness=rng.uniform(50, 500, L), eps_top=1.0,
                        eps_bottom=float(rng.uniform(1.5, 4.0)), period=(1000.0, 1000.0),
                        orders=(M, M),
                        points=[dict(wavelength=float(rng.uniform(400, 800)),
                                     theta=float(np.deg2rad(rng.uniform(0, 30))),
                                     phi=float(np.deg2rad(rng.uniform(0, 360))),
                                     polarization='s' if rng.random() < 0.5 else 'p')]))
    return out


def flops_per_point(orders, layers, two_d=True):
    """Algorithmic real FP64 flops of one point: 186 L n^3 + 96 (L+1) n^3, n = 2N (2D) or N.

    Counting rule: GEMM(n) = INV(n) = 8 n^3, EIG(n) = 100 n^3 (complex multiply-add = 8 flops).
    """
    N = (2 * orders[0] + 1) * (2 * orders[1] + 1)
    n = 2 * N if two_d else N
    return 186.0 * layers * n ** 3 + 96.0 * (layers + 1) * n ** 3


def serving_requests(count, seed=0, m_range=(5, 15), l_range=(1, 10), sizes=(64, 128, 256),
                     wl_range=(1, 64), grad_fraction=0.3):
    """A serving stream: orders per axis U{m_range}, layers U{l_range} of n x n rasters
    (n from ``sizes``) with eps' U[1,16], eps'' U[0,1], and U{wl_range} wavelengths in
    [400, 1000] nm; a ``grad_fraction`` of the requests also ask for the raster gradient."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        M = int(rng.integers(m_range[0], m_range[1] + 1))
        L = int(rng.integers(l_range[0], l_range[1] + 1))
        n = int(rng.choice(sizes))
        rasters = (rng.uniform(1, 16, (L, n, n)) + 1j * rng.uniform(0, 1, (L, n, n))).astype(np.complex128)
        k = int(rng.integers(wl_range[0], wl_range[1] + 1))
        out.append(dict(rasters=rasters, thickness=rng.uniform(50, 500, L), eps_top=1.0,
                        eps_bottom=2.25, period=(1000.0, 1000.0), orders=(M, M),
                        wavelengths=np.sort(rng.uniform(400, 1000, k)),
                        theta=float(np.deg2rad(rng.uniform(0, 30))),
                        phi=float(np.deg2rad(rng.uniform(0, 360))),
                        polarization='s' if rng.random() < 0.5 else 'p',
                        gradient=bool(rng.random() < grad_fraction)))
    return out


def to_problem(spec, units='nm'):
    """A ``wavebit_amd.rcwa.Problem`` and ``Numerics`` from a workload dict."""
    import wavebit_amd
    from wavebit_amd import rcwa
    layers = [rcwa.Layer(thickness=float(t), epsilon=e) for t, e in zip(spec['thickness'], spec['rasters'])]
    wls = spec['wavelengths'] if 'wavelengths' in spec else [p['wavelength'] for p in spec['points']]
    p0 = spec['points'][0] if 'points' in spec else spec
    problem = rcwa.Problem(
        stack=rcwa.Stack(layers=layers, top=wavebit_amd.Material(spec['eps_top']),
                         bottom=wavebit_amd.Material(spec['eps_bottom']), period=spec['period']),
        excitation=rcwa.Excitation(wavelength=[float(w) for w in wls],
                                   sources=(rcwa.PlaneWave(polarization=p0['polarization'],
                                                           theta=p0['theta'], phi=p0['phi']),)),
        units=wavebit_amd.Units(length=units))
    return problem, rcwa.Numerics(orders=tuple(spec['orders']))
