"""Host side of the MI355X RCWA engine: points, buckets, device buffers and result unpacking.

A *point* is one (problem, wavelength, theta, phi, polarization) plane-wave solve. Points that
share an order window, a layer count and a solve path form a *bucket*, which is one
``wb_solve`` launch sequence on the device; convolution matrices depend only on a layer's
raster and the order window, so they are assembled once per (raster stack, window) and
indexed by every point that uses them.

Scalar host logic restated from the reference (paths under its ``src/wavebit/``):
``_incident_kxy`` (rcwa/_engine/engine.py:2706-2795) with ``incident_phase_index``
(_common/conventions.py:84-106), the routing of ``_routed_grating_type``
(rcwa/_engine/engine.py:3148-3181) and the polarization resolution of ``_adapt.polarization``
(rcwa/_adapt.py:264-296) and ``incident_sp_amplitudes`` (_common/conventions.py:1069-1115).
"""
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from .. import _native

_PI = 3.141592653589793
_HALF_PI = 1.5707963267948966
_HALF_PI_F32 = float(np.float32(_HALF_PI))
_HALF_PI_F32_BELOW = float(np.nextafter(np.float32(_HALF_PI), np.float32(0)))

PATH_2D, PATH_TE, PATH_TM = '2d', '1d_te', '1d_tm'
_MODE = {PATH_2D: _native.MODE_2D, PATH_TE: _native.MODE_1D_TE, PATH_TM: _native.MODE_1D_TM}


def incident_phase_index(eps_mu, theta):
    """``|Re(k)|/k0`` for a real angle (conventions.py:84-106)."""
    a, b = eps_mu.real, eps_mu.imag
    loss = b / math.cos(theta)
    radius = math.hypot(a, loss)
    if a >= 0:
        return math.sqrt((radius + a) / 2)
    large = math.sqrt((radius - a) / 2)
    return abs(loss) / (2 * (large if large > 0 else 1.0))


def incident_kxy(eps_inc, theta, phi):
    """Incident in-plane wavevector / k0 for real angles (engine.py:2706-2795)."""
    theta = 0.0 if theta is None else float(theta)
    phi = 0.0 if phi is None else float(phi)
    eps_inc = complex(eps_inc)
    if (eps_inc.imag == 0 and eps_inc.real <= 0) or (eps_inc.imag != 0 and abs(theta) >= _PI / 2):
        raise ValueError('real-angle incidence requires a propagating incident wave: a lossless lit '
                         'half-space must carry one (Re(eps*mu) > 0), and in an absorbing one '
                         'abs(theta) must be less than pi/2 radians')
    index = incident_phase_index(eps_inc, theta)
    if theta >= _HALF_PI_F32:
        theta = _HALF_PI_F32_BELOW
    elif theta <= -_HALF_PI_F32:
        theta = -_HALF_PI_F32_BELOW
    s = math.sin(theta)
    return index * s * math.cos(phi), index * s * math.sin(phi)


def route(n_rows, polarization, phi, hv):
    """The solve path a raster takes (engine.py:3148-3181, 1D conical excluded)."""
    if isinstance(polarization, str) and n_rows == 1 and phi is None and hv == 0:
        return PATH_TE if polarization == 's' else PATH_TM
    return PATH_2D


def jones(polarization):
    """The unit ``(a_s, a_p)`` pair (``_adapt.polarization``)."""
    if isinstance(polarization, str):
        return (1.0 + 0j, 0.0 + 0j) if polarization == 's' else (0.0 + 0j, 1.0 + 0j)
    a = np.asarray(polarization, dtype=np.complex128).reshape(-1)
    if a.shape != (2,):
        raise ValueError('Jones polarization must have exactly two components.')
    norm = math.sqrt(float(np.sum(np.abs(a) ** 2)))
    if norm == 0:
        raise ValueError('polarization has zero norm')
    return complex(a[0] / norm), complex(a[1] / norm)


@dataclass
class Point:
    """One plane-wave solve of one layer stack."""
    stack: int              # index of the conv-matrix stack in its bucket
    wavelength: float
    theta: float
    phi: object              # None keeps the pure-1D routing available
    polarization: object
    eps_top: complex
    eps_bottom: complex
    period: tuple
    thickness: tuple
    k_direct: object = None  # (kx, ky) in radians per length unit: PlaneWave(kx=, ky=)


def incident_k(eps_inc, theta, phi, wavelength, k_direct=None):
    """Incident in-plane wavevector / k0: from the angles (``incident_kxy``), or from direct
    physical wavenumbers divided by k0 = 2 pi / wavelength (reference engine.py:2755-2758)."""
    if k_direct is not None:
        scale = float(wavelength) / (2 * _PI)
        return k_direct[0] * scale, k_direct[1] * scale
    return incident_kxy(eps_inc, theta, phi)


@dataclass
class Bucket:
    """Points sharing (orders, layer count, path): one device launch sequence."""
    hu: int
    hv: int
    path: str
    n_layers: int
    rasters: list = field(default_factory=list)       # per stack: (L, Ny, Nx) complex tensor
    points: list = field(default_factory=list)


class ConvCache:
    """Device-resident convolution matrices keyed by (raster identity, orders)."""

    def __init__(self, device):
        self.device = device
        self._store = {}

    def get(self, key, eps, hu, hv, factorization='li', period=None, host=None):
        hit = self._store.get(key)
        if hit is not None:
            return hit
        self._store[key] = assemble_set(eps, hu, hv, factorization, period, host)
        return self._store[key]

    def __len__(self):
        return len(self._store)

    def clear(self):
        self._store.clear()


def lattice_of(period):
    """The cell basis (2 x 2, columns a1 and a2) of a period ``(Lx, Ly)`` or of a skewed cell's
    ``(a1x, a2y, a1y, a2x)``: what the NVF normal field measures its pixels in (the reference
    passes ``diag(period)`` for an axis-aligned cell, ``engine.py:1975-2006``)."""
    a1x, a2y, a1y, a2x = cell_of(period)
    return np.array([[a1x, a2x], [a1y, a2y]], dtype=np.float64)


def assemble_set(eps, hu, hv, factorization='li', period=None, host=None):
    """One raster stack's conv matrices on the device: ``(epx, epy, epz_i, epxy)``, ``epxy``
    (the NVF off-diagonal slot, epxy = epyx) None for the diagonal rules. ``host``: the rasters
    as a host array, which decides whether NVF needs its singular-value floor (a value that is
    not a passive dielectric); without it the floor stage runs. Raises ``FloatingPointError``
    when a layer's matrix is singular."""
    lattice = lattice_of(period) if (factorization == 'nvf' and period is not None) else None
    floor = host is None or not bool(np.all((host.real > 0) & (host.imag >= 0)))
    epx, epy, epz_i, epxy, info = _native.assemble(eps, hu, hv, factorization, lattice, floor)
    L = eps.shape[0]
    inv_info = info[:2 * L] if factorization == 'nvf' else info
    if int(inv_info.abs().sum()) != 0:
        raise FloatingPointError('a layer\'s convolution matrix is singular')
    return epx, epy, epz_i, epxy


def offdiag_cat(sets):
    """The NVF off-diagonal slots of several stacks' conv sets, concatenated as their epx are
    (zeros for a stack of a diagonal rule), or None when no stack has one."""
    if all(len(c) < 4 or c[3] is None for c in sets):
        return None
    return torch.cat([c[3] if len(c) > 3 and c[3] is not None else torch.zeros_like(c[0])
                      for c in sets]).contiguous()


PT_STRIDE = 16   # doubles per point record (WB_PT_STRIDE)


def point_record(pt, path):
    """The WB_PT_* record of one point."""
    kx0, ky0 = incident_k(pt.eps_top, pt.theta, pt.phi, pt.wavelength, pt.k_direct)
    a_s, a_p = jones(pt.polarization) if path == PATH_2D else (1.0 + 0j, 0.0 + 0j)
    eps_t, eps_b = complex(pt.eps_top), complex(pt.eps_bottom)
    py = pt.period[1] if len(pt.period) > 1 else 1.0
    a1y, a2x = (float(pt.period[2]), float(pt.period[3])) if len(pt.period) == 4 else (0.0, 0.0)
    return [pt.wavelength, kx0, ky0, 0.0 if pt.phi is None else float(pt.phi),
            float(pt.period[0]), float(py), eps_t.real, eps_t.imag, eps_b.real, eps_b.imag,
            a_s.real, a_s.imag, a_p.real, a_p.imag, a1y, a2x]


def cell_of(period):
    """``(a1x, a2y, a1y, a2x)`` of a period ``(Lx, Ly)`` or of a skewed cell's 4-tuple."""
    if len(period) == 4:
        return tuple(float(v) for v in period)
    return (float(period[0]), float(period[1]) if len(period) > 1 else 1.0, 0.0, 0.0)


def order_kxy(kx0, ky0, wavelength, period, hu, hv):
    """kx/k0, ky/k0 of the retained orders: vectors (2hu+1,), (2hv+1,) on a rectangular cell;
    on a skewed one (``period`` = (a1x, a2y, a1y, a2x)) two (2hv+1, 2hu+1) grids (reference
    ``lattice.py:899-944``)."""
    m = np.arange(-hu, hu + 1, dtype=np.float64)
    n = np.arange(-hv, hv + 1, dtype=np.float64)
    if len(period) != 4 or (period[2] == 0.0 and period[3] == 0.0):
        return kx0 + m * (wavelength / period[0]), ky0 + n * (wavelength / period[1])
    a1x, a2y, a1y, a2x = (float(v) for v in period)
    lam = wavelength / (a1x * a2y - a2x * a1y)
    return (kx0 + lam * (a2y * m[None, :] - a1y * n[:, None]),
            ky0 + lam * (-a2x * m[None, :] + a1x * n[:, None]))


def bucket_inputs(bucket, conv, device):
    """The device inputs of a bucket's solve: (pts, thick, conv_idx, epx, epy, epz_i)."""
    P, L = len(bucket.points), bucket.n_layers
    # concatenate the stacks' matrices: stack s layer l -> s * L + l (one stack: as they are)
    if len(conv) == 1:
        epx, epy, epz = (c.contiguous() for c in conv[0][:3])
    else:
        epx = torch.cat([c[0] for c in conv]).contiguous()
        epy = torch.cat([c[1] for c in conv]).contiguous()
        epz = torch.cat([c[2] for c in conv]).contiguous()
    # the point records, thicknesses and conv indices in one host buffer: one upload
    nr, nt, ni = P * PT_STRIDE, P * L, (P * L + 1) // 2
    host = np.zeros(nr + nt + ni, dtype=np.float64)
    host[:nr] = np.array([point_record(p, bucket.path) for p in bucket.points], dtype=np.float64).ravel()
    host[nr:nr + nt] = np.array([p.thickness for p in bucket.points], dtype=np.float64).ravel()
    host[nr + nt:].view(np.int32)[:P * L] = np.array(
        [[p.stack * L + l for l in range(L)] for p in bucket.points], dtype=np.int32).ravel()
    dev = torch.from_numpy(host).to(device)
    pts = dev[:nr].view(P, PT_STRIDE)
    thick = dev[nr:nr + nt].view(P, L)
    idx = dev[nr + nt:].view(torch.int32)[:P * L].view(P, L)
    return pts, thick, idx, epx, epy, epz


def rebuild_grazing(bucket, conv, device, eff, amp, pe, flags, smat=None):
    """Solve again the points of a 2D bucket with a grazing layer (``FLAG_GRAZING``), every such
    layer's mode cluster re-derived from the coupled first-order problem (``wb_solve_rebuild``;
    reference ``solve_2d.py:2463-2485``, which always does so). ``eff, amp, pe, flags`` (and
    ``smat``) are the bucket's host arrays, updated in place for those points. Returns the
    number of points solved again."""
    if bucket.path != PATH_2D:
        return 0
    fl = np.asarray(flags)
    untried = (fl & _native.FLAG_GRAZING) & ~(fl & (_native.FLAG_REBUILT | _native.FLAG_DEFECTIVE))
    bad = np.flatnonzero(np.any(untried != 0, axis=1))
    N = (2 * bucket.hu + 1) * (2 * bucket.hv + 1)
    if not bad.size or 2 * N > _native.GRAZE_NMAX:
        return 0
    sub = Bucket(hu=bucket.hu, hv=bucket.hv, path=bucket.path, n_layers=bucket.n_layers,
                 points=[bucket.points[i] for i in bad])
    out = _native.solve_rebuild(*bucket_inputs(sub, conv, device), bucket.hu, bucket.hv,
                                scattering=smat is not None, epxy=offdiag_cat(conv))
    host = [t.cpu().numpy() for t in out]
    for dst, src in zip((eff, amp, pe, flags), host[:4]):
        dst[bad] = src
    if smat is not None:
        smat[bad] = host[4]
    return int(bad.size)


# the point-record entries a layer's modes depend on: the wavelength, the incident in-plane
# wavevector and the cell (the frame azimuth, the half-spaces and the polarization enter later)
_MODES_FIELDS = (0, 1, 2, 4, 5, 14, 15)


def modes_key(bucket):
    """What the layer modes of a bucket's points depend on besides their conv matrices: the
    records' wavelength, incident wavevector and cell entries, as bytes."""
    rec = np.array([point_record(p, bucket.path) for p in bucket.points], dtype=np.float64)
    return (bucket.hu, bucket.hv, bucket.n_layers, rec[:, _MODES_FIELDS].tobytes(),
            tuple(p.stack for p in bucket.points))


def _pair_inputs(buckets, convs, device):
    """One pair per (bucket, point, layer): its point record, the index of its layer's conv
    matrices in the concatenation of every bucket's stacks, and per bucket the (P, L) table of
    pair indices. Returns (pts (pairs, 16), conv_idx (pairs, 1), epx, epy, epz, tables, epxy),
    epxy None unless a stack holds NVF layers."""
    recs, cidx, tables = [], [], []
    base = 0
    for bucket, conv in zip(buckets, convs):
        L = bucket.n_layers
        table = np.empty((len(bucket.points), L), dtype=np.int32)
        for p, pt in enumerate(bucket.points):
            rec = point_record(pt, bucket.path)
            for l in range(L):
                table[p, l] = len(recs)
                recs.append(rec)
                cidx.append(base + pt.stack * L + l)
        tables.append(table)
        base += sum(int(c[0].shape[0]) for c in conv)
    epx = torch.cat([c[0] for conv in convs for c in conv]).contiguous()
    epy = torch.cat([c[1] for conv in convs for c in conv]).contiguous()
    epz = torch.cat([c[2] for conv in convs for c in conv]).contiguous()
    epxy = offdiag_cat([c for conv in convs for c in conv])
    pts = torch.from_numpy(np.array(recs, dtype=np.float64)).to(device)
    idx = torch.from_numpy(np.array(cidx, dtype=np.int32).reshape(-1, 1)).to(device)
    return pts, idx, epx, epy, epz, tables, epxy


def window_modes(buckets, convs, device, rebuild=True, inputs=None):
    """The layer modes of every (point, layer) pair of several buckets that share an order window
    and path, whatever their layer counts, in one ``wb_layer_modes`` launch sequence (one
    eigensolve batch); with ``rebuild``, pairs with a grazing layer are solved again with the
    cluster rebuilt (that check reads the flags, so it waits for the eigensolves; the server
    leaves it to the completion of its launch instead). Returns (W, V, q, flags (pairs,), tables):
    ``tables[i][p, l]`` is the pair index of layer l of point p of bucket i."""
    b0 = buckets[0]
    mode = _MODE[b0.path]
    pts, idx, epx, epy, epz, tables, epxy = inputs or _pair_inputs(buckets, convs, device)
    W, V, q, flags = _native.layer_modes(pts, idx, epx, epy, epz, b0.hu, b0.hv, mode, epxy=epxy)
    flags = flags.reshape(-1)
    N = (2 * b0.hu + 1) * (2 * b0.hv + 1)
    if rebuild and mode == _native.MODE_2D and 2 * N <= _native.GRAZE_NMAX:
        bad = torch.nonzero(flags & _native.FLAG_GRAZING).reshape(-1)
        if bad.numel():
            W2, V2, q2, f2 = _native.layer_modes(pts[bad].contiguous(), idx[bad].contiguous(), epx,
                                                 epy, epz, b0.hu, b0.hv, mode, rebuild=True,
                                                 epxy=epxy)
            W[bad], V[bad], q[bad], flags[bad] = W2, V2, q2, f2.reshape(-1)
    return W, V, q, flags, tables


def _ragged_inputs(buckets, tables, device):
    """Host side of one ragged cascade over several buckets: the points ordered by non-increasing
    layer count (the buckets' points kept together), pairs numbered layer-major (layer l of point
    p at base_l + p). Returns the staged (pts, thick (B,), modes idx (B,)), the layer counts, each
    bucket's first point and its (P, L) table of pair numbers (device int64)."""
    order = sorted(range(len(buckets)), key=lambda i: -buckets[i].n_layers)
    start, recs, lays = {}, [], []
    for i in order:
        start[i] = len(lays)
        recs.extend(point_record(p, buckets[i].path) for p in buckets[i].points)
        lays.extend([buckets[i].n_layers] * len(buckets[i].points))
    P, Lmax = len(lays), max(lays)
    thick = np.zeros((P, Lmax))
    tab = np.zeros((P, Lmax), dtype=np.int32)
    for i, b in enumerate(buckets):
        s, L = start[i], b.n_layers
        thick[s:s + len(b.points), :L] = np.array([p.thickness for p in b.points]).reshape(-1, L)
        tab[s:s + len(b.points), :L] = tables[i]
    _, (pp, ll) = _native.ragged_order(lays)
    pair = np.full((P, Lmax), -1, dtype=np.int64)
    pair[pp, ll] = np.arange(pp.size)
    pairs = [torch.from_numpy(pair[start[i]:start[i] + len(b.points), :b.n_layers].copy()).to(device)
             for i, b in enumerate(buckets)]
    staged = (torch.from_numpy(np.array(recs, dtype=np.float64)).to(device),
              torch.from_numpy(np.ascontiguousarray(thick[pp, ll])).to(device),
              torch.from_numpy(np.ascontiguousarray(tab[pp, ll])).to(device))
    return staged, lays, [start[i] for i in range(len(buckets))], pairs


def run_window(buckets, convs, device, modes=None, rebuild=True, ragged=True):
    """Solve buckets of one order window and path with different layer counts: their modes in one
    batch (``window_modes``, or ``modes`` from an earlier call on the same points), then the
    S-matrices and one cascade over every point (``wb_solve_from_modes_ragged``: cascade step l
    takes layer l of each point that has it). With ``ragged=False``, a single bucket or more than
    ``RAGGED_MAX`` layers, one ``wb_solve_from_modes`` per bucket instead, whose results equal
    ``run_bucket``'s to the bit; the ragged cascade does the same operations per point in
    batches of other sizes (equal within rounding). Every host input goes to the device before
    the first launch, so with ``rebuild=False`` nothing here waits for the device. Returns
    [(eff, amp, pe, flags, N)] per bucket, device tensors, flags combining the modes' and the
    cascade's, and the modes used."""
    b0 = buckets[0]
    mode = _MODE[b0.path]
    N = (2 * b0.hu + 1) * (2 * b0.hv + 1)
    inputs = None
    if modes is None:
        inputs = _pair_inputs(buckets, convs, device)
    tables = inputs[5] if modes is None else modes[4]
    tabs = [torch.from_numpy(t).to(device) for t in tables]
    ragged = (ragged and len(buckets) > 1 and
              max(b.n_layers for b in buckets) <= _native.RAGGED_MAX)
    if ragged:
        (pts_r, thick_r, idx_r), lays, starts, pairs = _ragged_inputs(buckets, tables, device)
    else:
        staged = []
        for bucket in buckets:
            P, L = len(bucket.points), bucket.n_layers
            rec = np.array([point_record(p, bucket.path) for p in bucket.points], dtype=np.float64)
            thick = np.array([p.thickness for p in bucket.points], dtype=np.float64).reshape(P, L)
            staged.append((torch.from_numpy(rec).to(device), torch.from_numpy(thick).to(device)))
    if modes is None:
        modes = window_modes(buckets, convs, device, rebuild=rebuild, inputs=inputs)
    W, V, q, mflags, _ = modes
    out = []
    if ragged:
        eff, amp, pe, flags = _native.solve_from_modes_ragged(pts_r, thick_r, W, V, q, idx_r, lays,
                                                              b0.hu, b0.hv, mode)
        for bucket, s, pair, tab in zip(buckets, starts, pairs, tabs):
            P, L = len(bucket.points), bucket.n_layers
            f = flags[pair] | mflags[tab.reshape(-1).long()].reshape(P, L)
            out.append((eff[s:s + P], amp[s:s + P], pe[s:s + P], f, N))
        return out, modes
    for bucket, (pts, thick), tab in zip(buckets, staged, tabs):
        P, L = len(bucket.points), bucket.n_layers
        eff, amp, pe, flags = _native.solve_from_modes(pts, thick, W, V, q, tab, bucket.hu,
                                                       bucket.hv, mode)
        flags = flags | mflags[tab.reshape(-1).long()].reshape(P, L)
        out.append((eff, amp, pe, flags, N))
    return out, modes


def run_fields(bucket, conv, device):
    """Solve a bucket keeping what its fields are sampled from: every pair's layer modes (grazing
    clusters rebuilt, ``window_modes``) and mode amplitudes for the point's own incidence
    (``wb_field_state``). Returns device tensors (eff, amp, pe, flags) and the state: pts, W, V,
    q, amps (P L, 2, n), the (P, L) pair table, the concatenated epz_i and the (P, L) index of
    each layer's conv matrices in it (device int32), and the (P, L) thicknesses."""
    mode = _MODE[bucket.path]
    P, L = len(bucket.points), bucket.n_layers
    W, V, q, mflags, tables = window_modes([bucket], [conv], device)
    tab = torch.from_numpy(tables[0]).to(device)
    rec = np.array([point_record(p, bucket.path) for p in bucket.points], dtype=np.float64)
    thick = np.array([p.thickness for p in bucket.points], dtype=np.float64).reshape(P, L)
    pts = torch.from_numpy(rec).to(device)
    thick_d = torch.from_numpy(thick).to(device)
    amps, eff, amp, pe, flags = _native.field_state(pts, thick_d, W, V, q, tab, bucket.hu, bucket.hv,
                                                    mode)
    flags = flags | mflags[tab.reshape(-1).long()].reshape(P, L)
    epz = torch.cat([c[2] for c in conv]).contiguous()
    conv_idx = np.array([[p.stack * L + l for l in range(L)] for p in bucket.points], dtype=np.int32)
    state = dict(pts=pts, thick=thick_d, W=W, V=V, q=q, amps=amps, table=tab,
                 epz=epz, conv_idx=torch.from_numpy(conv_idx).to(device))
    return eff, amp, pe, flags, state


def run_bucket(bucket, conv, device, workspace=None, graph_slot=None):
    """Solve every point of a bucket; returns device tensors (eff, amp, pe, flags) and N.

    ``conv`` is a list of (epx, epy, epz_i, epxy) per stack, each (L, N, N) on the device (epxy
    None for the diagonal rules). ``graph_slot`` (a two-element list [graph, last shapes], or
    None) holds a ``graphs.SolveGraph`` reused while the shapes match: launch-bound sizes run
    eagerly the first time, are captured the second time the same shapes come, and are replayed
    afterwards (diagonal rules only).
    """
    P, L = len(bucket.points), bucket.n_layers
    N = (2 * bucket.hu + 1) * (2 * bucket.hv + 1)
    pts_d, thick_d, idx_d, epx, epy, epz = bucket_inputs(bucket, conv, device)
    epxy = offdiag_cat(conv)
    mode = _MODE[bucket.path]
    if graph_slot is not None and epxy is None:
        from . import graphs
        g = graph_slot[0]
        if g is not None and g.matches(pts_d, thick_d, epx, bucket.hu, bucket.hv, mode):
            eff, amp, pe, flags = g.run(pts_d, thick_d, idx_d, epx, epy, epz)
            return eff, amp, pe, flags, N
        key = (tuple(pts_d.shape), tuple(thick_d.shape), tuple(epx.shape), bucket.hu, bucket.hv, mode)
        if graph_slot[1] == key and graphs.worth_capturing(P, L, bucket.hu, bucket.hv, mode):
            # the second run of the same shapes: capture, then replay from now on
            g = graphs.SolveGraph(pts_d, thick_d, idx_d, epx, epy, epz, bucket.hu, bucket.hv, mode)
            graph_slot[0] = g
            eff, amp, pe, flags = g.run(pts_d, thick_d, idx_d, epx, epy, epz)
            return eff, amp, pe, flags, N
        graph_slot[1] = key
    eff, amp, pe, flags = _native.solve(pts_d, thick_d, idx_d, epx, epy, epz, bucket.hu,
                                        bucket.hv, mode, workspace, epxy=epxy)
    return eff, amp, pe, flags, N


def run_bucket_scattering(bucket, conv, device):
    """``run_bucket`` that also returns the stack's published scattering blocks per point:
    (eff, amp, pe, flags, N, smat) with smat (P, 4, n, n) on the device."""
    N = (2 * bucket.hu + 1) * (2 * bucket.hv + 1)
    eff, amp, pe, flags, smat = _native.solve_scattering(
        *bucket_inputs(bucket, conv, device), bucket.hu, bucket.hv, _MODE[bucket.path],
        epxy=offdiag_cat(conv))
    return eff, amp, pe, flags, N, smat


def unpack(eff, amp, pe, p, hu, hv, path):
    """One point's outputs as ``{'incident', 's_inc', 'p_inc'}`` dicts of order-grid arrays."""
    nv, nu = 2 * hv + 1, 2 * hu + 1
    variants = ('incident',) if path != PATH_2D else ('incident', 's_inc', 'p_inc')
    out = {}
    for v, name in enumerate(variants):
        e = eff[p, v].reshape(4, nv, nu)
        a = amp[p, v].reshape(4, nv, nu)
        sub = {'R_s': e[0], 'R_p': e[1], 'T_s': e[2], 'T_p': e[3],
               'r_s': a[0], 'r_p': a[1], 't_s': a[2], 't_p': a[3],
               'power_entering': pe[p, v]}
        sub['R'] = sub['R_s'] + sub['R_p']
        sub['T'] = sub['T_s'] + sub['T_p']
        out[name] = sub
    return out
