"""In-layer fields of a solved run: ``Result.field(x=, y=, z=)`` (reference
``src/wavebit/rcwa/results.py:149-227``, ``rcwa/_engine/field_result.py``, ``rcwa/_engine/fields.py``).

A run with ``Outputs(fields=True)`` keeps, on the device, every layer's modes (``W``, ``V``,
``q`` from ``wb_layer_modes``) and mode amplitudes (``wb_field_state``); sampling is one batched
``wb_field_sample`` launch sequence over every (wavelength, layer) pair (chunked to bound its
workspace), each layer taking the heights that fall in it, and the values come back as host
arrays of shape ``(wavelength, z, y, x, component)``, ``E`` in V/m per unit incident amplitude
times the source's ``amplitude``, ``H`` in A/m.

Sampling rules (the reference's, ``field_result.py:85-114``, ``:117-160``, ``:330-363``):
``x`` and ``y`` are measured from the cell centre; an integer count spans the cell
(``linspace(-1/2, 1/2, count)`` of each lattice vector), anything else is taken as
coordinates. ``z`` is measured up from the bottom interface; a count samples every layer at
``linspace(0, 1, count)`` of its thickness, both faces included (an internal face appears
twice, once per layer); explicit heights within a few roundings of a face are moved onto it,
a face belongs to the layer above it and the top face to the top layer, and a height outside
the stack is refused. On a skewed cell counts are needed on both axes (or coordinates on both).
"""
from dataclasses import dataclass, field
from types import MappingProxyType

import numpy as np

from ..errors import UnsupportedFeatureError, ValidationError
from ..units import Z0

COMPONENTS = ('Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz')
RECONSTRUCTIONS = ('plain', 'lalanne', 'split', 'splitw')
_FACE_ULPS_PER_LAYER = 4          # field_result.py:122
_MIRROR = (1.0, 1.0, -1.0, -1.0, -1.0, 1.0)   # z -> -z: E polar, H axial
_WORKSPACE_BYTES = 2 << 30        # largest workspace of one sampling launch sequence


@dataclass(frozen=True)
class FieldSamples:
    """Field values with their sample coordinates and components (reference
    ``observations/fields.py:22-41``): ``values`` is ``(wavelength, z, y, x, component)``;
    ``units`` maps each component to its unit and ``'coordinates'`` to the length unit."""
    coordinates: object
    values: object
    components: tuple
    units: object
    sampling: str
    metadata: object = field(default_factory=dict)

    def __post_init__(self):
        object.__setattr__(self, 'components', tuple(self.components))
        for name in ('coordinates', 'units', 'metadata'):
            object.__setattr__(self, name, MappingProxyType(dict(getattr(self, name))))


def is_count(spec):
    """An integer (not a bool) is a number of samples; anything else is coordinates."""
    return isinstance(spec, (int, np.integer)) and not isinstance(spec, bool)


def check_coordinates(name, value):
    """A positive integer count, or a nonempty 1-D array of finite real coordinates
    (reference ``results.py:48-69``)."""
    if is_count(value):
        if value >= 1:
            return
    else:
        try:
            a = np.asarray(value)
            usable = (a.dtype.kind in 'biufc' and a.ndim == 1 and a.size > 0
                      and bool(np.all(np.isfinite(a))) and not np.any(np.imag(a)))
        except (TypeError, ValueError):
            usable = False
        if usable:
            return
    raise ValidationError(f'{name} must be a positive integer count or a nonempty one-dimensional '
                          'array of finite real coordinates.', code='field.coordinates', path=name)


def xy_grid(cell, x, y):
    """Centre-origin sample points ``(X, Y)`` of shape ``(ny, nx)`` and the coordinates to report.
    ``cell`` is ``(a1x, a2y, a1y, a2x)``."""
    a1x, a2y, a1y, a2x = cell
    skewed = a1y != 0.0 or a2x != 0.0
    if is_count(x) and is_count(y):
        u = np.linspace(-0.5, 0.5, int(x))
        v = np.linspace(-0.5, 0.5, int(y))
        X = u[None, :] * a1x + v[:, None] * a2x
        Y = u[None, :] * a1y + v[:, None] * a2y
        if skewed:
            return X, Y, X, Y
        return X, Y, X[0].copy(), Y[:, 0].copy()
    if skewed and (is_count(x) or is_count(y)):
        raise ValidationError('On a skewed lattice, use counts for both x/y axes or explicit '
                              'Cartesian coordinates for both.', code='field.coordinates', path='x')
    xs = np.linspace(-0.5, 0.5, int(x)) * a1x if is_count(x) else np.asarray(x, dtype=np.float64)
    ys = np.linspace(-0.5, 0.5, int(y)) * a2y if is_count(y) else np.asarray(y, dtype=np.float64)
    X, Y = np.broadcast_arrays(xs[None, :], ys[:, None])
    return X, Y, (xs if is_count(x) else x), (ys if is_count(y) else y)


def heights(thickness, z):
    """``(z, owner, edges)``: the public heights sampled, the layer owning each, the faces."""
    ds = np.asarray(thickness, dtype=np.float64)
    edges = np.concatenate([[0.0], np.cumsum(ds)])
    L = ds.shape[0]
    if is_count(z):
        frac = np.linspace(0.0, 1.0, int(z))
        zz = np.concatenate([edges[j] + frac * ds[j] for j in range(L)])
        return zz, np.repeat(np.arange(L), int(z)), edges
    zz = np.asarray(z, dtype=np.float64).reshape(-1).copy()
    tol = _FACE_ULPS_PER_LAYER * max(L, 1) * np.finfo(np.float64).eps * float(np.max(np.abs(edges)))
    if tol > 0:
        gap = edges[None, :] - zz[:, None]
        near = np.argmin(np.abs(gap), axis=1)
        snap = np.abs(gap[np.arange(zz.size), near]) <= tol
        zz = np.where(snap, edges[near], zz)
    if zz.min() < 0.0 or zz.max() > edges[-1]:
        raise ValidationError(
            f'z coordinates must lie inside the stack, 0 to {edges[-1]:g} measured upward from its '
            f'bottom; got {zz.min():g} to {zz.max():g}.', code='field.coordinates', path='z')
    owner = np.minimum(np.sum(zz[:, None] >= edges[None, 1:], axis=1), L - 1)
    return zz, owner, edges


class FieldRecord:
    """The device state a run's fields are sampled from: per wavelength (point ``p``) and layer
    (kernel index ``l``, bottom to top of the solved stack) the modes at pair ``table[p, l]`` of
    ``W``, ``V``, ``q``, the amplitudes at ``amps[p L + l]``, the inverse-rule matrix
    ``epz[conv_idx[p, l]]`` and the thickness ``thick[p, l]``. ``mirrored``: the stack was solved
    upside down (bottom incidence), so public layer ``j`` is kernel layer ``L - 1 - j``, heights
    run down from its top face and the components take the mirror parity."""

    def __init__(self, *, device, hu, hv, mode, pts, thick, W, V, q, amps, table, epz, conv_idx,
                 thickness, cell, mirrored, amplitude, bare=False):
        self.device, self.hu, self.hv, self.mode = device, hu, hv, mode
        self.pts, self.thick, self.W, self.V, self.q, self.amps = pts, thick, W, V, q, amps
        self.table, self.epz, self.conv_idx = table, epz, conv_idx
        self.thickness = np.asarray(thickness, dtype=np.float64)       # public order
        self.cell, self.mirrored, self.amplitude, self.bare = cell, mirrored, complex(amplitude), bare

    def sample(self, x, y, z):
        """``(values (wavelength, z, y, x, 6), coordinates)``, values on the device."""
        import torch
        from .. import _native
        X, Y, x_rep, y_rep = xy_grid(self.cell, x, y)
        zz, owner, edges = heights(self.thickness, z)
        ny, nx = X.shape
        xy = torch.from_numpy(np.stack([X.reshape(-1), Y.reshape(-1)], axis=1).copy()).to(self.device)
        nxy = nx * ny
        P = self.pts.shape[0]
        L = self.thickness.shape[0]
        # each kernel layer's samples, padded to a common count: (public index, local height)
        per = []
        for lk in range(L):
            j = L - 1 - lk if self.mirrored else lk
            sel = np.flatnonzero(owner == j)
            d = float(self.thickness[j])
            local = np.clip(zz[sel] - edges[j], 0.0, d)
            per.append((sel, d - local if self.mirrored else local))
        nz = max(1, max(sel.size for sel, _ in per))
        par = _MIRROR if self.mirrored else (1.0,) * 6
        scale = [self.amplitude * par[c] * (1.0 if c < 3 else 1.0 / Z0) for c in range(6)]
        # chunks of heights and of points that keep the workspace under its bound
        nzc = nz
        while nzc > 1 and _native.field_sample_bytes(1, L, self.hu, self.hv, self.mode, nzc, nxy) > _WORKSPACE_BYTES:
            nzc = (nzc + 1) // 2
        pc = P
        while pc > 1 and _native.field_sample_bytes(pc, L, self.hu, self.hv, self.mode, nzc, nxy) > _WORKSPACE_BYTES:
            pc = (pc + 1) // 2
        values = torch.empty((P, zz.shape[0], nxy, 6), dtype=torch.complex128, device=self.device)
        epz = None if self.mode == _native.MODE_1D_TE else self.epz
        workspace = None
        for z0 in range(0, nz, nzc):
            zl = np.zeros((L, nzc))
            src, dst = [], []
            for lk, (sel, local) in enumerate(per):
                part = sel[z0:z0 + nzc]
                zl[lk, :part.size] = local[z0:z0 + nzc]
                src.extend(lk * nzc + np.arange(part.size))
                dst.extend(part)
            if not src:
                continue
            zl_d = torch.from_numpy(zl).to(self.device)
            src_d = torch.as_tensor(np.asarray(src, dtype=np.int64), device=self.device)
            dst_d = torch.as_tensor(np.asarray(dst, dtype=np.int64), device=self.device)
            for p0 in range(0, P, pc):
                p1 = min(P, p0 + pc)
                if workspace is None:
                    workspace = torch.empty(_native.field_sample_bytes(p1 - p0, L, self.hu, self.hv,
                                                                       self.mode, nzc, nxy),
                                            dtype=torch.uint8, device=self.device)
                out = _native.field_sample(self.pts[p0:p1], self.thick[p0:p1], self.W, self.V, self.q,
                                           self.table[p0:p1], epz, self.conv_idx[p0:p1],
                                           self.amps[p0 * L:p1 * L], zl_d, xy, self.hu, self.hv,
                                           self.mode, scale, workspace=workspace)
                values[p0:p1].index_copy_(1, dst_d, out.index_select(1, src_d))
        values = values.reshape(P, zz.shape[0], ny, nx, 6)
        coords = {'x': x_rep, 'y': y_rep, 'z': zz if is_count(z) else z}
        return values, coords


def field(result, x, y, z, reconstruction='plain'):
    """``Result.field``: see the module docstring."""
    record = result._fields
    if record is None:
        raise UnsupportedFeatureError('This result did not retain field reconstruction state. '
                                      'Request Outputs(fields=True).', code='outputs.unavailable',
                                      path='outputs.fields')
    if reconstruction not in RECONSTRUCTIONS:
        raise ValidationError("reconstruction must be 'plain', 'lalanne', 'split' or 'splitw', not "
                              f'{reconstruction!r}.', code='field.reconstruction', path='reconstruction')
    if reconstruction != 'plain':
        raise UnsupportedFeatureError(
            f"reconstruction={reconstruction!r} is not implemented by this engine; "
            "reconstruction='plain' sums each component's own Fourier series.",
            code='field.reconstruction', path='reconstruction')
    if record.bare:
        raise ValidationError(
            'This stack has no layers, so field() has nothing to sample: it reconstructs inside '
            'finite layers only, and here z = 0 is the interface between the two half-spaces. '
            'Add a layer to sample inside it.', code='field.coordinates', path='z')
    for name, value in (('x', x), ('y', y), ('z', z)):
        check_coordinates(name, value)
    values, coords = record.sample(x, y, z)
    coords['wavelength'] = np.asarray(result.wavelengths)
    units = {c: ('A/m' if c.startswith('H') else 'V/m') for c in COMPONENTS}
    units['coordinates'] = result.provenance.get('length_units')
    return FieldSamples(coordinates=coords, values=values.cpu().numpy(), components=COMPONENTS,
                        units=units, sampling='collocated Cartesian phasor',
                        metadata={'axis_order': ('wavelength', 'z', 'y', 'x', 'component'),
                                  'time_convention': 'exp(-i*omega*t)'})
