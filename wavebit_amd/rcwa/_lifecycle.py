"""``check`` / ``prepare`` / ``Prepared.run`` / ``run`` on the MI355X engine.

Mirrors reference ``src/wavebit/rcwa/_lifecycle.py``: ``check`` (:421-457) admits a problem
and reports resource estimates without solving; ``prepare`` (:736-797) validates concrete
values and freezes what a run reuses; ``Prepared.run`` (:698-733) executes a live problem of
the same structure after the runtime guards (:515-561) and packs a ``Result``; ``run``
(:800-803) is prepare-then-run. Error classes and codes follow the reference: bad inputs are
``ValidationError`` (``rcwa.invalid``), unsupported requests ``UnsupportedFeatureError``
(``rcwa.unsupported`` and friends), reuse violations ``IncompatibleProblemError``
(``prepared.*``), failed factorizations or eigensolves ``NumericalError('rcwa.singular')``
and non-finite observables ``NumericalError('rcwa.nonfinite')``.

Execution: ``backend='torch', device='cuda[:N]'`` (PyTorch-ROCm exposes HIP as ``cuda``) or
the explicit ``backend='hip'`` run the HIP engine; ``precision='double'`` only. The
reference's ``'numpy'`` and ``'jax'`` backends are CPU/JAX paths this engine does not have:
they are refused with ``UnsupportedFeatureError`` rather than silently rerouted.

Unlike the reference, which builds one solver per wavelength in a Python loop
(:712-727), a run here is one batched device solve over all its wavelengths, and the
convolution matrices of each raster stack are kept on the device and reused by every later
run of the same rasters.
"""
import hashlib
from dataclasses import dataclass, field
from types import MappingProxyType

import numpy as np

from ..errors import (IncompatibleProblemError, NumericalError, UnsupportedFeatureError,
                      ValidationError, WavebitError)
from ..materials import Material
from .numerics import Numerics
from .problem import Problem
from .sources import PlaneWave
from . import _engine
from .fields import FieldRecord
from .results import pack


@dataclass(frozen=True)
class Execution:
    backend: str = 'numpy'
    device: str = 'cpu'
    precision: str = 'double'

    def validate(self):
        if self.backend not in ('numpy', 'jax', 'torch', 'hip'):
            raise ValidationError("backend must be one of 'numpy', 'jax', 'torch', or 'hip'.",
                                  code='execution.backend', path='execution.backend',
                                  actual=self.backend)
        if self.precision not in ('double', 'single'):
            raise ValidationError("precision must be 'double' or 'single'.",
                                  code='execution.precision', path='execution.precision')
        if self.backend in ('numpy', 'jax'):
            raise UnsupportedFeatureError(
                f"backend={self.backend!r} is a CPU path; this engine runs the RCWA solve on the "
                "MI355X through HIP: use backend='torch', device='cuda' (or backend='hip').",
                code='rcwa.backend', path='execution.backend')
        if not str(self.device).startswith('cuda'):
            raise UnsupportedFeatureError(
                f"device={self.device!r}: the HIP engine runs on a GPU device ('cuda' or 'cuda:N').",
                code='rcwa.device', path='execution.device')
        if self.precision != 'double':
            raise UnsupportedFeatureError('Only precision="double" (complex128) is implemented.',
                                          code='rcwa.precision', path='execution.precision')


@dataclass(frozen=True)
class CheckReport:
    issues: tuple = ()
    estimates: object = field(default_factory=dict)
    resolved: object = field(default_factory=dict)
    capabilities: object = field(default_factory=dict)

    @property
    def ok(self):
        return not any(i.severity == 'error' for i in self.issues)

    def raise_if_invalid(self):
        for issue in self.issues:
            if issue.severity == 'error':
                raise ValidationError(issue.message, code=issue.code, path=issue.path)


# --------------------------------------------------------------------------- admission

def _host_array(value, path):
    """A host complex128 snapshot of a raster; a live (requires_grad) tensor is read detached --
    its graph is taken up by the differentiable run."""
    if hasattr(value, 'detach'):
        value = value.detach().cpu().numpy()
    try:
        return np.asarray(value, dtype=np.complex128)
    except (TypeError, ValueError):
        raise ValidationError('expected numbers', code='rcwa.invalid', path=path) from None


def _scalar(value, path, real=True, live_ok=False):
    if hasattr(value, 'detach'):
        if getattr(value, 'requires_grad', False) and not live_ok:
            raise UnsupportedFeatureError(
                'Gradients are implemented with respect to raster permittivities and layer '
                'thicknesses; this value carries requires_grad.', code='rcwa.gradient', path=path)
        value = value.detach().cpu().item()
    z = complex(value)
    if real:
        if z.imag != 0:
            raise UnsupportedFeatureError('complex values are not implemented here', code='rcwa.unsupported',
                                          path=path)
        return z.real
    return z


def _is_live(value):
    return bool(getattr(value, 'requires_grad', False))


def _live(problem):
    """True when a layer's raster or thickness is a tensor that requires a gradient (reference
    live contract, _lifecycle.py:785-786)."""
    return any(_is_live(l.epsilon) or _is_live(l.thickness) for l in problem.stack.layers)


def _material(m, path):
    if isinstance(m, (int, float, complex)):
        raise ValidationError('Use a common Material model, not an untyped material number.',
                              code='material.type', path=path)
    if not hasattr(m, 'at'):
        raise ValidationError('Use a common Material model.', code='material.type', path=path)
    return complex(m.at(None).epsilon)


def _wavelengths(problem):
    value = problem.excitation.wavelength
    if hasattr(value, 'detach'):
        value = value.detach().cpu().numpy()
    arr = np.asarray(value, dtype=np.float64) if not isinstance(value, tuple) else np.asarray(
        [_scalar(v, 'excitation.wavelength') for v in value], dtype=np.float64)
    if arr.ndim == 0:
        arr = arr.reshape(1)
    if arr.ndim != 1 or arr.size == 0:
        raise ValidationError('wavelength must be a scalar or a nonempty one-dimensional spectrum.',
                              code='wavelength.shape', path='excitation.wavelength')
    return arr


@dataclass
class _Resolved:
    rasters: np.ndarray        # (L, Ny, Nx) complex, bottom-to-top
    thickness: np.ndarray      # (L,)
    eps_top: complex
    eps_bottom: complex
    period: tuple
    wavelengths: np.ndarray
    source: PlaneWave
    polarization: object
    path: str
    hu: int
    hv: int
    uniform: bool
    factorization: str = 'li'    # 'li' (the default), 'laurent' or 'nvf'
    coefficients: str = 'sampled'    # 'sampled' (the default on a raster) or 'integrated'
    scattering: bool = False         # Outputs(scattering=True): keep the stack's S-matrix
    fields: bool = False             # Outputs(fields=True): keep the in-layer field state
    bare: bool = False               # the stack declared no layers
    thetas: tuple = None             # per-point effective theta after a detuning ladder
    side: str = 'top'                # the lit half-space; 'bottom' is solved on the mirrored stack
    k_direct: tuple = None           # PlaneWave(kx=, ky=): radians per length unit


def _resolve(problem, numerics, execution):
    if not isinstance(problem, Problem):
        raise ValidationError('rcwa lifecycle functions require rcwa.Problem.', code='problem.type',
                              path='problem')
    if not isinstance(numerics, Numerics):
        raise ValidationError('numerics must be rcwa.Numerics.', code='numerics.type', path='numerics')
    execution.validate()
    bad = numerics.unsupported()
    if bad is not None:
        raise UnsupportedFeatureError(
            f'numerics.{bad[0]}={bad[1]!r} is not implemented by this engine (it implements the '
            'raster defaults: SMM, Li rules with fixed nesting, sampled or integrated '
            "coefficients, and the opt-in factorization='laurent' and factorization='nvf' with "
            "its default normal field ('distance') and projector ('linear')).",
            code='rcwa.unsupported', path=f'numerics.{bad[0]}')
    factorization = numerics.factorization or 'li'
    if factorization == 'laurent' and numerics.li_nesting is not None:
        # reference engine.py:5887-5899: an explicit nesting has no reader under Laurent's rule
        raise UnsupportedFeatureError(
            f"li_nesting={numerics.li_nesting!r} has no reader here: factorization='laurent' "
            "nests no rules -- it takes the direct rule for every product, so there is no "
            "inverse for an order to matter to. It selects the order Li's rules are nested in, "
            "so it needs factorization='li'.", code='rcwa.unsupported', path='numerics.li_nesting')
    if factorization == 'nvf' and numerics.li_nesting is not None:
        raise UnsupportedFeatureError(
            f"li_nesting={numerics.li_nesting!r} has no reader here: factorization='nvf' nests no "
            "rules -- it takes the direct rule both ways and closes the discontinuity with the "
            "normal-field projector. It selects the order Li's rules are nested in, so it needs "
            "factorization='li'.", code='rcwa.unsupported', path='numerics.li_nesting')
    for name, value in (('raster_normal_method', numerics.raster_normal_method),
                        ('nvf_projector', numerics.nvf_projector)):
        # reference engine.py:5807-5850: the normal-field options are read by NVF only
        if value is not None and factorization != 'nvf':
            raise UnsupportedFeatureError(
                f"{name}={value!r} has no reader here: factorization={factorization!r} builds no "
                "normal field. It selects the raster NVF construction, so it needs "
                "factorization='nvf'.", code='rcwa.unsupported', path=f'numerics.{name}')
    if not problem.outputs.diffraction:
        raise UnsupportedFeatureError('Outputs(diffraction=False) leaves nothing this engine computes.',
                                      code='rcwa.unsupported', path='outputs.diffraction')
    stack = problem.stack
    if stack.lattice is not None:
        period = _cell(stack.lattice)
    else:
        period = tuple(_scalar(v, 'stack.period') for v in stack.period)
        if len(period) != 2 or not all(p > 0 for p in period):
            raise ValidationError('period must be two positive lengths.', code='stack.period',
                                  path='stack.period')
    skewed = len(period) == 4
    if skewed and factorization == 'li':
        # reference: Li's inverse rules run along the lattice axes and are refused on a skewed
        # cell; Laurent's rule names no axis (factorization.py:991-994), and NVF's tensor is
        # written in Cartesian components (factorization.py:1283-1295)
        raise UnsupportedFeatureError(
            "factorization='li' runs its inverse rules along the cell's axes, which a skewed "
            "lattice does not have: pass factorization='laurent' or 'nvf' on this cell.",
            code='rcwa.unsupported', path='numerics.factorization')
    if skewed and factorization == 'nvf' and numerics.fourier_coefficients == 'integrated':
        # reference factorization.py:3264-3268: ('integrated', 'nvf') is refused on a skewed cell
        raise UnsupportedFeatureError(
            "fourier_coefficients='integrated' with factorization='nvf' is not solved on a skewed "
            "cell: use the sampled coefficients there.", code='rcwa.unsupported',
            path='numerics.fourier_coefficients')
    sources = problem.excitation.sources
    if len(sources) != 1:
        raise UnsupportedFeatureError('One incident plane wave per problem is implemented.',
                                      code='rcwa.unsupported', path='excitation.sources')
    src = sources[0]
    k_direct = None
    if src.kx is not None or src.ky is not None:
        k_direct = tuple(0.0 if v is None else _scalar(v, 'excitation.sources.kx' if i == 0 else
                                                      'excitation.sources.ky')
                         for i, v in enumerate((src.kx, src.ky)))
    theta = None if src.theta is None else _scalar(src.theta, 'excitation.sources.theta')
    phi = None if src.phi is None else _scalar(src.phi, 'excitation.sources.phi')
    layers = stack.layers
    raster = [l.epsilon is not None for l in layers]
    if any(raster) and not all(raster):
        raise UnsupportedFeatureError(
            'A stack cannot mix sampled layers and region layers. Represent homogeneous layers on '
            'the same raster, or use regions throughout.', code='rcwa.mixed_layer_encoding',
            path='stack.layers')
    eps_top = _material(stack.top, 'stack.top')
    eps_bottom = _material(stack.bottom, 'stack.bottom')
    if any(l.regions for l in layers):
        raise UnsupportedFeatureError('Region (shape) layers are not implemented; give sampled '
                                      'epsilon rasters.', code='rcwa.unsupported', path='stack.layers')
    if any(l.mu is not None for l in layers):
        raise UnsupportedFeatureError('Magnetic rasters are not implemented.', code='rcwa.unsupported',
                                      path='stack.layers.mu')
    uniform = not any(raster)
    if not layers:
        rasters = np.full((1, 1, 1), eps_top, dtype=np.complex128)
        thickness = np.zeros(1)
        uniform = True
    elif uniform:
        rasters = np.stack([np.full((1, 1), _material(l.background, f'stack.layers[{i}].background'),
                                    dtype=np.complex128) for i, l in enumerate(layers)])
        thickness = np.array([_scalar(l.thickness, f'stack.layers[{i}].thickness', live_ok=True)
                              for i, l in enumerate(layers)])
    else:
        arrs = [_host_array(l.epsilon, f'stack.layers[{i}].epsilon') for i, l in enumerate(layers)]
        for i, a in enumerate(arrs):
            if a.ndim != 2 or a.size == 0:
                raise ValidationError('epsilon must be a two-dimensional (Ny, Nx) raster.',
                                      code='layer.epsilon_shape', path=f'stack.layers[{i}].epsilon')
        if len({a.shape for a in arrs}) != 1:
            raise ValidationError('Every raster layer of a stack must have the same (Ny, Nx) shape.',
                                  code='layer.epsilon_shape', path='stack.layers')
        rasters = np.stack(arrs)
        thickness = np.array([_scalar(l.thickness, f'stack.layers[{i}].thickness', live_ok=True)
                              for i, l in enumerate(layers)])
    hu, hv = numerics.orders
    if rasters.shape[1] == 1 and hv != 0 and not uniform:
        raise ValidationError(
            'A one-row raster is invariant along y, so it has no y orders: use orders=(M, 0).',
            code='rcwa.invalid', path='numerics.orders')
    if hv > 15:
        raise UnsupportedFeatureError('orders[1] above 15 is not implemented by this engine.',
                                      code='rcwa.unsupported', path='numerics.orders')
    pol = src.polarization
    if not isinstance(pol, str):
        pol = tuple(_scalar(v, 'excitation.sources.polarization', real=False) for v in pol)
        if len(pol) != 2:
            raise ValidationError('Jones polarization must have exactly two components.',
                                  code='source.polarization_shape',
                                  path='excitation.sources.polarization')
    if src.side == 'bottom':
        # Bottom incidence reads the kernel cascade's front side (reference scattering.py:943-949
        # with the backside flag off). The gap-embedded layer S-matrices are symmetric, so that
        # is the top-incidence read of the stack mirrored in z: the half-spaces swapped and the
        # layers in the opposite order, every order's kx / ky unchanged.
        rasters, thickness = rasters[::-1].copy(), thickness[::-1].copy()
        eps_top, eps_bottom = eps_bottom, eps_top
    n_rows = rasters.shape[1] if not uniform else 2
    # a direct ky off zero is conical: route it as an azimuth would be
    route_phi = 0.0 if (k_direct is not None and k_direct[1] != 0.0) else phi
    path = _engine.route(n_rows, pol, route_phi, hv)
    if skewed and path != _engine.PATH_2D:
        raise UnsupportedFeatureError('A skewed lattice is solved on the 2D coupled path: give a '
                                      'Jones polarization or phi, or more than one raster row.',
                                      code='rcwa.unsupported', path='stack.lattice')
    ws = _wavelengths(problem)
    return _Resolved(rasters=rasters, thickness=thickness, eps_top=eps_top, eps_bottom=eps_bottom,
                     period=period, wavelengths=ws, source=src, polarization=pol, path=path,
                     hu=hu, hv=hv, uniform=uniform, factorization=factorization,
                     coefficients=numerics.fourier_coefficients or 'sampled',
                     scattering=bool(problem.outputs.scattering), side=src.side, k_direct=k_direct,
                     fields=bool(problem.outputs.fields), bare=not layers)


def _cell(lattice):
    """A lattice (2 x 2, columns a1 and a2, or an object with ``u`` / ``v``) as ``(a1x, a2y)``
    when axis-aligned -- the rectangular cell, bit-identical to ``period`` -- else
    ``(a1x, a2y, a1y, a2x)``."""
    try:
        if hasattr(lattice, 'u') and hasattr(lattice, 'v'):
            A = np.stack([np.asarray(lattice.u, dtype=np.float64), np.asarray(lattice.v, dtype=np.float64)], axis=1)
        else:
            A = np.asarray(lattice, dtype=np.float64)
    except (TypeError, ValueError):
        raise ValidationError('lattice must be a 2 x 2 basis whose columns are the lattice vectors.',
                              code='stack.lattice', path='stack.lattice') from None
    if A.shape != (2, 2) or not np.all(np.isfinite(A)):
        raise ValidationError('lattice must be a finite 2 x 2 basis whose columns are the lattice vectors.',
                              code='stack.lattice', path='stack.lattice')
    a1x, a1y, a2x, a2y = float(A[0, 0]), float(A[1, 0]), float(A[0, 1]), float(A[1, 1])
    if a1x <= 0 or a2y <= 0 or a1x * a2y - a2x * a1y <= 0:
        raise ValidationError('lattice vectors must form a right-handed basis with a1x, a2y > 0.',
                              code='stack.lattice', path='stack.lattice')
    return (a1x, a2y) if a1y == 0.0 and a2x == 0.0 else (a1x, a2y, a1y, a2x)


def _structure_key(r):
    return (r.rasters.shape, r.thickness.shape, r.wavelengths.shape, r.path, r.hu, r.hv,
            isinstance(r.polarization, str))


def _runtime_checks(r):
    finite = bool(np.all(np.isfinite(r.rasters)) and np.all(np.isfinite(r.thickness))
                  and np.all(np.isfinite(r.wavelengths)) and np.isfinite(r.eps_top)
                  and np.isfinite(r.eps_bottom))
    pol_ok = True
    if not isinstance(r.polarization, str):
        pol_ok = sum(abs(complex(v)) ** 2 for v in r.polarization) > 0
    amp = r.source.amplitude
    amp_ok = abs(complex(amp.item() if hasattr(amp, 'item') else amp)) > 0
    return {'finite_inputs': finite,
            'positive_wavelength': bool(np.all(r.wavelengths > 0)),
            'nonnegative_thickness': bool(np.all(r.thickness >= 0)),
            'nonzero_polarization': bool(pol_ok and amp_ok)}


def _estimates(r):
    N = (2 * r.hu + 1) * (2 * r.hv + 1)
    n = 2 * N if r.path == _engine.PATH_2D else N
    P, L = r.wavelengths.size, r.rasters.shape[0]
    ws = _engine._native.solve_workspace_bytes(P, L, r.hu, r.hv, _engine._MODE[r.path]) if _native_ok() else None
    return {'retained_spatial_modes': N, 'one_dense_material_matrix_bytes': N * N * 16,
            'one_mode_matrix_bytes': n * n * 16, 'solve_workspace_bytes': ws,
            'scope': 'device workspace of one batched run over every wavelength'}


def _native_ok():
    try:
        _engine._native.load_library()
        return True
    except Exception:
        return False


def _report(problem, numerics, execution, r):
    return CheckReport(
        issues=(),
        estimates=MappingProxyType(_estimates(r)),
        resolved=MappingProxyType({'order_window': ((0, r.hu), (0, r.hv)),
                                   'wavelength_count': int(r.wavelengths.size),
                                   'layer_count': len(problem.stack.layers),
                                   'fourier_coefficients': r.coefficients,
                                   'factorization': r.factorization,
                                   'li_nesting': 'fixed' if r.factorization == 'li' else None,
                                   'algorithm': 'smm',
                                   'formulation': 'second_order', 'solve_path': r.path,
                                   'backend': execution.backend, 'device': execution.device,
                                   'precision': execution.precision, 'engine': 'wavebit_amd/hip-gfx950',
                                   'length_units': problem.units.length}),
        capabilities=MappingProxyType({'eager': True, 'batched_wavelengths': True,
                                       'material_gradients': 'wavebit_amd.rcwa.grad',
                                       'diffraction': True}))


def check(problem, *, numerics, backend='numpy', device='cpu', precision='double'):
    """Admission and scoped resource estimates without solving (reference _lifecycle.py:421)."""
    execution = Execution(backend=backend, device=device, precision=precision)
    try:
        r = _resolve(problem, numerics, execution)
        _engine.incident_k(r.eps_top, r.source.theta, r.source.phi, r.wavelengths[0], r.k_direct)
        return _report(problem, numerics, execution, r)
    except WavebitError as error:
        return CheckReport(issues=(error.issue,))
    except (TypeError, ValueError) as error:
        return CheckReport(issues=(ValidationError(str(error), code='rcwa.invalid', path='problem').issue,))


class _Cache:
    """Device-resident convolution matrices per (raster bytes, window) for one preparation."""

    def __init__(self, device):
        self.conv = _engine.ConvCache(device)
        self.graph = [None, None]    # captured solve of this preparation (launch-bound sizes)
        self.modes = None            # (key, modes) of the last run's layer eigensystems

    def sharing_conv(self):
        """A cache with the same convolution matrices and no captured solve or modes (for the
        points a detuning ladder solves again)."""
        other = _Cache.__new__(_Cache)
        other.conv, other.graph, other.modes = self.conv, [None, None], None
        return other

    @staticmethod
    def key(rasters, hu, hv, factorization='li', period=None):
        # NVF's normal field measures the pixels in the cell's metric: the cell is part of the key
        return (hashlib.blake2b(np.ascontiguousarray(rasters).tobytes(), digest_size=16).hexdigest(),
                rasters.shape, hu, hv, factorization,
                _engine.cell_of(period) if factorization == 'nvf' else None)

    def get(self, rasters, hu, hv, factorization='li', period=None):
        import torch
        eps = torch.from_numpy(np.ascontiguousarray(rasters)).to(self.conv.device)
        return self.conv.get(self.key(rasters, hu, hv, factorization, period), eps, hu, hv,
                             factorization, period, rasters)


@dataclass(frozen=True)
class Prepared:
    """A fixed numerical contract and structure; device caches of convolution matrices."""
    numerics: Numerics
    execution: Execution
    report: CheckReport
    contract: object
    _structure: object = field(repr=False)
    _cache: object = field(repr=False, default=None)

    def _resolve_live(self, problem):
        r = _resolve(problem, self.numerics, self.execution)
        if _structure_key(r) != self._structure:
            raise IncompatibleProblemError(
                'The live problem changed a static part of the preparation (layer count, raster '
                'shape, wavelength count, solve path or order window). Prepare again.',
                code='prepared.structure', path='problem')
        return r

    def validate(self, problem):
        try:
            r = self._resolve_live(problem)
            for name, ok in _runtime_checks(r).items():
                if not ok:
                    raise IncompatibleProblemError('The prepared runtime guard failed.',
                                                   code=f'prepared.{name}', path='problem')
            return _report(problem, self.numerics, self.execution, r)
        except WavebitError as error:
            return CheckReport(issues=(error.issue,))

    def run(self, problem):
        """Execute a live problem of the prepared structure; returns a ``Result``."""
        r = self._resolve_live(problem)
        checks = _runtime_checks(r)
        for name, ok in checks.items():
            if not ok:
                raise IncompatibleProblemError('The prepared runtime guard failed.',
                                               code=f'prepared.{name}', path='problem')
        if _live(problem):
            return _run_live(self, problem, r, checks)
        try:
            per_wl, flags, smats, record = _solve_resolved(r, self._cache)
            r2, per_wl, flags, smats = _detune(r, self._cache, per_wl, flags, smats, self.numerics)
            if r.fields and r2 is not r:
                # the ladder moved points: the field state is taken at the effective points
                record = _solve_resolved(r2, self._cache.sharing_conv())[3]
            r = r2
        except FloatingPointError as error:
            raise NumericalError(str(error), code='rcwa.singular', path='problem') from error
        except ValueError as error:
            raise ValidationError(str(error), code='rcwa.invalid', path='problem') from error
        error = flag_error(flags, self.numerics)
        if error is not None:
            raise error
        notice, ep_counts = flag_notice(flags, self.numerics)
        if notice is not None:
            import warnings
            from .diagnostics import ExceptionalPointWarning
            warnings.warn(notice, ExceptionalPointWarning, stacklevel=2)
        prov = {'solver': 'rcwa', 'engine': 'wavebit_amd/hip-gfx950', 'length_units': problem.units.length,
                'execution': self.execution, 'numerics': self.numerics,
                'coordinate_frame': '+z up; bottom interface z=0; layers bottom-to-top',
                'time_convention': 'exp(-i*omega*t)'}
        diag = {'report': self.report, 'solve_path': tuple(r.path for _ in r.wavelengths),
                'effective_wavelengths': tuple(float(w) for w in r.wavelengths),
                'effective_theta': r.thetas or tuple(r.source.theta for _ in r.wavelengths),
                'exceptional_points': ep_counts}
        result = pack(per_wl, r.hu, r.hv, r.wavelengths, checks, diag, prov, scattering=smats,
                      side=r.side, fields=record)
        if not result.validity.checks['numerical_finite']:
            raise NumericalError('The requested solve produced nonfinite observables; no implicit '
                                 'detuning was applied.', code='rcwa.nonfinite', path='result')
        return result


def _run_live(prepared, problem, r, checks):
    """Differentiable run: efficiencies as torch tensors carrying the graph to the layer rasters
    and thicknesses (the hand-written reverse pass of ``adjoint.RasterSolve``)."""
    import torch
    from dataclasses import replace
    from . import adjoint
    one_d = r.path in (_engine.PATH_TE, _engine.PATH_TM)
    if r.uniform or not (r.path == _engine.PATH_2D or one_d):
        raise UnsupportedFeatureError(
            'Gradients are implemented for raster stacks; this problem routes to '
            f'{r.path!r} with uniform layers.', code='rcwa.gradient', path='problem')
    if prepared.numerics.detuning:
        raise UnsupportedFeatureError(
            'A differentiable run solves the requested point; detuning is applied to concrete '
            'runs only.', code='rcwa.gradient', path='numerics.detuning')
    if r.scattering or r.fields:
        what = 'scattering' if r.scattering else 'fields'
        raise UnsupportedFeatureError(
            'Gradients are implemented for the diffraction outputs; a differentiable run cannot '
            f'also keep the {"scattering matrix" if r.scattering else "field state"}. Solve it with '
            'concrete values.', code='rcwa.gradient', path=f'outputs.{what}')
    device = torch.device(prepared._cache.conv.device)
    _engine._native.context(device)
    layers = problem.stack.layers
    eps_list = [l.epsilon if torch.is_tensor(l.epsilon) else torch.as_tensor(np.asarray(l.epsilon))
                for l in layers]
    cplx = any(e.is_complex() for e in eps_list)
    dt = torch.complex128 if cplx else torch.float64
    eps = torch.stack([e.to(dtype=dt) for e in eps_list])
    thick = torch.stack([l.thickness.to(torch.float64).reshape(()) if torch.is_tensor(l.thickness)
                         else torch.tensor(float(l.thickness), dtype=torch.float64) for l in layers])
    if r.side == 'bottom':                   # the mirrored stack (see _resolve); flip is differentiable
        eps, thick = torch.flip(eps, [0]), torch.flip(thick, [0])
    if r.factorization == 'nvf' and not adjoint.passive(eps):
        raise UnsupportedFeatureError(
            "Gradients under factorization='nvf' are implemented for passive dielectric layers "
            '(Re eps > 0, Im eps >= 0 everywhere); a stack with metal or gain is solved with '
            'concrete values.', code='rcwa.gradient', path='numerics.factorization')
    # The pure TE/TM routes differentiate through the coupled path at hv = 0, phi = 0: a one-row
    # (y-invariant) raster there decouples into the same TE (s) and TM (p) problems, the 2D Li
    # rules reducing to the 1D ones, so its efficiencies and their gradients are the 1D route's
    # (the reference differentiates that route by stock autograd, solve_1d.py:151-253). The
    # amplitudes, whose frame signs differ between the routes, come from a concrete 1D solve.
    points = [_engine.Point(stack=0, wavelength=float(wl), theta=r.source.theta,
                            phi=0.0 if one_d else r.source.phi,
                            polarization=r.polarization, eps_top=r.eps_top, eps_bottom=r.eps_bottom,
                            period=r.period, thickness=tuple(float(t) for t in r.thickness), k_direct=r.k_direct)
              for wl in r.wavelengths]
    lattice = _engine.lattice_of(r.period) if r.factorization == 'nvf' else None
    eff, amp, pe, flags = adjoint.solve(eps, thick.to(device), points, r.hu, r.hv, device,
                                        factorization=r.factorization, lattice=lattice)
    fl_h = flags.cpu().numpy()
    error = flag_error(fl_h, prepared.numerics)
    if error is not None:
        raise error
    notice, _ = flag_notice(fl_h, prepared.numerics)
    if notice is not None:
        import warnings
        from .diagnostics import ExceptionalPointWarning
        warnings.warn(notice, ExceptionalPointWarning, stacklevel=3)
    eff_h, amp_h, pe_h = (t.detach().cpu().numpy() for t in (eff, amp, pe))
    if one_d:
        per_wl = _solve_resolved(r, prepared._cache)[0]
    else:
        per_wl = [_engine.unpack(eff_h, amp_h, pe_h, p, r.hu, r.hv, r.path) for p in range(len(points))]
    prov = {'solver': 'rcwa', 'engine': 'wavebit_amd/hip-gfx950 (differentiable)',
            'length_units': problem.units.length, 'execution': prepared.execution,
            'numerics': prepared.numerics}
    diag = {'report': prepared.report, 'solve_path': tuple(r.path for _ in r.wavelengths),
            'gradient': 'hand-written adjoint (raster permittivity, layer thickness)'
                        + (' through the coupled path at hv = 0' if one_d else '')}
    result = pack(per_wl, r.hu, r.hv, r.wavelengths, checks, diag, prov, side=r.side)
    if not result.validity.checks['numerical_finite']:
        raise NumericalError('The requested solve produced nonfinite observables.',
                             code='rcwa.nonfinite', path='result')
    P = len(points)
    nu, nv = 2 * r.hu + 1, 2 * r.hv + 1
    e0 = eff[:, 0]
    live = {}
    for port, (s_i, p_i) in (('reflected', (0, 1)), ('transmitted', (2, 3))):
        es = e0[:, s_i].reshape(P, nv, nu)
        ep = e0[:, p_i].reshape(P, nv, nu)
        live[port] = replace(result.outputs[port], efficiency=es + ep, efficiency_s=es,
                             efficiency_p=ep)
    outputs = dict(result.outputs)
    outputs.update(live)
    return replace(result, outputs=outputs)


def _fatal_bits(numerics):
    """Flag bits that leave no valid answer under ``numerics``: a failed eigensolve or
    factorization, a numerically defective grazing cluster (reference ``solve_2d.py:190-201``
    raises there) and, under ``exceptional_points='rebuild'``, a near-defective pair whose modes
    as computed may be off by more than 1e-10 (the reference would rebuild it)."""
    fl = _engine._native
    bits = fl.FLAG_NOCONVERGE | fl.FLAG_SINGULAR | fl.FLAG_DEFECTIVE
    if numerics.exceptional_points == 'rebuild':
        bits |= fl.FLAG_EP_LOUD
    return bits


def _kept_grazing(flags):
    """Grazing layers whose cluster was not rebuilt (kept as computed, as the reference keeps
    a cluster it cannot restate)."""
    fl = _engine._native
    flags = np.asarray(flags)
    return ((flags & fl.FLAG_GRAZING) != 0) & ((flags & (fl.FLAG_REBUILT | fl.FLAG_DEFECTIVE)) == 0)


def flag_error(flags, numerics):
    """The error a solve's per-pair flags (``WB_FLAG_*``) call for, or None. Every case is
    ``NumericalError('rcwa.singular')``, the code the reference gives a linear-algebra failure
    and its defective-cluster refusal (``_lifecycle.py:718-719`` of the reference)."""
    fl = _engine._native
    flags = np.asarray(flags)
    if np.any(flags & (fl.FLAG_NOCONVERGE | fl.FLAG_SINGULAR)):
        return NumericalError('A layer eigensolve did not converge or a scattering-matrix factorization '
                              'was singular.', code='rcwa.singular', path='problem')
    if np.any(flags & fl.FLAG_DEFECTIVE):
        return NumericalError('A 2D layer eigenmode cluster is numerically defective at a grazing '
                              'condition (an order on the layer light line): the coupled mode rates '
                              'split by less than 1e-10 of the layer modal scale, and no exponential '
                              'mode basis represents the layer at working precision. Move off the '
                              'condition by a relative step of the wavelength or angle; '
                              'Numerics(detuning=step) does that retry.',
                              code='rcwa.singular', path='problem')
    if numerics.exceptional_points == 'rebuild' and np.any(flags & fl.FLAG_EP_LOUD):
        return NumericalError('A 2D layer holds a near-defective mode pair (eigenvectors at a sine '
                              'below 1.05e-3, an exceptional point of the layer operator) whose '
                              'modes as computed may be off by more than 1e-10; this engine does not '
                              'rebuild such a pair. Move off the point by a relative step of the '
                              "wavelength or angle (Numerics(detuning=step)), or take the modes as "
                              "computed with Numerics(exceptional_points='off').",
                              code='rcwa.singular', path='problem')
    return None


def flag_notice(flags, numerics):
    """``(message, counts)`` for the layer solves that keep modes the reference would rebuild,
    or ``(None, counts)``.

    A grazing layer (|w| below 1e-9 of the largest) has its cluster re-derived from the coupled
    first-order problem (``wb_solve_rebuild``; reference ``solve_2d.py:2463-2485``); a cluster
    the rebuild cannot restate (above 8 modes, or products that disagree on its size) is kept as
    computed with a warning, as the reference keeps it. Near-defective pairs above the loud sine
    are counted and taken as computed: their error is below 1e-10 by the bound of the pair
    screen."""
    fl = _engine._native
    flags = np.asarray(flags)
    kept = _kept_grazing(flags)
    counts = {'grazing_layer_solves': int(np.count_nonzero(flags & fl.FLAG_GRAZING)),
              'grazing_rebuilt_layer_solves': int(np.count_nonzero(flags & fl.FLAG_REBUILT)),
              'grazing_kept_layer_solves': int(np.count_nonzero(kept)),
              'defective_pair_layer_solves': int(np.count_nonzero(flags & fl.FLAG_EP))}
    if not counts['grazing_kept_layer_solves']:
        return None, counts
    return (f"{counts['grazing_kept_layer_solves']} layer solve(s) hold a grazing mode cluster the "
            'rebuild could not restate (more than 8 modes, or the two operator products disagree '
            'on its size); their modes are kept as computed and results near such a point can '
            'differ from the reference beyond round-off.'), counts


def _detune(r, cache, per_wl, flags, smats, numerics):
    """``Numerics(detuning=step)``: the reference's Rayleigh-point ladder (reference
    ``engine.py:6382-6483``). A point whose solve holds a grazing mode or failed is solved again
    with its wavelength times ``1 + step``; one still flagged also gets theta times ``1 + step``
    (theta = 0 has no second rung). The knobs multiply, each point keeps the first rung that
    clears it (or the last that solved), and the moved points are reported:
    ``AutoDetuneWarning``, the effective wavelengths as the result's wavelengths and the
    effective angles in its diagnostics; a condition that outlives the ladder warns
    ``RayleighPointWarning``. Returns ``(r, per_wl, flags, smats)`` with ``r`` carrying the
    effective wavelengths and a per-point ``thetas``."""
    import warnings
    from dataclasses import replace
    from .diagnostics import AutoDetuneWarning, RayleighPointWarning
    step = numerics.detuning
    fatal = _fatal_bits(numerics)
    if not step:
        return r, per_wl, flags, smats
    flags = np.array(flags)

    def retry(row):
        return bool(np.any(row & fatal) or np.any(_kept_grazing(row)))

    bad = [p for p in range(len(per_wl)) if retry(flags[p])]
    if not bad:
        return r, per_wl, flags, smats
    per_wl, smats = list(per_wl), (list(smats) if smats is not None else None)
    wl0 = np.array(r.wavelengths, dtype=np.float64)
    wls = wl0.copy()
    theta0 = r.source.theta
    thetas = [theta0 for _ in per_wl]
    knobs = ['wavelength'] + (['theta'] if theta0 not in (None, 0, 0.0) else [])
    wl_applied = set()                 # points whose wavelength rung took: the knob stays applied
    moved = {}
    for knob in knobs:
        if not bad:
            break
        trial_wl = np.array([wl0[p] * (1 + step) if (knob == 'wavelength' or p in wl_applied) else wl0[p]
                             for p in bad])
        trial_theta = theta0 * (1 + step) if knob == 'theta' else None
        sub = replace(r, wavelengths=trial_wl,
                      source=replace(r.source, theta=trial_theta) if trial_theta is not None else r.source)
        sub_per, sub_flags, sub_smats, _ = _solve_resolved(replace(sub, fields=False),
                                                           cache.sharing_conv())
        still = []
        for j, p in enumerate(bad):
            if np.any(sub_flags[j] & fatal):
                still.append(p)            # this rung did not take; the point keeps its last answer
                continue
            wls[p] = trial_wl[j]
            per_wl[p] = sub_per[j]
            flags[p] = sub_flags[j]
            if smats is not None:
                smats[p] = sub_smats[j]
            if trial_theta is not None:
                thetas[p] = trial_theta
            moved.setdefault(p, []).append(knob)
            if knob == 'wavelength':
                wl_applied.add(p)
            if retry(sub_flags[j]):
                still.append(p)
        bad = still
    if moved:
        where = ', '.join(f'{len(v)} point(s) by {"+".join(k)}' for k, v in
                          _group_moves(moved).items())
        warnings.warn(f'detuning={step:g} moved {where} off the requested point; the result '
                      'reports the effective wavelengths and angles.', AutoDetuneWarning, stacklevel=3)
    if bad:
        warnings.warn(f'{len(bad)} point(s) still hold a layer mode on its light line after the '
                      'detuning ladder.', RayleighPointWarning, stacklevel=3)
    return replace(r, wavelengths=wls, thetas=tuple(thetas)), per_wl, flags, smats


def _group_moves(moved):
    out = {}
    for p, knobs in moved.items():
        out.setdefault(tuple(knobs), []).append(p)
    return out


def _solve_resolved(r, cache):
    """One batched device solve over all wavelengths of a resolved problem."""
    import torch
    device = torch.device(cache.conv.device)
    _engine._native.context(device)          # raises EngineUnavailable without a GPU or library
    conv = cache.get(r.rasters, r.hu, r.hv, r.factorization, r.period)
    bucket = _engine.Bucket(hu=r.hu, hv=r.hv, path=r.path, n_layers=r.rasters.shape[0])
    for p, wl in enumerate(r.wavelengths):
        bucket.points.append(_engine.Point(
            stack=0, wavelength=float(wl), theta=r.thetas[p] if r.thetas else r.source.theta,
            phi=r.source.phi, polarization=r.polarization, eps_top=r.eps_top,
            eps_bottom=r.eps_bottom, period=r.period, thickness=tuple(float(t) for t in r.thickness),
            k_direct=r.k_direct))
    smats = None
    smat = None
    record = None
    if r.fields:
        eff, amp, pe, flags, state = _engine.run_fields(bucket, [conv], device)
        record = FieldRecord(device=device, hu=r.hu, hv=r.hv, mode=_engine._MODE[r.path],
                             thickness=r.thickness[::-1] if r.side == 'bottom' else r.thickness,
                             cell=_engine.cell_of(r.period), mirrored=r.side == 'bottom',
                             amplitude=_scalar(r.source.amplitude, 'excitation.sources.amplitude',
                                               real=False),
                             bare=r.bare, **state)
        if r.scattering:
            smat = _engine.run_bucket_scattering(bucket, [conv], device)[5].cpu().numpy()
    elif r.scattering:
        eff, amp, pe, flags, _, smat = _engine.run_bucket_scattering(bucket, [conv], device)
        smat = smat.cpu().numpy()
    elif _resident_modes(bucket):
        # the layer eigensystems stay on the device between runs: a run that changes only
        # thicknesses (or the polarization) takes them from the last run and skips the eigensolves
        mkey = (_Cache.key(r.rasters, r.hu, r.hv, r.factorization, r.period), r.path,
                _engine.modes_key(bucket))
        kept = cache.modes[1] if cache.modes is not None and cache.modes[0] == mkey else None
        outs, modes = _engine.run_window([bucket], [[conv]], device, modes=kept)
        cache.modes = (mkey, modes)
        eff, amp, pe, flags, _ = outs[0]
    else:
        eff, amp, pe, flags, _ = _engine.run_bucket(bucket, [conv], device, graph_slot=cache.graph)
    eff, amp, pe, flags = _engine._native.host_copy(eff, amp, pe, flags)
    _engine.rebuild_grazing(bucket, [conv], device, eff, amp, pe, flags, smat)
    if smat is not None:
        smats = _scattering_matrices(r, smat)
    per_wl = [_engine.unpack(eff, amp, pe, p, r.hu, r.hv, r.path) for p in range(len(bucket.points))]
    return per_wl, flags, smats, record


# Preparations whose fused solve needs more workspace than a captured graph is kept for, and whose
# modes take at most this many bytes, keep their layer eigensystems resident between runs.
RESIDENT_MODES_BYTES = 32 << 30


def _resident_modes(bucket):
    mode = _engine._MODE[bucket.path]
    P, L = len(bucket.points), bucket.n_layers
    N = (2 * bucket.hu + 1) * (2 * bucket.hv + 1)
    n = 2 * N if mode == _engine._native.MODE_2D else N
    from . import graphs
    if graphs.worth_capturing(P, L, bucket.hu, bucket.hv, mode):
        return False
    return P * L * (2 * n * n + n) * 16 <= RESIDENT_MODES_BYTES


def _scattering_matrices(r, smat):
    """One ``ScatteringMatrix`` per wavelength from the device blocks (P, 4, n, n), with the ports
    the reference attaches (``scattering.py:607-624``): the retained orders' kx/k0 (and ky/k0 on
    the coupled path), the top medium's index at the entry, the bottom medium's at the exit."""
    from .scattering import Port, ScatteringMatrix
    channel = {_engine.PATH_2D: 'sp', _engine.PATH_TE: 'te', _engine.PATH_TM: 'tm'}[r.path]
    n_top, n_bot = np.sqrt(complex(r.eps_top)), np.sqrt(complex(r.eps_bottom))
    out = []
    for p, wl in enumerate(r.wavelengths):
        kx0, ky0 = _engine.incident_k(r.eps_top, r.source.theta, r.source.phi, wl, r.k_direct)
        kx, ky = _engine.order_kxy(kx0, ky0, float(wl), r.period, r.hu, r.hv)
        ky = ky if channel == 'sp' else None
        S = smat[p]
        out.append(ScatteringMatrix(S[0].copy(), S[1].copy(), S[2].copy(), S[3].copy(),
                                    entry_port=Port(kx, ky, n_top, 1.0, channel),
                                    exit_port=Port(kx, ky, n_bot, 1.0, channel)))
    return out


def prepare(problem, *, numerics, backend='numpy', device='cpu', precision='double'):
    """Validate concrete values and freeze the structure a run reuses (reference :736)."""
    execution = Execution(backend=backend, device=device, precision=precision)
    try:
        r = _resolve(problem, numerics, execution)
        _engine.incident_k(r.eps_top, r.source.theta, r.source.phi, r.wavelengths[0], r.k_direct)
    except WavebitError:
        raise
    except NotImplementedError as error:
        raise UnsupportedFeatureError(str(error), code='rcwa.unsupported', path='problem') from error
    except (TypeError, ValueError) as error:
        raise ValidationError(str(error), code='rcwa.invalid', path='problem') from error
    for name, ok in _runtime_checks(r).items():
        if not ok:
            raise ValidationError(f'runtime guard {name} failed on the prepared problem.',
                                  code=f'prepared.{name}', path='problem')
    report = _report(problem, numerics, execution, r)
    contract = MappingProxyType({
        'static': ('numerics', 'execution', 'units', 'layer count', 'raster shapes',
                   'wavelength count', 'solve path', 'order window'),
        'live': ('raster values', 'half-space materials', 'wavelength values', 'incident angles',
                 'polarization', 'layer thickness'),
        'cache_policy': 'convolution matrices stay on the device per distinct raster stack; '
                        'modes and propagation are rebuilt every run; launch-bound sizes replay '
                        'a captured HIP graph of the solve'})
    return Prepared(numerics=numerics, execution=execution, report=report, contract=contract,
                    _structure=_structure_key(r), _cache=_Cache(device))


def run(problem, *, numerics, backend='numpy', device='cpu', precision='double'):
    """Prepare and run one problem (reference :800-803)."""
    return prepare(problem, numerics=numerics, backend=backend, device=device,
                   precision=precision).run(problem)
