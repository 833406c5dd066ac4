"""ctypes binding of the HIP engine's C ABI (``include/wavebit_amd.h``).

Device memory and streams come from PyTorch-ROCm (``torch.cuda`` is HIP on ROCm); every
buffer handed across the boundary is a contiguous tensor on the engine's device, complex128
as interleaved doubles. There is no fallback: if ``libwavebit_hip.so`` is missing or no GPU is
visible, these functions raise.
"""
import ctypes
import threading
from pathlib import Path

import numpy as np
import torch

_LIB_PATH = Path(__file__).resolve().parent / 'libwavebit_hip.so'
_lock = threading.Lock()
_lib = None
_ctxs = {}

WB_OK, WB_BADARG, WB_OOM, WB_SINGULAR, WB_NONFINITE, WB_NOCONVERGE, WB_FLAGGED, WB_HIP = range(8)
MODE_2D, MODE_1D_TE, MODE_1D_TM = 0, 1, 2
FLAG_NOCONVERGE, FLAG_SINGULAR, FLAG_EP, FLAG_GRAZING = 1, 2, 4, 8
FLAG_REBUILT, FLAG_DEFECTIVE, FLAG_EP_LOUD = 16, 32, 64
GRAZE_NMAX = 2048            # largest mode order wb_solve_rebuild takes
RAGGED_MAX = 64              # largest layer count wb_solve_from_modes_ragged takes
PT_STRIDE = 16

EXPORTS = ('wb_version', 'wb_ctx_create', 'wb_ctx_destroy', 'wb_last_error', 'wb_zgemm_batched',
           'wb_zgetrf_batched', 'wb_zgetrs_batched', 'wb_zgeev_workspace_size',
           'wb_zgeev_batched', 'wb_assemble_li_workspace_size', 'wb_assemble_li',
           'wb_solve_workspace_size', 'wb_solve', 'wb_zgemm_batched_ops', 'wb_zgetrs_batched_ops',
           'wb_solve_checkpoint_size', 'wb_solve_keep', 'wb_solve_adjoint_workspace_size',
           'wb_solve_adjoint', 'wb_assemble_li_adjoint_workspace_size', 'wb_assemble_li_adjoint',
           'wb_assemble_laurent_workspace_size', 'wb_assemble_laurent',
           'wb_assemble_laurent_adjoint_workspace_size', 'wb_assemble_laurent_adjoint',
           'wb_assemble_nvf_adjoint_workspace_size', 'wb_assemble_nvf_adjoint',
           'wb_solve_scattering', 'wb_timing_enable', 'wb_timing_read', 'wb_gemm_tuning_configs',
           'wb_gemm_tuning_config_name', 'wb_gemm_tuning_record', 'wb_gemm_autotune_size',
           'wb_gemm_autotune', 'wb_gemm_tuning_table', 'wb_gemm_tuning_load', 'wb_gemm_tuning_set',
           'wb_gemm_tuning_clear', 'wb_solve_noalloc', 'wb_solve_rebuild_workspace_size',
           'wb_solve_rebuild', 'wb_layer_modes_workspace_size', 'wb_layer_modes',
           'wb_solve_from_modes', 'wb_field_state_workspace_size', 'wb_field_state',
           'wb_field_sample_workspace_size', 'wb_field_sample', 'wb_flags_status',
           'wb_solve_ragged_workspace_size', 'wb_solve_from_modes_ragged',
           'wb_assemble_nvf_workspace_size', 'wb_assemble_nvf')

TIMER_IDS = {'zgemm': 0, 'getrf_panel': 1, 'laswp_trsm': 2, 'trsm_diag': 3, 'gehrd': 4,
             'unghr': 5, 'hqr': 6, 'trevc': 7, 'colnorm': 8, 'assemble': 9, 'stage_eig': 10,
             'stage_smat': 11}


class EngineUnavailable(RuntimeError):
    """The HIP engine cannot run here (library not built, or no GPU visible)."""


def library_path():
    return _LIB_PATH


def load_library():
    """Load ``libwavebit_hip.so`` and declare its signatures (no GPU needed)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not _LIB_PATH.exists():
            raise EngineUnavailable(
                f'{_LIB_PATH} is not built; run `python -m wavebit_amd.build` (hipcc, gfx950)')
        lib = ctypes.CDLL(str(_LIB_PATH))
        c_int, c_ll, c_sz, c_vp, c_dp = (ctypes.c_int, ctypes.c_longlong, ctypes.c_size_t,
                                         ctypes.c_void_p, ctypes.POINTER(ctypes.c_double))
        sig = {
            'wb_version': (c_int, []),
            'wb_ctx_create': (c_int, [c_int, ctypes.POINTER(c_vp)]),
            'wb_ctx_destroy': (c_int, [c_vp]),
            'wb_last_error': (ctypes.c_char_p, [c_vp]),
            'wb_zgemm_batched': (c_int, [c_vp, c_int, c_int, c_int, c_dp, c_vp, c_int, c_ll, c_vp,
                                         c_int, c_ll, c_dp, c_vp, c_int, c_ll, c_int, c_vp]),
            'wb_zgetrf_batched': (c_int, [c_vp, c_int, c_vp, c_int, c_ll, c_vp, c_vp, c_int, c_vp]),
            'wb_zgetrs_batched': (c_int, [c_vp, c_int, c_int, c_vp, c_int, c_ll, c_vp, c_vp, c_int,
                                          c_ll, c_int, c_vp]),
            'wb_zgeev_workspace_size': (c_sz, [c_int, c_int]),
            'wb_zgeev_batched': (c_int, [c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp, c_int,
                                         c_vp]),
            'wb_assemble_li_workspace_size': (c_sz, [c_int, c_int, c_int, c_int, c_int]),
            'wb_assemble_li': (c_int, [c_vp, c_vp, c_int, c_int, c_int, c_int, c_int, c_vp, c_vp,
                                       c_vp, c_vp, c_sz, c_vp, c_vp]),
            'wb_solve_workspace_size': (c_sz, [c_int, c_int, c_int, c_int, c_int]),
            'wb_solve': (c_int, [c_vp, c_int, c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp,
                                 c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_sz,
                                 c_vp]),
            'wb_assemble_nvf_workspace_size': (c_sz, [c_int, c_int, c_int, c_int, c_int]),
            'wb_assemble_nvf': (c_int, [c_vp, c_vp, c_int, c_int, c_int, c_int, c_int, c_dp, c_int,
                                        c_vp, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp, c_vp]),
            'wb_solve_rebuild_workspace_size': (c_sz, [c_int, c_int, c_int, c_int]),
            'wb_solve_rebuild': (c_int, [c_vp, c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp,
                                         c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                         c_sz, c_vp]),
            'wb_layer_modes_workspace_size': (c_sz, [c_int, c_int, c_int, c_int, c_int, c_int]),
            'wb_layer_modes': (c_int, [c_vp, c_int, c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp,
                                       c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_vp, c_sz,
                                       c_vp]),
            'wb_solve_from_modes': (c_int, [c_vp, c_int, c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp,
                                            c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                            c_sz, c_vp]),
            'wb_solve_ragged_workspace_size': (c_sz, [c_int, c_vp, c_int, c_int, c_int]),
            'wb_solve_from_modes_ragged': (c_int, [c_vp, c_int, c_vp, c_int, c_int, c_int, c_vp, c_vp,
                                                   c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                                   c_vp, c_sz, c_vp]),
            'wb_field_state_workspace_size': (c_sz, [c_int, c_int, c_int, c_int, c_int]),
            'wb_field_state': (c_int, [c_vp, c_int, c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp,
                                       c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_sz,
                                       c_vp]),
            'wb_field_sample_workspace_size': (c_sz, [c_int, c_int, c_int, c_int, c_int, c_int, c_int]),
            'wb_field_sample': (c_int, [c_vp, c_int, c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp,
                                        c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_vp, c_int, c_vp, c_dp,
                                        c_vp, c_vp, c_sz, c_vp]),
            'wb_flags_status': (c_int, [c_vp, c_ll, c_int]),
            'wb_solve_noalloc': (c_int, [c_vp, c_int, c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp,
                                         c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
            'wb_zgemm_batched_ops': (c_int, [c_vp, c_int, c_int, c_int, c_int, c_int, c_dp, c_vp,
                                             c_int, c_ll, c_vp, c_int, c_ll, c_dp, c_vp, c_int,
                                             c_ll, c_int, c_vp]),
            'wb_zgetrs_batched_ops': (c_int, [c_vp, c_int, c_int, c_int, c_vp, c_int, c_ll, c_vp,
                                              c_vp, c_int, c_ll, c_int, c_vp]),
            'wb_solve_checkpoint_size': (c_sz, [c_int, c_int, c_int, c_int, c_int]),
            'wb_solve_keep': (c_int, [c_vp, c_int, c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp,
                                      c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                      c_sz, c_vp, c_sz, c_vp]),
            'wb_solve_adjoint_workspace_size': (c_sz, [c_int, c_int, c_int, c_int, c_int]),
            'wb_solve_adjoint': (c_int, [c_vp, c_int, c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp,
                                         c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                         c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]),
            'wb_assemble_li_adjoint_workspace_size': (c_sz, [c_int, c_int, c_int, c_int, c_int]),
            'wb_assemble_li_adjoint': (c_int, [c_vp, c_vp, c_int, c_int, c_int, c_int, c_int, c_vp,
                                               c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]),
            'wb_assemble_laurent_workspace_size': (c_sz, [c_int, c_int, c_int, c_int, c_int]),
            'wb_assemble_laurent': (c_int, [c_vp, c_vp, c_int, c_int, c_int, c_int, c_int, c_vp,
                                            c_vp, c_vp, c_vp, c_sz, c_vp, c_vp]),
            'wb_assemble_laurent_adjoint_workspace_size': (c_sz, [c_int, c_int, c_int, c_int,
                                                                  c_int]),
            'wb_assemble_laurent_adjoint': (c_int, [c_vp, c_vp, c_int, c_int, c_int, c_int, c_int,
                                                    c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_sz,
                                                    c_vp]),
            'wb_assemble_nvf_adjoint_workspace_size': (c_sz, [c_int, c_int, c_int, c_int, c_int]),
            'wb_assemble_nvf_adjoint': (c_int, [c_vp, c_vp, c_int, c_int, c_int, c_int, c_int,
                                                c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_sz,
                                                c_vp]),
            'wb_solve_scattering': (c_int, [c_vp, c_int, c_int, c_int, c_int, c_int, c_vp, c_vp,
                                            c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                            c_vp, c_vp, c_vp, c_sz, c_vp]),
            'wb_timing_enable': (c_int, [c_int]),
            'wb_gemm_tuning_configs': (c_int, []),
            'wb_gemm_tuning_config_name': (ctypes.c_char_p, [c_int]),
            'wb_gemm_tuning_record': (c_int, [c_int]),
            'wb_gemm_autotune_size': (c_sz, []),
            'wb_gemm_autotune': (c_int, [c_vp, c_int, c_vp, c_sz, c_vp]),
            'wb_gemm_tuning_table': (c_sz, [ctypes.c_char_p, c_sz]),
            'wb_gemm_tuning_load': (c_int, [ctypes.c_char_p]),
            'wb_gemm_tuning_set': (c_int, [c_int, c_int, c_int, c_int, c_int, c_int, c_int]),
            'wb_gemm_tuning_clear': (None, []),
            'wb_timing_read': (c_int, [c_int, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(c_ll), ctypes.POINTER(ctypes.c_double),
                                       c_int]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


class NativeError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f'wavebit_amd status {status}: {message}')
        self.status = status


class Context:
    """One ``wb_ctx`` per device, created on first use and kept for the process."""

    def __init__(self, device):
        self.lib = load_library()
        self.device = device
        handle = ctypes.c_void_p()
        status = self.lib.wb_ctx_create(device, ctypes.byref(handle))
        if status != WB_OK:
            raise EngineUnavailable(f'wb_ctx_create({device}) failed with status {status}')
        self.handle = handle

    def check(self, status, what):
        if status != WB_OK:
            msg = self.lib.wb_last_error(self.handle)
            raise NativeError(status, f'{what}: {msg.decode() if msg else ""}')


def device_index(device):
    device = torch.device(device)
    if device.type != 'cuda':
        raise EngineUnavailable(f'the HIP engine runs on a GPU device, got {device}')
    return torch.cuda.current_device() if device.index is None else device.index


def context(device):
    if not torch.cuda.is_available():
        raise EngineUnavailable('no GPU is visible to PyTorch-ROCm; the HIP engine cannot run')
    idx = device_index(device)
    with _lock:
        ctx = _ctxs.get(idx)
    if ctx is None:
        ctx = Context(idx)
        with _lock:
            _ctxs.setdefault(idx, ctx)
            ctx = _ctxs[idx]
    return ctx


def _stream(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _require(t, dtype, device, name):
    if t.dtype != dtype or t.device != device or not t.is_contiguous():
        raise ValueError(f'{name}: expected a contiguous {dtype} tensor on {device}, got '
                         f'{t.dtype} on {t.device} (contiguous={t.is_contiguous()})')


def _offdiag_ok(epxy, epx):
    if epxy is not None:
        _require(epxy, torch.complex128, epx.device, 'epxy')
        if epxy.shape != epx.shape:
            raise ValueError(f'epxy: expected the shape of epx {tuple(epx.shape)}, got '
                             f'{tuple(epxy.shape)}')


def _workspace(nbytes, device):
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


# ----------------------------------------------------------------------------- primitives

def zgemm(A, B, C=None, alpha=1.0, beta=0.0):
    """Batched ``alpha A @ B + beta C`` for (batch, m, k) x (batch, k, n) complex128."""
    device = A.device
    ctx = context(device)
    _require(A, torch.complex128, device, 'A')
    _require(B, torch.complex128, device, 'B')
    batch, m, k = A.shape
    n = B.shape[-1]
    if B.shape != (batch, k, n):
        raise ValueError(f'shape mismatch {tuple(A.shape)} @ {tuple(B.shape)}')
    if C is None:
        C = torch.zeros((batch, m, n), dtype=torch.complex128, device=device)
        beta = 0.0
    _require(C, torch.complex128, device, 'C')
    al = (ctypes.c_double * 2)(complex(alpha).real, complex(alpha).imag)
    be = (ctypes.c_double * 2)(complex(beta).real, complex(beta).imag)
    ctx.check(ctx.lib.wb_zgemm_batched(ctx.handle, m, n, k, al, _ptr(A), k, m * k, _ptr(B), n,
                                       k * n, be, _ptr(C), n, m * n, batch, _stream(device)),
              'zgemm')
    return C


def zgetrf(A):
    """In-place batched LU with partial pivoting: returns (A, ipiv, info)."""
    device = A.device
    ctx = context(device)
    _require(A, torch.complex128, device, 'A')
    batch, n, _ = A.shape
    ipiv = torch.empty((batch, n), dtype=torch.int32, device=device)
    info = torch.empty((batch,), dtype=torch.int32, device=device)
    ctx.check(ctx.lib.wb_zgetrf_batched(ctx.handle, n, _ptr(A), n, n * n, _ptr(ipiv), _ptr(info),
                                        batch, _stream(device)), 'zgetrf')
    return A, ipiv, info


def zgetrs(LU, ipiv, B):
    """Solve ``A X = B`` in place in B (batch, n, nrhs) from ``zgetrf`` factors."""
    device = LU.device
    ctx = context(device)
    _require(B, torch.complex128, device, 'B')
    batch, n, _ = LU.shape
    nrhs = B.shape[-1]
    ctx.check(ctx.lib.wb_zgetrs_batched(ctx.handle, n, nrhs, _ptr(LU), n, n * n, _ptr(ipiv),
                                        _ptr(B), nrhs, n * nrhs, batch, _stream(device)), 'zgetrs')
    return B


def zgeev(A):
    """Batched general eigendecomposition; A (batch, n, n) is consumed. Returns (w, W, info)."""
    device = A.device
    ctx = context(device)
    _require(A, torch.complex128, device, 'A')
    batch, n, _ = A.shape
    w = torch.empty((batch, n), dtype=torch.complex128, device=device)
    W = torch.empty((batch, n, n), dtype=torch.complex128, device=device)
    info = torch.empty((batch,), dtype=torch.int32, device=device)
    nbytes = ctx.lib.wb_zgeev_workspace_size(n, batch)
    ws = _workspace(nbytes, device)
    ctx.check(ctx.lib.wb_zgeev_batched(ctx.handle, n, _ptr(A), _ptr(w), _ptr(W), _ptr(ws), nbytes,
                                       _ptr(info), batch, _stream(device)), 'zgeev')
    return w, W, info


def assemble_li(eps, hu, hv):
    """Li convolution matrices of a raster stack eps (L, Ny, Nx) -> (epx, epy, epz_i, info)."""
    device = eps.device
    ctx = context(device)
    _require(eps, torch.complex128, device, 'eps')
    L, Ny, Nx = eps.shape
    N = (2 * hu + 1) * (2 * hv + 1)
    out = [torch.empty((L, N, N), dtype=torch.complex128, device=device) for _ in range(3)]
    info = torch.empty((2 * L,), dtype=torch.int32, device=device)
    nbytes = ctx.lib.wb_assemble_li_workspace_size(L, Ny, Nx, hu, hv)
    ws = _workspace(nbytes, device)
    ctx.check(ctx.lib.wb_assemble_li(ctx.handle, _ptr(eps), L, Ny, Nx, hu, hv, _ptr(out[0]),
                                     _ptr(out[1]), _ptr(out[2]), _ptr(ws), nbytes, _ptr(info),
                                     _stream(device)), 'assemble_li')
    return out[0], out[1], out[2], info


def assemble_laurent(eps, hu, hv):
    """Laurent convolution matrices of a raster stack eps (L, Ny, Nx) -> (epx, epy, epz_i, info):
    epx and epy are the same direct-rule matrix, epz_i its inverse; info holds L ints."""
    device = eps.device
    ctx = context(device)
    _require(eps, torch.complex128, device, 'eps')
    L, Ny, Nx = eps.shape
    N = (2 * hu + 1) * (2 * hv + 1)
    out = [torch.empty((L, N, N), dtype=torch.complex128, device=device) for _ in range(3)]
    info = torch.empty((L,), dtype=torch.int32, device=device)
    nbytes = ctx.lib.wb_assemble_laurent_workspace_size(L, Ny, Nx, hu, hv)
    ws = _workspace(nbytes, device)
    ctx.check(ctx.lib.wb_assemble_laurent(ctx.handle, _ptr(eps), L, Ny, Nx, hu, hv, _ptr(out[0]),
                                          _ptr(out[1]), _ptr(out[2]), _ptr(ws), nbytes, _ptr(info),
                                          _stream(device)), 'assemble_laurent')
    return out[0], out[1], out[2], info


def _cell(lattice):
    """The cell basis (2 x 2, columns a1 and a2) as the C-ABI's four doubles, or None."""
    if lattice is None:
        return None
    A = np.asarray(lattice, dtype=np.float64).reshape(2, 2)
    return (ctypes.c_double * 4)(A[0, 0], A[1, 0], A[0, 1], A[1, 1])


def assemble_nvf(eps, hu, hv, lattice=None, floor=True):
    """NVF convolution matrices of a raster stack eps (L, Ny, Nx) -> (epx, epy, epz_i, epxy,
    info): epxy is the solve's epxy and epyx slot alike; ``lattice`` the cell basis (2 x 2,
    columns a1 and a2) or None for square pixels; ``floor``: floor [[1/eps]]'s singular values on
    the layers that are not passive dielectrics (False skips the work for an all-passive stack);
    info holds 3 L ints (the last L: layers that are not passive dielectrics)."""
    device = eps.device
    ctx = context(device)
    _require(eps, torch.complex128, device, 'eps')
    L, Ny, Nx = eps.shape
    N = (2 * hu + 1) * (2 * hv + 1)
    out = [torch.empty((L, N, N), dtype=torch.complex128, device=device) for _ in range(4)]
    info = torch.empty((3 * L,), dtype=torch.int32, device=device)
    nbytes = ctx.lib.wb_assemble_nvf_workspace_size(L, Ny, Nx, hu, hv)
    ws = _workspace(nbytes, device)
    ctx.check(ctx.lib.wb_assemble_nvf(ctx.handle, _ptr(eps), L, Ny, Nx, hu, hv, _cell(lattice),
                                      int(bool(floor)),
                                      _ptr(out[0]),
                                      _ptr(out[1]), _ptr(out[2]), _ptr(out[3]), _ptr(ws), nbytes,
                                      _ptr(info), _stream(device)), 'assemble_nvf')
    return out[0], out[1], out[2], out[3], info


def assemble(eps, hu, hv, factorization='li', lattice=None, floor=True):
    """The convolution matrices of ``factorization`` ('li', 'laurent' or 'nvf') for a raster
    stack: (epx, epy, epz_i, epxy, info), epxy None for the diagonal rules."""
    if factorization == 'laurent':
        return _with_offdiag(assemble_laurent(eps, hu, hv))
    if factorization == 'li':
        return _with_offdiag(assemble_li(eps, hu, hv))
    if factorization == 'nvf':
        return assemble_nvf(eps, hu, hv, lattice, floor)
    raise ValueError(f'no assembly for factorization={factorization!r}')


def _with_offdiag(res):
    epx, epy, epz_i, info = res
    return epx, epy, epz_i, None, info


def solve_workspace_bytes(P, L, hu, hv, mode):
    return int(load_library().wb_solve_workspace_size(P, L, hu, hv, mode))


def solve_ragged_workspace_bytes(layers, hu, hv, mode):
    """Workspace of ``solve_from_modes_ragged`` for points with these layer counts (any order)."""
    lay = np.ascontiguousarray(sorted(layers, reverse=True), dtype=np.int32)
    return int(load_library().wb_solve_ragged_workspace_size(
        len(lay), lay.ctypes.data_as(ctypes.c_void_p), hu, hv, mode))


def _packed_outputs(P, L, N, device):
    """eff (P, 3, 4, N), amp (P, 3, 4, N) complex, pe (P, 3) and flags (P, L) int32 as views of
    one device buffer, so that ``host_copy`` brings a solve's results back in a single copy."""
    nf = (P * L + 1) // 2
    buf = torch.empty(36 * P * N + 3 * P + nf, dtype=torch.float64, device=device)
    return packed_views(buf, P, L, N)


def packed_views(buf, P, L, N):
    """(eff, amp, pe, flags) views of a packed output buffer (see ``_packed_outputs``)."""
    na, ne = 24 * P * N, 12 * P * N
    amp = buf[:na].view(torch.complex128).view(P, 3, 4, N)
    eff = buf[na:na + ne].view(P, 3, 4, N)
    pe = buf[na + ne:na + ne + 3 * P].view(P, 3)
    flags = buf[na + ne + 3 * P:].view(torch.int32)[:P * L].view(P, L)
    eff._wb_packed = (buf, amp, pe, flags)
    return eff, amp, pe, flags


def host_copy(eff, amp, pe, flags):
    """Host numpy arrays of a solve's (eff, amp, pe, flags): one device-to-host copy when the
    four are ``solve``'s packed views, else one copy each."""
    packed = getattr(eff, '_wb_packed', None)
    if packed is None or packed[1] is not amp or packed[2] is not pe or packed[3] is not flags:
        return tuple(t.cpu().numpy() for t in (eff, amp, pe, flags))
    buf = packed[0]
    h = buf.cpu()
    P, _, _, N = eff.shape
    L = flags.shape[1]
    na, ne = 24 * P * N, 12 * P * N
    return (h[na:na + ne].view(P, 3, 4, N).numpy(),
            h[:na].view(torch.complex128).view(P, 3, 4, N).numpy(),
            h[na + ne:na + ne + 3 * P].view(P, 3).numpy(),
            h[na + ne + 3 * P:].view(torch.int32)[:P * L].view(P, L).numpy())


def solve(pts, thick, conv_idx, epx, epy, epz_i, hu, hv, mode, workspace=None, epxy=None):
    """Batch RCWA solve; see ``wb_solve``. ``epxy``: the NVF off-diagonal slot (epxy = epyx),
    or None. Returns (eff, amp, pe, flags)."""
    device = epx.device
    ctx = context(device)
    P, L = thick.shape
    N = (2 * hu + 1) * (2 * hv + 1)
    _require(pts, torch.float64, device, 'pts')
    _require(thick, torch.float64, device, 'thick')
    _require(conv_idx, torch.int32, device, 'conv_idx')
    for name, t in (('epx', epx), ('epy', epy), ('epz_i', epz_i)):
        _require(t, torch.complex128, device, name)
    eff, amp, pe, flags = _packed_outputs(P, L, N, device)
    nbytes = ctx.lib.wb_solve_workspace_size(P, L, hu, hv, mode)
    if workspace is None or workspace.numel() < nbytes:
        workspace = _workspace(nbytes, device)
    _offdiag_ok(epxy, epx)
    ctx.check(ctx.lib.wb_solve(ctx.handle, P, L, hu, hv, mode, _ptr(pts), _ptr(thick),
                               _ptr(conv_idx), _ptr(epx), _ptr(epy), _ptr(epz_i), _ptr(epxy),
                               _ptr(epxy), _ptr(eff),
                               _ptr(amp), _ptr(pe), _ptr(flags), _ptr(workspace), nbytes,
                               _stream(device)), 'solve')
    return eff, amp, pe, flags


def solve_rebuild(pts, thick, conv_idx, epx, epy, epz_i, hu, hv, scattering=False, epxy=None):
    """``solve`` of 2D points with every grazing layer's mode cluster rebuilt; see
    ``wb_solve_rebuild``. Returns (eff, amp, pe, flags) and, with ``scattering``, smat."""
    device = epx.device
    ctx = context(device)
    P, L = thick.shape
    N = (2 * hu + 1) * (2 * hv + 1)
    n = 2 * N
    _require(pts, torch.float64, device, 'pts')
    _require(thick, torch.float64, device, 'thick')
    _require(conv_idx, torch.int32, device, 'conv_idx')
    for name, t in (('epx', epx), ('epy', epy), ('epz_i', epz_i)):
        _require(t, torch.complex128, device, name)
    eff = torch.empty((P, 3, 4, N), dtype=torch.float64, device=device)
    amp = torch.empty((P, 3, 4, N), dtype=torch.complex128, device=device)
    pe = torch.empty((P, 3), dtype=torch.float64, device=device)
    flags = torch.empty((P, L), dtype=torch.int32, device=device)
    smat = torch.empty((P, 4, n, n), dtype=torch.complex128, device=device) if scattering else None
    nbytes = ctx.lib.wb_solve_rebuild_workspace_size(P, L, hu, hv)
    ws = _workspace(nbytes, device)
    _offdiag_ok(epxy, epx)
    ctx.check(ctx.lib.wb_solve_rebuild(ctx.handle, P, L, hu, hv, _ptr(pts), _ptr(thick),
                                       _ptr(conv_idx), _ptr(epx), _ptr(epy), _ptr(epz_i),
                                       _ptr(epxy), _ptr(epxy),
                                       _ptr(eff), _ptr(amp), _ptr(pe), _ptr(flags), _ptr(smat),
                                       _ptr(ws), nbytes, _stream(device)), 'solve_rebuild')
    return (eff, amp, pe, flags, smat) if scattering else (eff, amp, pe, flags)


def layer_modes_bytes(P, L, hu, hv, mode, rebuild=False):
    return int(load_library().wb_layer_modes_workspace_size(P, L, hu, hv, mode, int(bool(rebuild))))


def layer_modes(pts, conv_idx, epx, epy, epz_i, hu, hv, mode, rebuild=False, workspace=None,
                epxy=None):
    """The raw layer modes of every (point, layer) pair; see ``wb_layer_modes``. ``conv_idx`` is
    (P, L). Returns (W, V, q, flags): (P L, n, n), (P L, n, n), (P L, n), (P, L)."""
    device = epx.device
    ctx = context(device)
    P, L = conv_idx.shape
    N = (2 * hu + 1) * (2 * hv + 1)
    n = 2 * N if mode == MODE_2D else N
    _require(pts, torch.float64, device, 'pts')
    _require(conv_idx, torch.int32, device, 'conv_idx')
    for name, t in (('epx', epx), ('epy', epy), ('epz_i', epz_i)):
        _require(t, torch.complex128, device, name)
    W = torch.empty((P * L, n, n), dtype=torch.complex128, device=device)
    V = torch.empty((P * L, n, n), dtype=torch.complex128, device=device)
    q = torch.empty((P * L, n), dtype=torch.complex128, device=device)
    flags = torch.empty((P, L), dtype=torch.int32, device=device)
    nbytes = ctx.lib.wb_layer_modes_workspace_size(P, L, hu, hv, mode, int(bool(rebuild)))
    if workspace is None or workspace.numel() < nbytes:
        workspace = _workspace(nbytes, device)
    _offdiag_ok(epxy, epx)
    ctx.check(ctx.lib.wb_layer_modes(ctx.handle, P, L, hu, hv, mode, _ptr(pts), _ptr(conv_idx),
                                     _ptr(epx), _ptr(epy), _ptr(epz_i), _ptr(epxy), _ptr(epxy),
                                     _ptr(W), _ptr(V), _ptr(q),
                                     _ptr(flags), int(bool(rebuild)), _ptr(workspace), nbytes,
                                     _stream(device)), 'layer_modes')
    return W, V, q, flags


def solve_from_modes(pts, thick, W, V, q, modes_idx, hu, hv, mode, scattering=False, workspace=None):
    """The S-matrix half of ``solve`` from modes of ``layer_modes``; see ``wb_solve_from_modes``.
    ``modes_idx`` (P, L) int32 or None. Returns (eff, amp, pe, flags) and, with ``scattering``,
    smat; flags report this call's factorizations only."""
    device = W.device
    ctx = context(device)
    P, L = thick.shape
    N = (2 * hu + 1) * (2 * hv + 1)
    n = 2 * N if mode == MODE_2D else N
    _require(pts, torch.float64, device, 'pts')
    _require(thick, torch.float64, device, 'thick')
    for name, t in (('W', W), ('V', V), ('q', q)):
        _require(t, torch.complex128, device, name)
    if modes_idx is not None:
        _require(modes_idx, torch.int32, device, 'modes_idx')
    eff = torch.empty((P, 3, 4, N), dtype=torch.float64, device=device)
    amp = torch.empty((P, 3, 4, N), dtype=torch.complex128, device=device)
    pe = torch.empty((P, 3), dtype=torch.float64, device=device)
    flags = torch.empty((P, L), dtype=torch.int32, device=device)
    smat = torch.empty((P, 4, n, n), dtype=torch.complex128, device=device) if scattering else None
    nbytes = ctx.lib.wb_solve_workspace_size(P, L, hu, hv, mode)
    if workspace is None or workspace.numel() < nbytes:
        workspace = _workspace(nbytes, device)
    ctx.check(ctx.lib.wb_solve_from_modes(ctx.handle, P, L, hu, hv, mode, _ptr(pts), _ptr(thick),
                                          _ptr(W), _ptr(V), _ptr(q), _ptr(modes_idx), _ptr(eff),
                                          _ptr(amp), _ptr(pe), _ptr(flags), _ptr(smat),
                                          _ptr(workspace), nbytes, _stream(device)), 'solve_from_modes')
    return (eff, amp, pe, flags, smat) if scattering else (eff, amp, pe, flags)


def solve_from_modes_ragged(pts, thick, W, V, q, modes_idx, layers, hu, hv, mode, workspace=None):
    """``solve_from_modes`` over points whose layer counts differ, as one cascade; see
    ``wb_solve_from_modes_ragged``. ``layers`` (host ints, non-increasing) are the points' layer
    counts; ``thick`` (B,) float64 and ``modes_idx`` (B,) int32 or None are numbered layer-major
    (layer l of point p at base_l + p). Returns (eff, amp, pe, flags), flags (B,) in the same
    numbering and reporting this call's factorizations only."""
    device = W.device
    ctx = context(device)
    lay = np.ascontiguousarray(layers, dtype=np.int32)
    P = int(lay.shape[0])
    B = int(lay.sum())
    N = (2 * hu + 1) * (2 * hv + 1)
    _require(pts, torch.float64, device, 'pts')
    _require(thick, torch.float64, device, 'thick')
    for name, t in (('W', W), ('V', V), ('q', q)):
        _require(t, torch.complex128, device, name)
    if modes_idx is not None:
        _require(modes_idx, torch.int32, device, 'modes_idx')
    if tuple(thick.shape) != (B,) or (modes_idx is not None and tuple(modes_idx.shape) != (B,)):
        raise ValueError(f'thick and modes_idx: expected ({B},) in the layer-major numbering')
    lay_p = lay.ctypes.data_as(ctypes.c_void_p)
    eff = torch.empty((P, 3, 4, N), dtype=torch.float64, device=device)
    amp = torch.empty((P, 3, 4, N), dtype=torch.complex128, device=device)
    pe = torch.empty((P, 3), dtype=torch.float64, device=device)
    flags = torch.empty((B,), dtype=torch.int32, device=device)
    nbytes = ctx.lib.wb_solve_ragged_workspace_size(P, lay_p, hu, hv, mode)
    if nbytes == 0:
        raise ValueError('layers: expected non-increasing layer counts in 1..64')
    if workspace is None or workspace.numel() < nbytes:
        workspace = _workspace(nbytes, device)
    ctx.check(ctx.lib.wb_solve_from_modes_ragged(ctx.handle, P, lay_p, hu, hv, mode, _ptr(pts),
                                                 _ptr(thick), _ptr(W), _ptr(V), _ptr(q),
                                                 _ptr(modes_idx), _ptr(eff), _ptr(amp), _ptr(pe),
                                                 _ptr(flags), _ptr(workspace), nbytes,
                                                 _stream(device)), 'solve_from_modes_ragged')
    return eff, amp, pe, flags


def ragged_order(layers):
    """The layer-major numbering of ``wb_solve_from_modes_ragged`` for layer counts in any order:
    (order, pair) with ``order`` the points sorted by non-increasing layer count (stable) and
    ``pair[k]`` the (sorted point, layer) of pair k, as two int arrays."""
    lay = np.asarray(layers, dtype=np.int64)
    order = np.argsort(-lay, kind='stable')
    sl = lay[order]
    pts_, lays_ = [], []
    for l in range(int(sl.max()) if sl.size else 0):
        cnt = int(np.count_nonzero(sl > l))
        pts_.append(np.arange(cnt))
        lays_.append(np.full(cnt, l))
    if not pts_:
        return order, (np.zeros(0, np.int64), np.zeros(0, np.int64))
    return order, (np.concatenate(pts_), np.concatenate(lays_))


def field_state(pts, thick, W, V, q, modes_idx, hu, hv, mode, workspace=None):
    """``solve_from_modes`` plus each pair's mode amplitudes for the point's own incidence; see
    ``wb_field_state``. Returns (amps, eff, amp, pe, flags), amps (P L, 2, n)."""
    device = W.device
    ctx = context(device)
    P, L = thick.shape
    N = (2 * hu + 1) * (2 * hv + 1)
    n = 2 * N if mode == MODE_2D else N
    _require(pts, torch.float64, device, 'pts')
    _require(thick, torch.float64, device, 'thick')
    for name, t in (('W', W), ('V', V), ('q', q)):
        _require(t, torch.complex128, device, name)
    if modes_idx is not None:
        _require(modes_idx, torch.int32, device, 'modes_idx')
    amps = torch.empty((P * L, 2, n), dtype=torch.complex128, device=device)
    eff = torch.empty((P, 3, 4, N), dtype=torch.float64, device=device)
    amp = torch.empty((P, 3, 4, N), dtype=torch.complex128, device=device)
    pe = torch.empty((P, 3), dtype=torch.float64, device=device)
    flags = torch.empty((P, L), dtype=torch.int32, device=device)
    nbytes = ctx.lib.wb_field_state_workspace_size(P, L, hu, hv, mode)
    if workspace is None or workspace.numel() < nbytes:
        workspace = _workspace(nbytes, device)
    ctx.check(ctx.lib.wb_field_state(ctx.handle, P, L, hu, hv, mode, _ptr(pts), _ptr(thick), _ptr(W),
                                     _ptr(V), _ptr(q), _ptr(modes_idx), _ptr(amps), _ptr(eff),
                                     _ptr(amp), _ptr(pe), _ptr(flags), _ptr(workspace), nbytes,
                                     _stream(device)), 'field_state')
    return amps, eff, amp, pe, flags


def field_sample_bytes(P, L, hu, hv, mode, nz, nxy):
    return int(load_library().wb_field_sample_workspace_size(P, L, hu, hv, mode, nz, nxy))


def field_sample(pts, thick, W, V, q, modes_idx, epz_i, conv_idx, amps, z, xy, hu, hv, mode, scale,
                 workspace=None):
    """Fields of P points x L layers; see ``wb_field_sample``. ``pts`` (P, 16), ``thick`` (P, L),
    ``W``/``V`` (pairs, n, n), ``q`` (pairs, n), ``modes_idx`` / ``conv_idx`` (P, L) int32 or None,
    ``epz_i`` (C, N, N) or None on the TE path, ``amps`` (P L, 2, n), ``z`` (L, nz) local heights,
    ``xy`` (nxy, 2); ``scale`` six complex factors. Returns (P, L nz, nxy, 6)."""
    device = W.device
    ctx = context(device)
    P, L = thick.shape
    _require(pts, torch.float64, device, 'pts')
    _require(thick, torch.float64, device, 'thick')
    _require(z, torch.float64, device, 'z')
    _require(xy, torch.float64, device, 'xy')
    for name, t in (('W', W), ('V', V), ('q', q), ('amps', amps)):
        _require(t, torch.complex128, device, name)
    if epz_i is not None:
        _require(epz_i, torch.complex128, device, 'epz_i')
    for name, t in (('modes_idx', modes_idx), ('conv_idx', conv_idx)):
        if t is not None:
            _require(t, torch.int32, device, name)
    nz, nxy = int(z.shape[1]), int(xy.shape[0])
    out = torch.empty((P, L * nz, nxy, 6), dtype=torch.complex128, device=device)
    sc = (ctypes.c_double * 12)(*[v for c in scale for v in (complex(c).real, complex(c).imag)])
    nbytes = ctx.lib.wb_field_sample_workspace_size(P, L, hu, hv, mode, nz, nxy)
    if workspace is None or workspace.numel() < nbytes:
        workspace = _workspace(nbytes, device)
    ctx.check(ctx.lib.wb_field_sample(ctx.handle, P, L, hu, hv, mode, _ptr(pts), _ptr(thick), _ptr(W),
                                      _ptr(V), _ptr(q), _ptr(modes_idx), _ptr(epz_i), _ptr(conv_idx),
                                      _ptr(amps), nz, _ptr(z), nxy, _ptr(xy), sc, _ptr(out),
                                      _ptr(workspace), nbytes, _stream(device)), 'field_sample')
    return out


STATUS_NAMES = {0: 'WB_OK', 1: 'WB_BADARG', 2: 'WB_OOM', 3: 'WB_SINGULAR', 4: 'WB_NONFINITE',
                5: 'WB_NOCONVERGE', 6: 'WB_FLAGGED', 7: 'WB_HIP', 8: 'WB_UNSUPPORTED'}


def flags_status(flags, notice_mask=0):
    """``wb_flags_status`` of a host array of flags: the batch-level status name."""
    import numpy as np
    a = np.ascontiguousarray(np.asarray(flags, dtype=np.int32)).reshape(-1)
    return STATUS_NAMES[int(load_library().wb_flags_status(a.ctypes.data_as(ctypes.c_void_p), a.size,
                                                             int(notice_mask)))]


SOLVE_NOALLOC_BYTES = 64 << 20   # WB_SOLVE_NOALLOC_BYTES


def solve_noalloc(pts, thick, conv_idx, epx, epy, epz_i, hu, hv, mode, out=None):
    """``solve`` through ``wb_solve_noalloc``: the engine runs on its own fixed device array, and
    with ``out`` (the ``(eff, amp, pe, flags)`` tensors of an earlier call of the same shapes)
    nothing is allocated on the device for the call. Returns (eff, amp, pe, flags)."""
    device = epx.device
    ctx = context(device)
    P, L = thick.shape
    N = (2 * hu + 1) * (2 * hv + 1)
    _require(pts, torch.float64, device, 'pts')
    _require(thick, torch.float64, device, 'thick')
    _require(conv_idx, torch.int32, device, 'conv_idx')
    for name, t in (('epx', epx), ('epy', epy), ('epz_i', epz_i)):
        _require(t, torch.complex128, device, name)
    if out is None:
        out = (torch.empty((P, 3, 4, N), dtype=torch.float64, device=device),
               torch.empty((P, 3, 4, N), dtype=torch.complex128, device=device),
               torch.empty((P, 3), dtype=torch.float64, device=device),
               torch.empty((P, L), dtype=torch.int32, device=device))
    eff, amp, pe, flags = out
    for name, t, shape in (('eff', eff, (P, 3, 4, N)), ('amp', amp, (P, 3, 4, N)),
                           ('pe', pe, (P, 3)), ('flags', flags, (P, L))):
        if tuple(t.shape) != shape:
            raise ValueError(f'{name}: expected shape {shape}, got {tuple(t.shape)}')
    ctx.check(ctx.lib.wb_solve_noalloc(ctx.handle, P, L, hu, hv, mode, _ptr(pts), _ptr(thick),
                                       _ptr(conv_idx), _ptr(epx), _ptr(epy), _ptr(epz_i),
                                       _ptr(eff), _ptr(amp), _ptr(pe), _ptr(flags),
                                       _stream(device)), 'solve_noalloc')
    return eff, amp, pe, flags


def solve_scattering(pts, thick, conv_idx, epx, epy, epz_i, hu, hv, mode, epxy=None):
    """``solve`` plus the stack's scattering matrix; see ``wb_solve_scattering``.
    Returns (eff, amp, pe, flags, smat) with smat (P, 4, n, n): S11, S12, S21, S22."""
    device = epx.device
    ctx = context(device)
    P, L = thick.shape
    N = (2 * hu + 1) * (2 * hv + 1)
    n = 2 * N if mode == MODE_2D else N
    _require(pts, torch.float64, device, 'pts')
    _require(thick, torch.float64, device, 'thick')
    _require(conv_idx, torch.int32, device, 'conv_idx')
    for name, t in (('epx', epx), ('epy', epy), ('epz_i', epz_i)):
        _require(t, torch.complex128, device, name)
    eff = torch.empty((P, 3, 4, N), dtype=torch.float64, device=device)
    amp = torch.empty((P, 3, 4, N), dtype=torch.complex128, device=device)
    pe = torch.empty((P, 3), dtype=torch.float64, device=device)
    flags = torch.empty((P, L), dtype=torch.int32, device=device)
    smat = torch.empty((P, 4, n, n), dtype=torch.complex128, device=device)
    nbytes = ctx.lib.wb_solve_workspace_size(P, L, hu, hv, mode)
    ws = _workspace(nbytes, device)
    _offdiag_ok(epxy, epx)
    ctx.check(ctx.lib.wb_solve_scattering(ctx.handle, P, L, hu, hv, mode, _ptr(pts), _ptr(thick),
                                          _ptr(conv_idx), _ptr(epx), _ptr(epy), _ptr(epz_i),
                                          _ptr(epxy), _ptr(epxy),
                                          _ptr(eff), _ptr(amp), _ptr(pe), _ptr(flags), _ptr(smat),
                                          _ptr(ws), nbytes, _stream(device)), 'solve_scattering')
    return eff, amp, pe, flags, smat


def zgemm_ops(A, B, opa=0, opb=0, C=None, alpha=1.0, beta=0.0):
    """Batched ``alpha op(A) @ op(B) + beta C``, op 1 = conjugate transpose of the stored
    (batch, rows, cols) operand."""
    device = A.device
    ctx = context(device)
    _require(A, torch.complex128, device, 'A')
    _require(B, torch.complex128, device, 'B')
    batch = A.shape[0]
    m, k = (A.shape[2], A.shape[1]) if opa else (A.shape[1], A.shape[2])
    kb, n = (B.shape[2], B.shape[1]) if opb else (B.shape[1], B.shape[2])
    if kb != k or B.shape[0] != batch:
        raise ValueError(f'shape mismatch op{opa}{tuple(A.shape)} @ op{opb}{tuple(B.shape)}')
    if C is None:
        C = torch.zeros((batch, m, n), dtype=torch.complex128, device=device)
        beta = 0.0
    _require(C, torch.complex128, device, 'C')
    al = (ctypes.c_double * 2)(complex(alpha).real, complex(alpha).imag)
    be = (ctypes.c_double * 2)(complex(beta).real, complex(beta).imag)
    ctx.check(ctx.lib.wb_zgemm_batched_ops(ctx.handle, opa, opb, m, n, k, al, _ptr(A), A.shape[2],
                                           A.shape[1] * A.shape[2], _ptr(B), B.shape[2],
                                           B.shape[1] * B.shape[2], be, _ptr(C), n, m * n, batch,
                                           _stream(device)), 'zgemm_ops')
    return C


def zgetrs_ops(LU, ipiv, B, op=0):
    """Solve ``A X = B`` (op 0) or ``A^H X = B`` (op 1) in place from ``zgetrf`` factors."""
    device = LU.device
    ctx = context(device)
    _require(B, torch.complex128, device, 'B')
    batch, n, _ = LU.shape
    nrhs = B.shape[-1]
    ctx.check(ctx.lib.wb_zgetrs_batched_ops(ctx.handle, op, n, nrhs, _ptr(LU), n, n * n, _ptr(ipiv),
                                            _ptr(B), nrhs, n * nrhs, batch, _stream(device)),
              'zgetrs_ops')
    return B


def checkpoint_bytes(P, L, hu, hv, mode=MODE_2D):
    return int(load_library().wb_solve_checkpoint_size(P, L, hu, hv, mode))


def solve_keep(pts, thick, conv_idx, epx, epy, epz_i, hu, hv, mode=MODE_2D, epxy=None,
               epyx=None):
    """``solve`` that also keeps the checkpoint of its reverse pass. Returns
    (eff, amp, pe, flags, state); ``state`` holds the workspace and checkpoint tensors and is
    consumed by ``solve_adjoint``. ``epxy``/``epyx``: the off-diagonal blocks of an NVF stack
    (both or neither)."""
    device = epx.device
    ctx = context(device)
    P, L = thick.shape
    N = (2 * hu + 1) * (2 * hv + 1)
    _require(pts, torch.float64, device, 'pts')
    _require(thick, torch.float64, device, 'thick')
    _require(conv_idx, torch.int32, device, 'conv_idx')
    for name, t in (('epx', epx), ('epy', epy), ('epz_i', epz_i)):
        _require(t, torch.complex128, device, name)
    if (epxy is None) != (epyx is None):
        raise ValueError('solve_keep: give both epxy and epyx, or neither')
    for name, t in (('epxy', epxy), ('epyx', epyx)):
        if t is not None:
            _require(t, torch.complex128, device, name)
    eff = torch.empty((P, 3, 4, N), dtype=torch.float64, device=device)
    amp = torch.empty((P, 3, 4, N), dtype=torch.complex128, device=device)
    pe = torch.empty((P, 3), dtype=torch.float64, device=device)
    flags = torch.empty((P, L), dtype=torch.int32, device=device)
    nbytes = ctx.lib.wb_solve_workspace_size(P, L, hu, hv, mode)
    ckbytes = ctx.lib.wb_solve_checkpoint_size(P, L, hu, hv, mode)
    ws = _workspace(nbytes, device)
    ck = _workspace(ckbytes, device)
    ctx.check(ctx.lib.wb_solve_keep(ctx.handle, P, L, hu, hv, mode, _ptr(pts), _ptr(thick),
                                    _ptr(conv_idx), _ptr(epx), _ptr(epy), _ptr(epz_i),
                                    _ptr(epxy), _ptr(epyx), _ptr(eff), _ptr(amp), _ptr(pe),
                                    _ptr(flags), _ptr(ws), nbytes, _ptr(ck), ckbytes,
                                    _stream(device)), 'solve_keep')
    state = dict(ws=ws, ck=ck, pts=pts, thick=thick, conv_idx=conv_idx, epx=epx, epy=epy,
                 epz_i=epz_i, epxy=epxy, epyx=epyx, hu=hu, hv=hv, mode=mode)
    return eff, amp, pe, flags, state


def solve_adjoint(state, eff_bar, n_conv=None):
    """Reverse pass of ``solve_keep``: cotangent of eff (P, 3, 4, N) -> (d_epx, d_epy, d_epz,
    d_thick), with (d_epxy, d_epyx) before d_thick when the solve had off-diagonal blocks;
    conv cotangents are summed per conv slot."""
    epx = state['epx']
    device = epx.device
    ctx = context(device)
    P, L = state['thick'].shape
    hu, hv, mode = state['hu'], state['hv'], state['mode']
    N = (2 * hu + 1) * (2 * hv + 1)
    C = epx.shape[0] if n_conv is None else n_conv
    eff_bar = eff_bar.to(device=device, dtype=torch.float64).contiguous()
    if tuple(eff_bar.shape) != (P, 3, 4, N):
        raise ValueError(f'eff_bar must have shape {(P, 3, 4, N)}, got {tuple(eff_bar.shape)}')
    offd = state.get('epxy') is not None
    d = [torch.empty((C, N, N), dtype=torch.complex128, device=device)
         for _ in range(5 if offd else 3)]
    d_thick = torch.empty((P, L), dtype=torch.float64, device=device)
    nbytes = ctx.lib.wb_solve_adjoint_workspace_size(P, L, hu, hv, mode)
    ws = _workspace(nbytes, device)
    ctx.check(ctx.lib.wb_solve_adjoint(ctx.handle, P, L, hu, hv, mode, _ptr(state['pts']),
                                       _ptr(state['thick']), _ptr(state['conv_idx']), C, _ptr(epx),
                                       _ptr(state['epy']), _ptr(state['epz_i']),
                                       _ptr(state.get('epxy')), _ptr(state.get('epyx')),
                                       _ptr(state['ws']), _ptr(state['ck']), _ptr(eff_bar),
                                       _ptr(d[0]), _ptr(d[1]), _ptr(d[2]),
                                       _ptr(d[3] if offd else None), _ptr(d[4] if offd else None),
                                       _ptr(d_thick), _ptr(ws), nbytes,
                                       _stream(device)), 'solve_adjoint')
    if offd:
        return d[0], d[1], d[2], d[3], d[4], d_thick
    return d[0], d[1], d[2], d_thick


def assemble_li_adjoint(eps, hu, hv, epx, epz_i, d_epx, d_epy, d_epz):
    """Cotangent of the raster stack eps (L, Ny, Nx) from those of ``assemble_li``'s outputs."""
    device = eps.device
    ctx = context(device)
    _require(eps, torch.complex128, device, 'eps')
    for name, t in (('epx', epx), ('epz_i', epz_i), ('d_epx', d_epx), ('d_epy', d_epy),
                    ('d_epz', d_epz)):
        _require(t, torch.complex128, device, name)
    L, Ny, Nx = eps.shape
    d_eps = torch.empty_like(eps)
    nbytes = ctx.lib.wb_assemble_li_adjoint_workspace_size(L, Ny, Nx, hu, hv)
    ws = _workspace(nbytes, device)
    ctx.check(ctx.lib.wb_assemble_li_adjoint(ctx.handle, _ptr(eps), L, Ny, Nx, hu, hv, _ptr(epx),
                                             _ptr(epz_i), _ptr(d_epx), _ptr(d_epy), _ptr(d_epz),
                                             _ptr(d_eps), _ptr(ws), nbytes, _stream(device)),
              'assemble_li_adjoint')
    return d_eps


def assemble_laurent_adjoint(eps, hu, hv, epz_i, d_epx, d_epy, d_epz):
    """Cotangent of the raster stack eps (L, Ny, Nx) from those of ``assemble_laurent``'s
    outputs."""
    device = eps.device
    ctx = context(device)
    _require(eps, torch.complex128, device, 'eps')
    for name, t in (('epz_i', epz_i), ('d_epx', d_epx), ('d_epy', d_epy), ('d_epz', d_epz)):
        _require(t, torch.complex128, device, name)
    L, Ny, Nx = eps.shape
    d_eps = torch.empty_like(eps)
    nbytes = ctx.lib.wb_assemble_laurent_adjoint_workspace_size(L, Ny, Nx, hu, hv)
    ws = _workspace(nbytes, device)
    ctx.check(ctx.lib.wb_assemble_laurent_adjoint(ctx.handle, _ptr(eps), L, Ny, Nx, hu, hv,
                                                  _ptr(epz_i), _ptr(d_epx), _ptr(d_epy),
                                                  _ptr(d_epz), _ptr(d_eps), _ptr(ws), nbytes,
                                                  _stream(device)), 'assemble_laurent_adjoint')
    return d_eps


def assemble_nvf_adjoint(eps, hu, hv, d_epx, d_epy, d_epz, d_off, lattice=None):
    """Cotangent of the raster stack eps (L, Ny, Nx) from those of ``assemble_nvf``'s outputs
    (without the floor; ``d_off`` the summed cotangent of the epxy and epyx slots), the normal
    field held fixed."""
    device = eps.device
    ctx = context(device)
    _require(eps, torch.complex128, device, 'eps')
    for name, t in (('d_epx', d_epx), ('d_epy', d_epy), ('d_epz', d_epz), ('d_off', d_off)):
        _require(t, torch.complex128, device, name)
    L, Ny, Nx = eps.shape
    d_eps = torch.empty_like(eps)
    nbytes = ctx.lib.wb_assemble_nvf_adjoint_workspace_size(L, Ny, Nx, hu, hv)
    ws = _workspace(nbytes, device)
    ctx.check(ctx.lib.wb_assemble_nvf_adjoint(ctx.handle, _ptr(eps), L, Ny, Nx, hu, hv,
                                              _cell(lattice), _ptr(d_epx), _ptr(d_epy),
                                              _ptr(d_epz), _ptr(d_off), _ptr(d_eps), _ptr(ws),
                                              nbytes, _stream(device)), 'assemble_nvf_adjoint')
    return d_eps


# ----------------------------------------------------------------------------- timing

def timing_enable(on=True, stages=False):
    """Per-kernel timing (``on``), or per-stage timing (``stages``: eigensolve, S-matrices,
    assembly gathers), of subsequent launches; ``on=False`` turns timing off."""
    load_library().wb_timing_enable((2 if stages else 1) if on else 0)


def timing_read(name, reset=True):
    """(total milliseconds, launch count, algorithmic flops) of kernel ``name`` since reset."""
    lib = load_library()
    ms = ctypes.c_double()
    cnt = ctypes.c_longlong()
    fl = ctypes.c_double()
    lib.wb_timing_read(TIMER_IDS[name], ctypes.byref(ms), ctypes.byref(cnt), ctypes.byref(fl),
                       1 if reset else 0)
    return ms.value, cnt.value, fl.value


# ----------------------------------------------------------------------------- GEMM autotuning

def gemm_tuning_configs():
    """Names of the GEMM tile configurations the autotuner chooses from (index = config id)."""
    lib = load_library()
    return [lib.wb_gemm_tuning_config_name(i).decode() for i in range(lib.wb_gemm_tuning_configs())]


def gemm_tuning_record(on):
    """Record (or stop recording) the shapes of subsequent GEMM launches."""
    load_library().wb_gemm_tuning_record(1 if on else 0)


def gemm_autotune(device, reps=3):
    """Time every tile configuration on each recorded, untuned GEMM shape and keep the fastest.
    Waits for the device; returns the number of shapes tuned."""
    ctx = context(device)
    nbytes = ctx.lib.wb_gemm_autotune_size()
    scratch = _workspace(nbytes, device) if nbytes > 0 else None
    r = ctx.lib.wb_gemm_autotune(ctx.handle, int(reps), _ptr(scratch), nbytes, _stream(device))
    if r < 0:
        ctx.check(-r, 'gemm_autotune')
    torch.cuda.synchronize(device)
    return int(r)


def gemm_tuning_table():
    """The tuning table as text lines 'opa opb m n k batch config'."""
    lib = load_library()
    n = lib.wb_gemm_tuning_table(None, 0)
    buf = ctypes.create_string_buffer(int(n))
    lib.wb_gemm_tuning_table(buf, n)
    return buf.value.decode()


def gemm_tuning_load(text):
    """Merge table lines (as produced by ``gemm_tuning_table``); returns how many were taken."""
    return int(load_library().wb_gemm_tuning_load(text.encode()))


def gemm_tuning_set(opa, opb, m, n, k, batch, config):
    """Force one GEMM shape's tile configuration (config < 0 removes the entry)."""
    r = load_library().wb_gemm_tuning_set(opa, opb, m, n, k, batch, config)
    if r != 0:
        raise ValueError(f'gemm_tuning_set: bad arguments ({opa}, {opb}, {m}, {n}, {k}, {batch}, {config})')


def gemm_tuning_clear():
    """Forget the tuning table and the recorded shapes."""
    load_library().wb_gemm_tuning_clear()
