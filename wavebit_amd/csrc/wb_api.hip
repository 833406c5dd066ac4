// C ABI of the wavebit_amd engine: contexts, argument checks, workspace carving and the
// launches behind each entry point declared in include/wavebit_amd.h.
#include <algorithm>
#include <climits>
#include <cmath>
#include <string>
#include <cstring>

#include "wb_rcwa.h"
#include "../../include/wavebit_amd.h"



struct wb_ctx {
  int device;
  std::string error;
};

namespace {

int fail(wb_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->error = msg;
  return code;
}

int hip_status(wb_ctx* ctx, hipError_t e, const char* where) {
  if (e == hipSuccess) {
    if (ctx) ctx->error.clear();
    return WB_OK;
  }
  return fail(ctx, WB_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

int enter(wb_ctx* ctx) {
  if (!ctx) return WB_BADARG;
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return fail(ctx, WB_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
  return WB_OK;
}

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

extern "C" {

int wb_version(void) { return 1; }

int wb_ctx_create(int device, wb_ctx** out) {
  if (!out) return WB_BADARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return WB_BADARG;
  // the allocation-free solve's device array: its address is taken here, once per device, with
  // that device current
  int prev = 0;
  if (hipGetDevice(&prev) != hipSuccess) return WB_HIP;
  if (hipSetDevice(device) != hipSuccess) return WB_HIP;
  const hipError_t attached = wb_noalloc_attach(device);
  hipSetDevice(prev);
  if (attached != hipSuccess) return WB_HIP;
  wb_ctx* c = new (std::nothrow) wb_ctx;
  if (!c) return WB_OOM;
  c->device = device;
  *out = c;
  return WB_OK;
}

int wb_ctx_destroy(wb_ctx* ctx) {
  delete ctx;
  return WB_OK;
}

const char* wb_last_error(const wb_ctx* ctx) { return ctx ? ctx->error.c_str() : "no context"; }

int wb_zgemm_batched(wb_ctx* ctx, int m, int n, int k, const double* alpha, const double* A,
                     int lda, long long strideA, const double* B, int ldb, long long strideB,
                     const double* beta, double* C, int ldc, long long strideC, int batch,
                     void* stream) {
  int s = enter(ctx);
  if (s) return s;
  if (m < 0 || n < 0 || k < 0 || batch < 0 || !alpha || !beta || lda < k || ldb < n || ldc < n)
    return fail(ctx, WB_BADARG, "wb_zgemm_batched: bad dimensions");
  return hip_status(ctx,
                    wb_zgemm_launch(m, n, k, mkc(alpha[0], alpha[1]), bp_strided(A, strideA), lda,
                                    bp_strided(B, strideB), ldb, mkc(beta[0], beta[1]),
                                    bp_strided(C, strideC), ldc, batch, (hipStream_t)stream),
                    "wb_zgemm_batched");
}

int wb_zgetrf_batched(wb_ctx* ctx, int n, double* A, int lda, long long strideA, int* ipiv,
                      int* info, int batch, void* stream) {
  int s = enter(ctx);
  if (s) return s;
  if (n < 0 || lda < n || batch < 0 || !ipiv || !info)
    return fail(ctx, WB_BADARG, "wb_zgetrf_batched: bad arguments");
  return hip_status(ctx,
                    wb_getrf_launch(n, bp_strided(A, strideA), lda, ipiv, info, batch,
                                    (hipStream_t)stream),
                    "wb_zgetrf_batched");
}

int wb_zgetrs_batched(wb_ctx* ctx, int n, int nrhs, const double* LU, int lda, long long strideA,
                      const int* ipiv, double* B, int ldb, long long strideB, int batch,
                      void* stream) {
  int s = enter(ctx);
  if (s) return s;
  if (n < 0 || nrhs < 0 || lda < n || ldb < nrhs || batch < 0 || !ipiv)
    return fail(ctx, WB_BADARG, "wb_zgetrs_batched: bad arguments");
  return hip_status(ctx,
                    wb_getrs_launch(n, nrhs, bp_strided(LU, strideA), lda, ipiv,
                                    bp_strided(B, strideB), ldb, batch, (hipStream_t)stream),
                    "wb_zgetrs_batched");
}

size_t wb_zgeev_workspace_size(int n, int batch) {
  if (n <= 0 || batch <= 0) return 0;
  const size_t nn = (size_t)n * n * sizeof(cd);
  return 2 * align_up(nn * batch) + align_up((size_t)n * batch * sizeof(cd)) +
         align_up(wb_hqr_ms_rot_bytes(n, batch)) + 256;
}

int wb_zgeev_batched(wb_ctx* ctx, int n, double* A, double* w, double* W, void* workspace,
                     size_t workspace_bytes, int* info, int batch, void* stream) {
  int s = enter(ctx);
  if (s) return s;
  if (n < 0 || batch < 0 || !info) return fail(ctx, WB_BADARG, "wb_zgeev_batched: bad arguments");
  if (n == 0 || batch == 0) return WB_OK;
  if (workspace_bytes < wb_zgeev_workspace_size(n, batch))
    return fail(ctx, WB_OOM, "wb_zgeev_batched: workspace too small");
  char* p = (char*)workspace;
  const size_t nn = (size_t)n * n * sizeof(cd);
  cd* Z = (cd*)p; p += align_up(nn * batch);
  cd* X = (cd*)p; p += align_up(nn * batch);
  cd* tau = (cd*)p; p += align_up((size_t)n * batch * sizeof(cd));
  void* rot = p;
  return hip_status(ctx,
                    wb_zgeev_launch(n, (cd*)A, (cd*)w, (cd*)W, Z, X, tau, rot, info, batch,
                                    (hipStream_t)stream),
                    "wb_zgeev_batched");
}

size_t wb_assemble_li_workspace_size(int L, int Ny, int Nx, int hu, int hv) {
  if (L <= 0 || Ny <= 0 || Nx <= 0 || hu < 0 || hv < 0) return 0;
  return wb_assemble_workspace_bytes(L, Ny, Nx, hu, hv);
}

int wb_assemble_li(wb_ctx* ctx, const double* eps, int L, int Ny, int Nx, int hu, int hv,
                   double* epx, double* epy, double* epz_i, void* workspace,
                   size_t workspace_bytes, int* info, void* stream) {
  int s = enter(ctx);
  if (s) return s;
  if (L <= 0 || Ny <= 0 || Nx <= 0 || hu < 0 || hv < 0 || hv > 15 || !info)
    return fail(ctx, WB_BADARG, "wb_assemble_li: bad dimensions (need L, Ny, Nx >= 1, 0 <= hv <= 15)");
  if (workspace_bytes < wb_assemble_li_workspace_size(L, Ny, Nx, hu, hv))
    return fail(ctx, WB_OOM, "wb_assemble_li: workspace too small");
  return hip_status(ctx,
                    wb_assemble_li_launch((const cd*)eps, L, Ny, Nx, hu, hv, (cd*)epx, (cd*)epy,
                                          (cd*)epz_i, workspace, info, (hipStream_t)stream),
                    "wb_assemble_li");
}

size_t wb_assemble_laurent_workspace_size(int L, int Ny, int Nx, int hu, int hv) {
  if (L <= 0 || Ny <= 0 || Nx <= 0 || hu < 0 || hv < 0) return 0;
  return wb_assemble_laurent_workspace_bytes(L, Ny, Nx, hu, hv);
}

int wb_assemble_laurent(wb_ctx* ctx, const double* eps, int L, int Ny, int Nx, int hu, int hv,
                        double* epx, double* epy, double* epz_i, void* workspace,
                        size_t workspace_bytes, int* info, void* stream) {
  int s = enter(ctx);
  if (s) return s;
  if (L <= 0 || Ny <= 0 || Nx <= 0 || hu < 0 || hv < 0 || hu > 15 || hv > 15 || !info)
    return fail(ctx, WB_BADARG, "wb_assemble_laurent: bad dimensions (need L, Ny, Nx >= 1, 0 <= hu, hv <= 15)");
  if (!eps || !epx || !epy || !epz_i)
    return fail(ctx, WB_BADARG, "wb_assemble_laurent: null buffer");
  if (workspace_bytes < wb_assemble_laurent_workspace_size(L, Ny, Nx, hu, hv))
    return fail(ctx, WB_OOM, "wb_assemble_laurent: workspace too small");
  return hip_status(ctx,
                    wb_assemble_laurent_launch((const cd*)eps, L, Ny, Nx, hu, hv, (cd*)epx, (cd*)epy,
                                               (cd*)epz_i, workspace, info, (hipStream_t)stream),
                    "wb_assemble_laurent");
}

// The NVF builder's geometry from the raster's shape and the cell basis: metric = J^-T (4) and
// the index steps (2), or a constant normal (cst, ux, uy) for a one-cell axis. False for a basis
// that is not right-handed.
static bool nvf_metric(int Ny, int Nx, int hu, int hv, const double* lattice, double* metric, int* cst,
                       double* ux, double* uy) {
  // the raster lifted to 4 h + 1 samples per axis (factorization.py:811-827)
  const int Kv = 4 * hv + 1, Ku = 4 * hu + 1;
  const int Y = Ny * ((Ny < Kv) ? Kv / Ny + 1 : 1), X = Nx * ((Nx < Ku) ? Ku / Nx + 1 : 1);
  // the cell basis, columns a1 = (a1x, a1y), a2 = (a2x, a2y); square pixels when NULL
  double a1x = 1.0, a1y = 0.0, a2x = 0.0, a2y = 1.0;
  if (lattice) {
    a1x = lattice[0]; a1y = lattice[1]; a2x = lattice[2]; a2y = lattice[3];
    if (!(a1x * a2y - a2x * a1y > 0.0))
      return false;
  } else {
    a1x = (double)X; a2y = (double)Y;
  }
  const double unit[6] = {1.0, 0.0, 0.0, 1.0, 1.0, 1.0};
  for (int k = 0; k < 6; ++k) metric[k] = unit[k];
  *cst = 0; *ux = 0.0; *uy = 0.0;
  if (Y == 1 || X == 1) {
    // a one-cell axis: the normal along b1 (one row) or b2 (one column), rows of the basis
    // inverse (factorization.py:1337-1375)
    const double det = a1x * a2y - a2x * a1y;
    const double b0 = (Y == 1) ? a2y / det : -a1y / det;
    const double b1 = (Y == 1) ? -a2x / det : a1x / det;
    const double h = std::hypot(b0, b1);
    *cst = 1; *ux = b0 / h; *uy = b1 / h;
  } else {
    // index steps in units of one u-step, and J^-T (factorization.py:1110-1133)
    double j00 = a1x / X, j10 = a1y / X, j01 = a2x / Y, j11 = a2y / Y;
    const double scale = std::hypot(j00, j10);
    j00 /= scale; j10 /= scale; j01 /= scale; j11 /= scale;
    const double step_j = std::hypot(j01, j11), step_i = std::hypot(j00, j10);
    const bool same = std::fabs(step_j - step_i) <= 1e-12 * std::max(std::fabs(step_j), std::fabs(step_i));
    const double det = j00 * j11 - j01 * j10;
    metric[0] = j11 / det; metric[1] = -j10 / det; metric[2] = -j01 / det; metric[3] = j00 / det;
    metric[4] = same ? 1.0 : step_j;
    metric[5] = same ? 1.0 : step_i;
  }
  return true;
}

size_t wb_assemble_nvf_workspace_size(int L, int Ny, int Nx, int hu, int hv) {
  if (L <= 0 || Ny <= 0 || Nx <= 0 || hu < 0 || hv < 0 || hu > 15 || hv > 15) return 0;
  return wb_assemble_nvf_workspace_bytes(L, Ny, Nx, hu, hv);
}

int wb_assemble_nvf(wb_ctx* ctx, const double* eps, int L, int Ny, int Nx, int hu, int hv,
                    const double* lattice, int floor, double* epx, double* epy, double* epz_i,
                    double* epxy, void* workspace, size_t workspace_bytes, int* info, void* stream) {
  int s = enter(ctx);
  if (s) return s;
  if (L <= 0 || Ny <= 0 || Nx <= 0 || hu < 0 || hv < 0 || hu > 15 || hv > 15 || !info)
    return fail(ctx, WB_BADARG, "wb_assemble_nvf: bad dimensions (need L, Ny, Nx >= 1, 0 <= hu, hv <= 15)");
  if (!eps || !epx || !epy || !epz_i || !epxy)
    return fail(ctx, WB_BADARG, "wb_assemble_nvf: null buffer");
  if (workspace_bytes < wb_assemble_nvf_workspace_size(L, Ny, Nx, hu, hv))
    return fail(ctx, WB_OOM, "wb_assemble_nvf: workspace too small");
  double metric[6];
  int cst = 0;
  double ux = 0.0, uy = 0.0;
  if (!nvf_metric(Ny, Nx, hu, hv, lattice, metric, &cst, &ux, &uy))
    return fail(ctx, WB_BADARG, "wb_assemble_nvf: the lattice columns must form a right-handed basis");
  return hip_status(ctx,
                    wb_assemble_nvf_launch((const cd*)eps, L, Ny, Nx, hu, hv, metric, cst, ux, uy,
                                           floor ? 1 : 0, (cd*)epx, (cd*)epy, (cd*)epz_i, (cd*)epxy, workspace, info,
                                           (hipStream_t)stream),
                    "wb_assemble_nvf");
}

size_t wb_assemble_nvf_adjoint_workspace_size(int L, int Ny, int Nx, int hu, int hv) {
  if (L <= 0 || Ny <= 0 || Nx <= 0 || hu < 0 || hv < 0 || hu > 15 || hv > 15) return 0;
  return wb_assemble_nvf_adjoint_workspace_bytes(L, Ny, Nx, hu, hv);
}

int wb_assemble_nvf_adjoint(wb_ctx* ctx, const double* eps, int L, int Ny, int Nx, int hu, int hv,
                            const double* lattice, const double* d_epx, const double* d_epy,
                            const double* d_epz, const double* d_off, double* d_eps, void* workspace,
                            size_t workspace_bytes, void* stream) {
  int s = enter(ctx);
  if (s) return s;
  if (L <= 0 || Ny <= 0 || Nx <= 0 || hu < 0 || hv < 0 || hu > 15 || hv > 15)
    return fail(ctx, WB_BADARG, "wb_assemble_nvf_adjoint: bad dimensions (need L, Ny, Nx >= 1, 0 <= hu, hv <= 15)");
  if (!eps || !d_epx || !d_epy || !d_epz || !d_off || !d_eps)
    return fail(ctx, WB_BADARG, "wb_assemble_nvf_adjoint: null buffer");
  if (workspace_bytes < wb_assemble_nvf_adjoint_workspace_size(L, Ny, Nx, hu, hv))
    return fail(ctx, WB_OOM, "wb_assemble_nvf_adjoint: workspace too small");
  double metric[6];
  int cst = 0;
  double ux = 0.0, uy = 0.0;
  if (!nvf_metric(Ny, Nx, hu, hv, lattice, metric, &cst, &ux, &uy))
    return fail(ctx, WB_BADARG, "wb_assemble_nvf_adjoint: the lattice columns must form a right-handed basis");
  return hip_status(ctx,
                    wb_assemble_nvf_adjoint_launch((const cd*)eps, L, Ny, Nx, hu, hv, metric, cst, ux, uy,
                                                   (const cd*)d_epx, (const cd*)d_epy, (const cd*)d_epz,
                                                   (const cd*)d_off, (cd*)d_eps, workspace, (hipStream_t)stream),
                    "wb_assemble_nvf_adjoint");
}

size_t wb_solve_workspace_size(int P, int L, int hu, int hv, int mode) {
  if (P <= 0 || L <= 0 || hu < 0 || hv < 0 || mode < 0 || mode > 2) return 0;
  return wb_solve_workspace_bytes(P, L, hu, hv, mode);
}

int wb_solve(wb_ctx* ctx, int P, int L, int hu, int hv, int mode, const double* pts,
             const double* thick, const int* conv_idx, const double* epx, const double* epy,
             const double* epz_i, const double* epxy,
             const double* epyx, double* eff, double* amp, double* pe, int* flags,
             void* workspace, size_t workspace_bytes, void* stream) {
  int s = enter(ctx);
  if (s) return s;
  if (P <= 0 || L <= 0 || hu < 0 || hv < 0 || mode < 0 || mode > 2 || (mode != 0 && hv != 0))
    return fail(ctx, WB_BADARG, "wb_solve: bad dimensions");
  if (!pts || !thick || !conv_idx || !epx || !epy || !epz_i || !eff || !amp || !pe)
    return fail(ctx, WB_BADARG, "wb_solve: null buffer");
  if (workspace_bytes < wb_solve_workspace_size(P, L, hu, hv, mode))
    return fail(ctx, WB_OOM, "wb_solve: workspace too small");
  wb_solve_args a;
  a.P = P; a.L = L; a.hu = hu; a.hv = hv; a.mode = mode;
  a.pts = pts; a.thick = thick; a.conv_idx = conv_idx;
  a.epx = (const cd*)epx; a.epy = (const cd*)epy; a.epz = (const cd*)epz_i;
  if ((epxy == nullptr) != (epyx == nullptr))
    return fail(ctx, WB_BADARG, "wb_solve: epxy and epyx are given together or not at all");
  a.epxy = (const cd*)epxy; a.epyx = (const cd*)epyx;
  a.eff = eff; a.amp = (cd*)amp; a.pe = pe; a.flags = flags; a.ck = nullptr;
  return hip_status(ctx, wb_solve_launch(a, workspace, (hipStream_t)stream), "wb_solve");
}

int wb_solve_noalloc(wb_ctx* ctx, int P, int L, int hu, int hv, int mode, const double* pts,
                     const double* thick, const int* conv_idx, const double* epx, const double* epy,
                     const double* epz_i, double* eff, double* amp, double* pe, int* flags,
                     void* stream) {
  int s = enter(ctx);
  if (s) return s;
  if (P <= 0 || L <= 0 || hu < 0 || hv < 0 || mode < 0 || mode > 2 || (mode != 0 && hv != 0))
    return fail(ctx, WB_BADARG, "wb_solve_noalloc: bad dimensions");
  if (!pts || !thick || !conv_idx || !epx || !epy || !epz_i || !eff || !amp || !pe)
    return fail(ctx, WB_BADARG, "wb_solve_noalloc: null buffer");
  const size_t need = wb_solve_workspace_size(P, L, hu, hv, mode);
  if (need > (size_t)WB_SOLVE_NOALLOC_BYTES)
    return fail(ctx, WB_OOM,
                "wb_solve_noalloc: the problem needs " + std::to_string(need) +
                    " bytes of scratch, above the fixed " + std::to_string((size_t)WB_SOLVE_NOALLOC_BYTES) +
                    " of the allocation-free solve (fewer points per call, or wb_solve)");
  void* scratch = wb_noalloc_scratch(ctx->device);
  if (!scratch) return fail(ctx, WB_HIP, "wb_solve_noalloc: the device array was not attached");
  wb_solve_args a;
  a.P = P; a.L = L; a.hu = hu; a.hv = hv; a.mode = mode;
  a.pts = pts; a.thick = thick; a.conv_idx = conv_idx;
  a.epx = (const cd*)epx; a.epy = (const cd*)epy; a.epz = (const cd*)epz_i;
  a.eff = eff; a.amp = (cd*)amp; a.pe = pe; a.flags = flags; a.ck = nullptr;
  return hip_status(ctx, wb_solve_launch(a, scratch, (hipStream_t)stream), "wb_solve_noalloc");
}

size_t wb_solve_rebuild_workspace_size(int P, int L, int hu, int hv) {
  if (P <= 0 || L <= 0 || hu < 0 || hv < 0) return 0;
  const int n = 2 * (2 * hu + 1) * (2 * hv + 1);
  return align_up(wb_solve_workspace_bytes(P, L, hu, hv, WB_MODE_2D)) + wb_graze_bytes(P * L, n);
}

int wb_solve_rebuild(wb_ctx* ctx, int P, int L, int hu, int hv, const double* pts,
                     const double* thick, const int* conv_idx, const double* epx, const double* epy,
                     const double* epz_i, const double* epxy,
             const double* epyx, double* eff, double* amp, double* pe, int* flags,
                     double* smat, void* workspace, size_t workspace_bytes, void* stream) {
  int s = enter(ctx);
  if (s) return s;
  if (P <= 0 || L <= 0 || hu < 0 || hv < 0)
    return fail(ctx, WB_BADARG, "wb_solve_rebuild: bad dimensions");
  if (2 * (2 * hu + 1) * (2 * hv + 1) > WB_GRAZE_NMAX)
    return fail(ctx, WB_UNSUPPORTED, "wb_solve_rebuild: mode order above 2048");
  if (!pts || !thick || !conv_idx || !epx || !epy || !epz_i || !eff || !amp || !pe || !flags)
    return fail(ctx, WB_BADARG, "wb_solve_rebuild: null buffer");
  if (workspace_bytes < wb_solve_rebuild_workspace_size(P, L, hu, hv))
    return fail(ctx, WB_OOM, "wb_solve_rebuild: workspace too small");
  wb_solve_args a;
  a.P = P; a.L = L; a.hu = hu; a.hv = hv; a.mode = WB_MODE_2D;
  a.pts = pts; a.thick = thick; a.conv_idx = conv_idx;
  a.epx = (const cd*)epx; a.epy = (const cd*)epy; a.epz = (const cd*)epz_i;
  if ((epxy == nullptr) != (epyx == nullptr))
    return fail(ctx, WB_BADARG, "wb_solve_rebuild: epxy and epyx are given together or not at all");
  a.epxy = (const cd*)epxy; a.epyx = (const cd*)epyx;
  a.eff = eff; a.amp = (cd*)amp; a.pe = pe; a.flags = flags; a.ck = nullptr; a.smat = (cd*)smat;
  a.graze = (char*)workspace + align_up(wb_solve_workspace_bytes(P, L, hu, hv, WB_MODE_2D));
  return hip_status(ctx, wb_solve_launch(a, workspace, (hipStream_t)stream), "wb_solve_rebuild");
}

size_t wb_layer_modes_workspace_size(int P, int L, int hu, int hv, int mode, int rebuild) {
  if (P <= 0 || L <= 0 || hu < 0 || hv < 0 || mode < 0 || mode > 2) return 0;
  const size_t base = align_up(wb_solve_workspace_bytes(P, L, hu, hv, mode));
  if (!rebuild || mode != WB_MODE_2D) return base;
  return base + wb_graze_bytes(P * L, 2 * (2 * hu + 1) * (2 * hv + 1));
}

int wb_layer_modes(wb_ctx* ctx, int P, int L, int hu, int hv, int mode, const double* pts,
                   const int* conv_idx, const double* epx, const double* epy, const double* epz_i, const double* epxy,
             const double* epyx,
                   double* W, double* V, double* q, int* flags, int rebuild, void* workspace,
                   size_t workspace_bytes, void* stream) {
  int s = enter(ctx);
  if (s) return s;
  if (P <= 0 || L <= 0 || hu < 0 || hv < 0 || mode < 0 || mode > 2 || (mode != 0 && hv != 0))
    return fail(ctx, WB_BADARG, "wb_layer_modes: bad dimensions");
  if (!pts || !conv_idx || !epx || !epy || !epz_i || !W || !V || !q)
    return fail(ctx, WB_BADARG, "wb_layer_modes: null buffer");
  if (rebuild && mode == WB_MODE_2D && 2 * (2 * hu + 1) * (2 * hv + 1) > WB_GRAZE_NMAX)
    return fail(ctx, WB_UNSUPPORTED, "wb_layer_modes: the grazing rebuild takes mode orders up to 2048");
  if (workspace_bytes < wb_layer_modes_workspace_size(P, L, hu, hv, mode, rebuild))
    return fail(ctx, WB_OOM, "wb_layer_modes: workspace too small");
  wb_solve_args a;
  a.P = P; a.L = L; a.hu = hu; a.hv = hv; a.mode = mode;
  a.pts = pts; a.thick = nullptr; a.conv_idx = conv_idx;
  a.epx = (const cd*)epx; a.epy = (const cd*)epy; a.epz = (const cd*)epz_i;
  if ((epxy == nullptr) != (epyx == nullptr))
    return fail(ctx, WB_BADARG, "wb_layer_modes: epxy and epyx are given together or not at all");
  a.epxy = (const cd*)epxy; a.epyx = (const cd*)epyx;
  a.eff = nullptr; a.amp = nullptr; a.pe = nullptr; a.flags = flags; a.ck = nullptr;
  a.modes_out = 1; a.modes_W = (cd*)W; a.modes_V = (cd*)V; a.modes_q = (cd*)q;
  if (rebuild && mode == WB_MODE_2D)
    a.graze = (char*)workspace + align_up(wb_solve_workspace_bytes(P, L, hu, hv, mode));
  return hip_status(ctx, wb_solve_launch(a, workspace, (hipStream_t)stream), "wb_layer_modes");
}

int wb_solve_from_modes(wb_ctx* ctx, int P, int L, int hu, int hv, int mode, const double* pts,
                        const double* thick, const double* W, const double* V, const double* q,
                        const int* modes_idx, double* eff, double* amp, double* pe, int* flags,
                        double* smat, void* workspace, size_t workspace_bytes, void* stream) {
  int s = enter(ctx);
  if (s) return s;
  if (P <= 0 || L <= 0 || hu < 0 || hv < 0 || mode < 0 || mode > 2 || (mode != 0 && hv != 0))
    return fail(ctx, WB_BADARG, "wb_solve_from_modes: bad dimensions");
  if (!pts || !thick || !W || !V || !q || !eff || !amp || !pe)
    return fail(ctx, WB_BADARG, "wb_solve_from_modes: null buffer");
  if (workspace_bytes < wb_solve_workspace_size(P, L, hu, hv, mode))
    return fail(ctx, WB_OOM, "wb_solve_from_modes: workspace too small");
  wb_solve_args a;
  a.P = P; a.L = L; a.hu = hu; a.hv = hv; a.mode = mode;
  a.pts = pts; a.thick = thick; a.conv_idx = nullptr;
  a.epx = nullptr; a.epy = nullptr; a.epz = nullptr;
  a.eff = eff; a.amp = (cd*)amp; a.pe = pe; a.flags = flags; a.ck = nullptr; a.smat = (cd*)smat;
  a.modes_in = 1; a.modes_W = (cd*)W; a.modes_V = (cd*)V; a.modes_q = (cd*)q; a.modes_idx = modes_idx;
  return hip_status(ctx, wb_solve_launch(a, workspace, (hipStream_t)stream), "wb_solve_from_modes");
}

// Step counts of a ragged call: steps[l] = points with more than l layers. Returns the pair
// count, or -1 if the layer counts are not a non-increasing sequence in [1, WB_RAGGED_MAX].
static long long ragged_steps(int P, const int* layers, int* steps, int* Lmax) {
  if (P <= 0 || !layers) return -1;
  long long B = 0;
  for (int p = 0; p < P; ++p) {
    if (layers[p] < 1 || layers[p] > WB_RAGGED_MAX || (p > 0 && layers[p] > layers[p - 1])) return -1;
    B += layers[p];
  }
  *Lmax = layers[0];
  for (int l = 0; l < *Lmax; ++l) {
    int c = 0;
    while (c < P && layers[c] > l) ++c;
    steps[l] = c;
  }
  return B;
}

size_t wb_solve_ragged_workspace_size(int P, const int* layers, int hu, int hv, int mode) {
  int steps[WB_RAGGED_MAX], Lmax = 0;
  const long long B = ragged_steps(P, layers, steps, &Lmax);
  if (B < 0 || B > INT_MAX || hu < 0 || hv < 0 || mode < 0 || mode > 2) return 0;
  return align_up(wb_solve_workspace_bytes_b(P, (int)B, hu, hv, mode)) + align_up((size_t)(B + P) * sizeof(int));
}

int wb_solve_from_modes_ragged(wb_ctx* ctx, int P, const int* layers, int hu, int hv, int mode,
                               const double* pts, const double* thick, const double* W,
                               const double* V, const double* q, const int* modes_idx, double* eff,
                               double* amp, double* pe, int* flags, void* workspace,
                               size_t workspace_bytes, void* stream) {
  int s = enter(ctx);
  if (s) return s;
  int steps[WB_RAGGED_MAX], Lmax = 0;
  const long long B = ragged_steps(P, layers, steps, &Lmax);
  if (B < 0 || B > INT_MAX || hu < 0 || hv < 0 || mode < 0 || mode > 2 || (mode != 0 && hv != 0))
    return fail(ctx, WB_BADARG,
                "wb_solve_from_modes_ragged: bad dimensions (layer counts must be non-increasing, "
                "1..64)");
  if (!pts || !thick || !W || !V || !q || !eff || !amp || !pe)
    return fail(ctx, WB_BADARG, "wb_solve_from_modes_ragged: null buffer");
  if (workspace_bytes < wb_solve_ragged_workspace_size(P, layers, hu, hv, mode))
    return fail(ctx, WB_OOM, "wb_solve_from_modes_ragged: workspace too small");
  wb_solve_args a;
  a.P = P; a.L = Lmax; a.hu = hu; a.hv = hv; a.mode = mode;
  a.pts = pts; a.thick = thick; a.conv_idx = nullptr;
  a.epx = nullptr; a.epy = nullptr; a.epz = nullptr;
  a.eff = eff; a.amp = (cd*)amp; a.pe = pe; a.flags = flags; a.ck = nullptr; a.smat = nullptr;
  a.modes_in = 1; a.modes_W = (cd*)W; a.modes_V = (cd*)V; a.modes_q = (cd*)q; a.modes_idx = modes_idx;
  a.steps = steps;
  a.ragged_tab = (int*)((char*)workspace + align_up(wb_solve_workspace_bytes_b(P, (int)B, hu, hv, mode)));
  return hip_status(ctx, wb_solve_launch(a, workspace, (hipStream_t)stream), "wb_solve_from_modes_ragged");
}

int wb_flags_status(const int* flags, long long count, int notice_mask) {
  if (!flags || count < 0) return WB_BADARG;
  int all = 0;
  for (long long i = 0; i < count; ++i) all |= flags[i];
  if (all & WB_FLAG_NOCONVERGE) return WB_NOCONVERGE;
  if (all & WB_FLAG_SINGULAR) return WB_SINGULAR;
  if (all & notice_mask) return WB_FLAGGED;
  return WB_OK;
}

size_t wb_field_state_workspace_size(int P, int L, int hu, int hv, int mode) {
  if (P <= 0 || L <= 0 || hu < 0 || hv < 0 || mode < 0 || mode > 2) return 0;
  const int N = (2 * hu + 1) * (2 * hv + 1);
  const int n = (mode == WB_MODE_2D) ? 2 * N : N;
  return align_up(wb_solve_workspace_bytes(P, L, hu, hv, mode)) +
         wb_field_extra_bytes(P, L, n, mode == WB_MODE_2D ? 2 : 1);
}

int wb_field_state(wb_ctx* ctx, int P, int L, int hu, int hv, int mode, const double* pts,
                   const double* thick, const double* W, const double* V, const double* q,
                   const int* modes_idx, double* amps, double* eff, double* amp, double* pe,
                   int* flags, void* workspace, size_t workspace_bytes, void* stream) {
  int s = enter(ctx);
  if (s) return s;
  if (P <= 0 || L <= 0 || hu < 0 || hv < 0 || mode < 0 || mode > 2 || (mode != 0 && hv != 0))
    return fail(ctx, WB_BADARG, "wb_field_state: bad dimensions");
  if (!pts || !thick || !W || !V || !q || !amps || !eff || !amp || !pe)
    return fail(ctx, WB_BADARG, "wb_field_state: null buffer");
  if (workspace_bytes < wb_field_state_workspace_size(P, L, hu, hv, mode))
    return fail(ctx, WB_OOM, "wb_field_state: workspace too small");
  wb_solve_args a;
  a.P = P; a.L = L; a.hu = hu; a.hv = hv; a.mode = mode;
  a.pts = pts; a.thick = thick; a.conv_idx = nullptr;
  a.epx = nullptr; a.epy = nullptr; a.epz = nullptr;
  a.eff = eff; a.amp = (cd*)amp; a.pe = pe; a.flags = flags; a.ck = nullptr;
  a.modes_in = 1; a.modes_W = (cd*)W; a.modes_V = (cd*)V; a.modes_q = (cd*)q; a.modes_idx = modes_idx;
  a.field_amps = (cd*)amps;
  a.field_ws = (char*)workspace + align_up(wb_solve_workspace_bytes(P, L, hu, hv, mode));
  return hip_status(ctx, wb_solve_launch(a, workspace, (hipStream_t)stream), "wb_field_state");
}

size_t wb_field_sample_workspace_size(int P, int L, int hu, int hv, int mode, int nz, int nxy) {
  if (P <= 0 || L <= 0 || hu < 0 || hv < 0 || mode < 0 || mode > 2 || nz <= 0 || nxy <= 0) return 0;
  return wb_field_sample_bytes(P, L, hu, hv, mode, nz, nxy);
}

int wb_field_sample(wb_ctx* ctx, int P, int L, int hu, int hv, int mode, const double* pts,
                    const double* thick, const double* W, const double* V, const double* q,
                    const int* modes_idx, const double* epz_i, const int* conv_idx,
                    const double* amps, int nz, const double* z, int nxy, const double* xy,
                    const double* scale, double* out, void* workspace, size_t workspace_bytes,
                    void* stream) {
  int s = enter(ctx);
  if (s) return s;
  if (P <= 0 || L <= 0 || hu < 0 || hv < 0 || mode < 0 || mode > 2 || (mode != 0 && hv != 0) ||
      nz <= 0 || nxy <= 0)
    return fail(ctx, WB_BADARG, "wb_field_sample: bad dimensions");
  if (!pts || !thick || !W || !V || !q || !amps || !z || !xy || !scale || !out ||
      (mode != WB_MODE_1D_TE && !epz_i))
    return fail(ctx, WB_BADARG, "wb_field_sample: null buffer");
  if (workspace_bytes < wb_field_sample_workspace_size(P, L, hu, hv, mode, nz, nxy))
    return fail(ctx, WB_OOM, "wb_field_sample: workspace too small");
  cd f[6];
  for (int k = 0; k < 6; ++k) f[k] = mkc(scale[2 * k], scale[2 * k + 1]);
  return hip_status(ctx,
                    wb_field_sample_launch(P, L, hu, hv, mode, pts, thick, (const cd*)W, (const cd*)V,
                                           (const cd*)q, modes_idx, (const cd*)epz_i, conv_idx,
                                           (const cd*)amps, nz, z, nxy, xy, f, (cd*)out, workspace,
                                           (hipStream_t)stream),
                    "wb_field_sample");
}

int wb_solve_scattering(wb_ctx* ctx, int P, int L, int hu, int hv, int mode, const double* pts,
                        const double* thick, const int* conv_idx, const double* epx,
                        const double* epy, const double* epz_i, const double* epxy,
             const double* epyx, double* eff, double* amp, double* pe,
                        int* flags, double* smat, void* workspace, size_t workspace_bytes,
                        void* stream) {
  int s = enter(ctx);
  if (s) return s;
  if (P <= 0 || L <= 0 || hu < 0 || hv < 0 || mode < 0 || mode > 2 || (mode != 0 && hv != 0))
    return fail(ctx, WB_BADARG, "wb_solve_scattering: bad dimensions");
  if (!pts || !thick || !conv_idx || !epx || !epy || !epz_i || !eff || !amp || !pe || !smat)
    return fail(ctx, WB_BADARG, "wb_solve_scattering: null buffer");
  if (workspace_bytes < wb_solve_workspace_size(P, L, hu, hv, mode))
    return fail(ctx, WB_OOM, "wb_solve_scattering: workspace too small");
  wb_solve_args a;
  a.P = P; a.L = L; a.hu = hu; a.hv = hv; a.mode = mode;
  a.pts = pts; a.thick = thick; a.conv_idx = conv_idx;
  a.epx = (const cd*)epx; a.epy = (const cd*)epy; a.epz = (const cd*)epz_i;
  if ((epxy == nullptr) != (epyx == nullptr))
    return fail(ctx, WB_BADARG, "wb_solve_scattering: epxy and epyx are given together or not at all");
  a.epxy = (const cd*)epxy; a.epyx = (const cd*)epyx;
  a.eff = eff; a.amp = (cd*)amp; a.pe = pe; a.flags = flags; a.ck = nullptr; a.smat = (cd*)smat;
  return hip_status(ctx, wb_solve_launch(a, workspace, (hipStream_t)stream), "wb_solve_scattering");
}

int wb_zgemm_batched_ops(wb_ctx* ctx, int opa, int opb, int m, int n, int k, const double* alpha,
                         const double* A, int lda, long long strideA, const double* B, int ldb,
                         long long strideB, const double* beta, double* C, int ldc,
                         long long strideC, int batch, void* stream) {
  int s = enter(ctx);
  if (s) return s;
  if (opa < 0 || opa > 1 || opb < 0 || opb > 1)
    return fail(ctx, WB_BADARG, "wb_zgemm_batched_ops: op must be 0 (none) or 1 (conjugate transpose)");
  if (m < 0 || n < 0 || k < 0 || batch < 0 || !alpha || !beta || lda < (opa ? m : k) ||
      ldb < (opb ? k : n) || ldc < n)
    return fail(ctx, WB_BADARG, "wb_zgemm_batched_ops: bad dimensions");
  return hip_status(ctx,
                    wb_zgemm_ops_launch(opa, opb, m, n, k, mkc(alpha[0], alpha[1]),
                                        bp_strided(A, strideA), lda, bp_strided(B, strideB), ldb,
                                        mkc(beta[0], beta[1]), bp_strided(C, strideC), ldc, batch,
                                        (hipStream_t)stream),
                    "wb_zgemm_batched_ops");
}

int wb_zgetrs_batched_ops(wb_ctx* ctx, int op, int n, int nrhs, const double* LU, int lda,
                          long long strideA, const int* ipiv, double* B, int ldb, long long strideB,
                          int batch, void* stream) {
  int s = enter(ctx);
  if (s) return s;
  if (op < 0 || op > 1 || n < 0 || nrhs < 0 || lda < n || ldb < nrhs || batch < 0 || !ipiv)
    return fail(ctx, WB_BADARG, "wb_zgetrs_batched_ops: bad arguments");
  hipError_t e = op ? wb_getrs_conj_launch(n, nrhs, bp_strided(LU, strideA), lda, ipiv,
                                           bp_strided(B, strideB), ldb, batch, (hipStream_t)stream)
                    : wb_getrs_launch(n, nrhs, bp_strided(LU, strideA), lda, ipiv,
                                      bp_strided(B, strideB), ldb, batch, (hipStream_t)stream);
  return hip_status(ctx, e, "wb_zgetrs_batched_ops");
}

size_t wb_solve_checkpoint_size(int P, int L, int hu, int hv, int mode) {
  if (P <= 0 || L <= 0 || hu < 0 || hv < 0 || mode != 0) return 0;
  return wb_checkpoint_bytes(P, L, hu, hv, mode);
}

int wb_solve_keep(wb_ctx* ctx, int P, int L, int hu, int hv, int mode, const double* pts,
                  const double* thick, const int* conv_idx, const double* epx, const double* epy,
                  const double* epz_i, const double* epxy, const double* epyx, double* eff,
                  double* amp, double* pe, int* flags,
                  void* workspace, size_t workspace_bytes, void* checkpoint,
                  size_t checkpoint_bytes, void* stream) {
  int s = enter(ctx);
  if (s) return s;
  if (mode != 0)
    return fail(ctx, WB_UNSUPPORTED, "wb_solve_keep: the adjoint is implemented for the 2D path (mode 0)");
  if (P <= 0 || L <= 0 || hu < 0 || hv < 0)
    return fail(ctx, WB_BADARG, "wb_solve_keep: bad dimensions");
  if (!pts || !thick || !conv_idx || !epx || !epy || !epz_i || !eff || !amp || !pe || !checkpoint)
    return fail(ctx, WB_BADARG, "wb_solve_keep: null buffer");
  if ((epxy == nullptr) != (epyx == nullptr))
    return fail(ctx, WB_BADARG, "wb_solve_keep: pass both off-diagonal slots or neither");
  if (workspace_bytes < wb_solve_workspace_size(P, L, hu, hv, mode))
    return fail(ctx, WB_OOM, "wb_solve_keep: workspace too small");
  if (checkpoint_bytes < wb_solve_checkpoint_size(P, L, hu, hv, mode))
    return fail(ctx, WB_OOM, "wb_solve_keep: checkpoint too small");
  wb_solve_args a;
  a.P = P; a.L = L; a.hu = hu; a.hv = hv; a.mode = mode;
  a.pts = pts; a.thick = thick; a.conv_idx = conv_idx;
  a.epx = (const cd*)epx; a.epy = (const cd*)epy; a.epz = (const cd*)epz_i;
  a.epxy = (const cd*)epxy; a.epyx = (const cd*)epyx;
  a.eff = eff; a.amp = (cd*)amp; a.pe = pe; a.flags = flags; a.ck = checkpoint;
  return hip_status(ctx, wb_solve_launch(a, workspace, (hipStream_t)stream), "wb_solve_keep");
}

size_t wb_solve_adjoint_workspace_size(int P, int L, int hu, int hv, int mode) {
  if (P <= 0 || L <= 0 || hu < 0 || hv < 0 || mode != 0) return 0;
  return wb_adjoint_workspace_bytes(P, L, hu, hv, mode);
}

int wb_solve_adjoint(wb_ctx* ctx, int P, int L, int hu, int hv, int mode, const double* pts,
                     const double* thick, const int* conv_idx, int n_conv, const double* epx,
                     const double* epy, const double* epz_i, const double* epxy,
                     const double* epyx, const void* solve_workspace, const void* checkpoint,
                     const double* eff_bar, double* d_epx, double* d_epy, double* d_epz,
                     double* d_epxy, double* d_epyx, double* d_thick, void* workspace,
                     size_t workspace_bytes, void* stream) {
  int s = enter(ctx);
  if (s) return s;
  if (mode != 0)
    return fail(ctx, WB_UNSUPPORTED, "wb_solve_adjoint: implemented for the 2D path (mode 0)");
  if (P <= 0 || L <= 0 || hu < 0 || hv < 0 || n_conv <= 0)
    return fail(ctx, WB_BADARG, "wb_solve_adjoint: bad dimensions");
  if (!pts || !thick || !conv_idx || !epx || !epy || !epz_i || !solve_workspace || !checkpoint ||
      !eff_bar || !d_epx || !d_epy || !d_epz || !d_thick)
    return fail(ctx, WB_BADARG, "wb_solve_adjoint: null buffer");
  if ((epxy == nullptr) != (epyx == nullptr) || (epxy == nullptr) != (d_epxy == nullptr) ||
      (epxy == nullptr) != (d_epyx == nullptr))
    return fail(ctx, WB_BADARG, "wb_solve_adjoint: the off-diagonal slots and their cotangents go together");
  if (workspace_bytes < wb_solve_adjoint_workspace_size(P, L, hu, hv, mode))
    return fail(ctx, WB_OOM, "wb_solve_adjoint: workspace too small");
  wb_solve_args f;
  f.P = P; f.L = L; f.hu = hu; f.hv = hv; f.mode = mode;
  f.pts = pts; f.thick = thick; f.conv_idx = conv_idx;
  f.epx = (const cd*)epx; f.epy = (const cd*)epy; f.epz = (const cd*)epz_i;
  f.epxy = (const cd*)epxy; f.epyx = (const cd*)epyx;
  f.eff = nullptr; f.amp = nullptr; f.pe = nullptr; f.flags = nullptr;
  f.ck = const_cast<void*>(checkpoint);
  wb_adjoint_args a;
  a.fwd = &f; a.n_conv = n_conv; a.solve_ws = solve_workspace; a.ck = checkpoint;
  a.eff_bar = eff_bar; a.d_epx = (cd*)d_epx; a.d_epy = (cd*)d_epy; a.d_epz = (cd*)d_epz;
  a.d_epxy = (cd*)d_epxy; a.d_epyx = (cd*)d_epyx;
  a.d_thick = d_thick;
  return hip_status(ctx, wb_adjoint_launch(a, workspace, (hipStream_t)stream), "wb_solve_adjoint");
}

size_t wb_assemble_li_adjoint_workspace_size(int L, int Ny, int Nx, int hu, int hv) {
  if (L <= 0 || Ny <= 0 || Nx <= 0 || hu < 0 || hv < 0) return 0;
  return wb_assemble_adjoint_workspace_bytes(L, Ny, Nx, hu, hv);
}

int wb_assemble_li_adjoint(wb_ctx* ctx, const double* eps, int L, int Ny, int Nx, int hu, int hv,
                           const double* epx, const double* epz_i, const double* d_epx,
                           const double* d_epy, const double* d_epz, double* d_eps,
                           void* workspace, size_t workspace_bytes, void* stream) {
  int s = enter(ctx);
  if (s) return s;
  if (L <= 0 || Ny <= 0 || Nx <= 0 || hu < 0 || hv < 0 || hu > 15 || hv > 15)
    return fail(ctx, WB_BADARG, "wb_assemble_li_adjoint: bad dimensions (need 0 <= hu, hv <= 15)");
  if (!eps || !epx || !epz_i || !d_epx || !d_epy || !d_epz || !d_eps)
    return fail(ctx, WB_BADARG, "wb_assemble_li_adjoint: null buffer");
  if (workspace_bytes < wb_assemble_li_adjoint_workspace_size(L, Ny, Nx, hu, hv))
    return fail(ctx, WB_OOM, "wb_assemble_li_adjoint: workspace too small");
  return hip_status(ctx,
                    wb_assemble_adjoint_launch((const cd*)eps, L, Ny, Nx, hu, hv, (const cd*)epx,
                                               (const cd*)epz_i, (const cd*)d_epx, (const cd*)d_epy,
                                               (const cd*)d_epz, (cd*)d_eps, workspace,
                                               (hipStream_t)stream),
                    "wb_assemble_li_adjoint");
}

size_t wb_assemble_laurent_adjoint_workspace_size(int L, int Ny, int Nx, int hu, int hv) {
  if (L <= 0 || Ny <= 0 || Nx <= 0 || hu < 0 || hv < 0) return 0;
  return wb_assemble_laurent_adjoint_workspace_bytes(L, Ny, Nx, hu, hv);
}

int wb_assemble_laurent_adjoint(wb_ctx* ctx, const double* eps, int L, int Ny, int Nx, int hu, int hv,
                                const double* epz_i, const double* d_epx, const double* d_epy,
                                const double* d_epz, double* d_eps, void* workspace,
                                size_t workspace_bytes, void* stream) {
  int s = enter(ctx);
  if (s) return s;
  if (L <= 0 || Ny <= 0 || Nx <= 0 || hu < 0 || hv < 0 || hu > 15 || hv > 15)
    return fail(ctx, WB_BADARG, "wb_assemble_laurent_adjoint: bad dimensions (need 0 <= hu, hv <= 15)");
  if (!eps || !epz_i || !d_epx || !d_epy || !d_epz || !d_eps)
    return fail(ctx, WB_BADARG, "wb_assemble_laurent_adjoint: null buffer");
  if (workspace_bytes < wb_assemble_laurent_adjoint_workspace_size(L, Ny, Nx, hu, hv))
    return fail(ctx, WB_OOM, "wb_assemble_laurent_adjoint: workspace too small");
  return hip_status(ctx,
                    wb_assemble_laurent_adjoint_launch((const cd*)eps, L, Ny, Nx, hu, hv, (const cd*)epz_i,
                                                       (const cd*)d_epx, (const cd*)d_epy,
                                                       (const cd*)d_epz, (cd*)d_eps, workspace,
                                                       (hipStream_t)stream),
                    "wb_assemble_laurent_adjoint");
}

int wb_gemm_tuning_configs(void) { return wb_gemm_config_count(); }
const char* wb_gemm_tuning_config_name(int config) { return wb_gemm_config_name(config); }
int wb_gemm_tuning_record(int on) {
  wb_gemm_record(on);
  return WB_OK;
}
size_t wb_gemm_autotune_size(void) { return wb_gemm_autotune_bytes(); }
int wb_gemm_autotune(wb_ctx* ctx, int reps, void* scratch, size_t scratch_bytes, void* stream) {
  int s = enter(ctx);
  if (s) return -s;
  if (reps < 1) return -fail(ctx, WB_BADARG, "wb_gemm_autotune: reps must be >= 1");
  const size_t need = wb_gemm_autotune_bytes();
  if (need > 0 && (!scratch || scratch_bytes < need))
    return -fail(ctx, WB_OOM, "wb_gemm_autotune: scratch smaller than wb_gemm_autotune_size()");
  hipError_t e = hipSuccess;
  const int tuned = wb_gemm_autotune_recorded(reps, scratch, scratch_bytes, (hipStream_t)stream, &e);
  if (tuned < 0) return -hip_status(ctx, e, "wb_gemm_autotune");
  ctx->error.clear();
  return tuned;
}
size_t wb_gemm_tuning_table(char* out, size_t cap) { return wb_gemm_table_text(out, cap); }
int wb_gemm_tuning_load(const char* text) { return wb_gemm_table_load(text); }
int wb_gemm_tuning_set(int opa, int opb, int m, int n, int k, int batch, int config) {
  if (opa < 0 || opa > 1 || opb < 0 || opb > 1 || m < 0 || n < 0 || k < 0 || batch < 0 ||
      config >= wb_gemm_config_count())
    return WB_BADARG;
  wb_gemm_table_set(opa, opb, m, n, k, batch, config);
  return WB_OK;
}
void wb_gemm_tuning_clear(void) { wb_gemm_table_clear(); }

}  // extern "C"
