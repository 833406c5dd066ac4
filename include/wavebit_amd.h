/* wavebit_amd -- C ABI of the MI355X RCWA engine (gfx950, complex FP64).
 *
 * Every buffer is caller-allocated device memory (for example a torch tensor's data_ptr()),
 * complex128 as interleaved {re, im} doubles, matrices row-major. The library never frees
 * caller memory and allocates no device memory of its own: each entry point that needs
 * scratch takes a workspace pointer and byte count, sized by its *_workspace_size query.
 * No C++ exception crosses this boundary; every call returns a wb_status and leaves a
 * message for wb_last_error(). `stream` is a hipStream_t (NULL = the null stream). Work is
 * enqueued asynchronously; outputs are valid once the stream has reached the call.
 *
 * A context belongs to one host thread and one device; share it across threads only under an
 * external lock.
 *
 * The reference (fieldalgo/wavebit) is pure Python; the functions below replace the internal
 * seams its raster RCWA path calls (paths relative to the reference's src/wavebit/):
 *   wb_assemble_li   rcwa/_engine/factorization.py:915-940  to_conv_mat_raster_sampled
 *                    (li_conv_matrices, :658-671), one call for a whole layer stack;
 *   wb_solve         rcwa/_engine/scattering.py:1386-1456 smm_2d and :704-902 smm_1d as the
 *                    engine calls them (rcwa/_engine/engine.py:3283-3316, 3336-3434), for a
 *                    batch of plane-wave points instead of one wavelength per call
 *                    (rcwa/_lifecycle.py:712-727);
 *   wb_zgeev_batched _common/backend.py:1767-1812 eig (numpy.linalg.eig / torch.linalg.eig);
 *   wb_zgetrf_batched / wb_zgetrs_batched / wb_zgemm_batched
 *                    the dense inv / solve / matmul calls of rcwa/_engine/scattering.py:116-200
 *                    (_ab_matrices, _s_layer, redheffer_star) and _common/backend.py:506-542
 *                    (linear_solve).
 */
#ifndef WAVEBIT_AMD_H
#define WAVEBIT_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wb_ctx wb_ctx;

typedef enum {
  WB_OK = 0,
  WB_BADARG = 1,      /* an argument is out of range or inconsistent */
  WB_OOM = 2,         /* the workspace given is smaller than the query returned */
  WB_SINGULAR = 3,    /* reserved: singular factorizations are reported through flags/info */
  WB_NONFINITE = 4,   /* reserved for callers that check outputs */
  WB_NOCONVERGE = 5,  /* reserved: eigensolver non-convergence is reported through flags/info */
  WB_FLAGGED = 6,     /* reserved */
  WB_HIP = 7          /* a HIP runtime call failed; see wb_last_error */
} wb_status;

/* wb_solve mode */
enum { WB_MODE_2D = 0, WB_MODE_1D_TE = 1, WB_MODE_1D_TM = 2 };

/* per-point record layout for wb_solve (doubles, WB_PT_STRIDE per point) */
enum {
  WB_PT_WAVELENGTH = 0,   /* vacuum wavelength, same length unit as the periods */
  WB_PT_KX0 = 1,          /* incident in-plane wavevector / k0 (real angles) */
  WB_PT_KY0 = 2,
  WB_PT_PHI = 3,          /* frame azimuth at k_par = 0 (radians) */
  WB_PT_PERIOD_X = 4,
  WB_PT_PERIOD_Y = 5,
  WB_PT_EPS_TOP = 6,      /* superstrate permittivity (re, im): the lit half-space */
  WB_PT_EPS_BOTTOM = 8,   /* substrate permittivity (re, im) */
  WB_PT_A_S = 10,         /* incident Jones pair, unit norm (re, im) each */
  WB_PT_A_P = 12,
  WB_PT_STRIDE = 16
};

/* flag bits written per (point, layer) by wb_solve */
enum { WB_FLAG_NOCONVERGE = 1, WB_FLAG_SINGULAR = 2 };

int wb_version(void);
int wb_ctx_create(int device, wb_ctx** out);
int wb_ctx_destroy(wb_ctx* ctx);
const char* wb_last_error(const wb_ctx* ctx);

/* C[b] = alpha A[b] B[b] + beta C[b]; A m x k, B k x n, C m x n; alpha/beta point to two
 * doubles on the host. stride* are element strides between batch entries (0 = shared). */
int wb_zgemm_batched(wb_ctx* ctx, int m, int n, int k, const double* alpha, const double* A,
                     int lda, long long strideA, const double* B, int ldb, long long strideB,
                     const double* beta, double* C, int ldc, long long strideC, int batch,
                     void* stream);

/* LU with partial pivoting in place; ipiv: batch x n (0-based row swapped with row j);
 * info: batch ints, 0 or j+1 for the first exactly-zero pivot. */
int wb_zgetrf_batched(wb_ctx* ctx, int n, double* A, int lda, long long strideA, int* ipiv,
                      int* info, int batch, void* stream);

/* Solve A X = B in place in B from wb_zgetrf_batched's factors. */
int wb_zgetrs_batched(wb_ctx* ctx, int n, int nrhs, const double* LU, int lda, long long strideA,
                      const int* ipiv, double* B, int ldb, long long strideB, int batch,
                      void* stream);

/* General complex eigendecomposition A W = W diag(w) of `batch` contiguous n x n matrices
 * (A is destroyed). W columns have unit 2-norm with their largest component real.
 * info: batch ints, 0 or i+1 when the QR iteration did not converge at eigenvalue i. */
size_t wb_zgeev_workspace_size(int n, int batch);
int wb_zgeev_batched(wb_ctx* ctx, int n, double* A, double* w, double* W, void* workspace,
                     size_t workspace_bytes, int* info, int batch, void* stream);

/* Li-rule convolution matrices (sampled coefficients, fixed nesting) of L raster layers
 * eps (L, Ny, Nx): epx, epy, epz_i, each (L, N, N), N = (2hu+1)(2hv+1). hv <= 15.
 * info: 2L ints (inversion of the outer epx operator per layer, then of the direct rule). */
size_t wb_assemble_li_workspace_size(int L, int Ny, int Nx, int hu, int hv);
int wb_assemble_li(wb_ctx* ctx, const double* eps, int L, int Ny, int Nx, int hu, int hv,
                   double* epx, double* epy, double* epz_i, void* workspace,
                   size_t workspace_bytes, int* info, void* stream);

/* Batch solve of P plane-wave points, each through L layers (bottom to top), top incidence.
 *   pts       P x WB_PT_STRIDE doubles (layout above)
 *   thick     P x L layer thicknesses
 *   conv_idx  P x L: which (epx, epy, epz_i) matrix set layer l of point p uses
 *   epx, epy, epz_i   C x N x N complex
 * Outputs:
 *   eff  P x 3 x 4 x N doubles: incidence variant (as given, pure s, pure p) x (R_s, R_p, T_s,
 *        T_p) per order, orders flattened row-major over (2hv+1, 2hu+1). 1D modes fill
 *        variant 0 only.
 *   amp  P x 3 x 4 x N complex: (r_s, r_p, t_s, t_p) Fresnel amplitudes, centre-referenced
 *   pe   P x 3 power entering the lit face / incident power
 *   flags P x L ints (WB_FLAG_*) or NULL */
size_t wb_solve_workspace_size(int P, int L, int hu, int hv, int mode);
int wb_solve(wb_ctx* ctx, int P, int L, int hu, int hv, int mode, const double* pts,
             const double* thick, const int* conv_idx, const double* epx, const double* epy,
             const double* epz_i, double* eff, double* amp, double* pe, int* flags,
             void* workspace, size_t workspace_bytes, void* stream);

/* Per-kernel timing (no reference counterpart: measurement support). When enabled, the
 * launches of the kernels below are bracketed by hipEvent records on their stream;
 * wb_timing_read returns the summed milliseconds, launch count and algorithmic real FP64 flops
 * (GEMM 8mnk; eigensolver stages from EIG(n) = 100 n^3) since the last reset. */
enum {
  WB_T_ZGEMM = 0, WB_T_GETRF_PANEL, WB_T_LASWP_TRSM, WB_T_TRSM_DIAG, WB_T_GEHRD, WB_T_UNGHR,
  WB_T_HQR, WB_T_TREVC, WB_T_COLNORM, WB_T_ASSEMBLE, WB_T_COUNT
};
int wb_timing_enable(int on);
int wb_timing_read(int id, double* total_ms, long long* count, double* flops, int reset);

#ifdef __cplusplus
}
#endif

#endif /* WAVEBIT_AMD_H */
