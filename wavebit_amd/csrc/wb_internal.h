// Internal launch interfaces shared by the wavebit_amd translation units.
#pragma once
#include "wb_common.h"

hipError_t wb_zgemm_launch(int m, int n, int k, cd alpha, BatchPtr A, int lda, BatchPtr B, int ldb,
                           cd beta, BatchPtr C, int ldc, int batch, hipStream_t stream);
hipError_t wb_getrf_launch(int n, BatchPtr A, int lda, int* ipiv, int* info, int batch,
                           hipStream_t stream);
hipError_t wb_getrs_launch(int n, int nrhs, BatchPtr LU, int lda, const int* ipiv, BatchPtr B,
                           int ldb, int batch, hipStream_t stream);
hipError_t wb_getrs_conj_launch(int n, int nrhs, BatchPtr LU, int lda, const int* ipiv, BatchPtr B,
                                int ldb, int batch, hipStream_t stream);
// GEMM tile-shape autotuner (wb_gemm.hip)
void wb_gemm_record(int on);
size_t wb_gemm_autotune_bytes();
int wb_gemm_autotune_recorded(int reps, void* scratch, size_t bytes, hipStream_t stream, hipError_t* err);
size_t wb_gemm_table_text(char* out, size_t cap);
int wb_gemm_table_load(const char* text);
void wb_gemm_table_clear();
void wb_gemm_table_set(int opa, int opb, int m, int n, int k, int batch, int cfg);
const char* wb_gemm_config_name(int cfg);
int wb_gemm_config_count();
hipError_t wb_zgemm_ops_launch(int opa, int opb, int m, int n, int k, cd alpha, BatchPtr A, int lda,
                               BatchPtr B, int ldb, cd beta, BatchPtr C, int ldc, int batch,
                               hipStream_t stream);
hipError_t wb_set_identity_launch(int n, BatchPtr B, int ldb, int batch, hipStream_t stream);
// Many-right-hand-side solves blocked by 64 rows on GEMMs (scratch: wb_getrs_ws_bytes).
size_t wb_getrs_ws_bytes(int n, int nrhs, int batch);
hipError_t wb_getrs_blocked(int n, int nrhs, BatchPtr LU, int lda, const int* ipiv, BatchPtr B,
                            int ldb, int batch, int conj, void* ws, hipStream_t stream);
hipError_t wb_inverse_blocked(int n, BatchPtr A, int lda, BatchPtr Ainv, int ldi, int* ipiv, int* info,
                              int batch, void* ws, hipStream_t stream);
hipError_t wb_inverse_launch(int n, BatchPtr A, int lda, BatchPtr Ainv, int ldi, int* ipiv,
                             int* info, int batch, hipStream_t stream);
hipError_t wb_zero_int_launch(int* x, int count, hipStream_t stream);
// The allocation-free solve's device array (wb_noalloc.hip): its address on `device`, taken once
// per device at context creation with that device current.
hipError_t wb_noalloc_attach(int device);
void* wb_noalloc_scratch(int device);
hipError_t wb_zgeev_launch(int n, cd* A, cd* w, cd* W, cd* Z, cd* X, cd* tau, void* rot_ws, int* info,
                           int batch, hipStream_t stream);
size_t wb_hqr_ms_rot_bytes(int n, int batch);
// The Schur form alone: T in A, the Schur vectors in Z (as Z^H when *zh is set on return);
// W, X: n x n scratch per matrix, tau: n per matrix.
hipError_t wb_schur_launch(int n, cd* A, cd* w, cd* W, cd* Z, cd* X, cd* tau, void* rot_ws, int* info,
                           int batch, int* zh, hipStream_t stream);
// zh: Z is held conjugate-transposed (Z^H) when nonzero.
hipError_t wb_hqr_ms_launch(int n, cd* H, cd* Z, int zh, cd* w, int* info, void* rot_ws, int batch,
                            hipStream_t stream);
size_t wb_assemble_workspace_bytes(int L, int Ny, int Nx, int hu, int hv);
hipError_t wb_assemble_li_launch(const cd* eps, int L, int Ny, int Nx, int hu, int hv, cd* epx,
                                 cd* epy, cd* epz_i, void* ws, int* info_out, hipStream_t stream);
size_t wb_assemble_laurent_workspace_bytes(int L, int Ny, int Nx, int hu, int hv);
hipError_t wb_assemble_laurent_launch(const cd* eps, int L, int Ny, int Nx, int hu, int hv, cd* epx,
                                      cd* epy, cd* epz_i, void* ws, int* info, hipStream_t stream);
size_t wb_assemble_laurent_adjoint_workspace_bytes(int L, int Ny, int Nx, int hu, int hv);
hipError_t wb_assemble_laurent_adjoint_launch(const cd* eps, int L, int Ny, int Nx, int hu, int hv,
                                              const cd* epz_i, const cd* d_epx, const cd* d_epy,
                                              const cd* d_epz, cd* d_eps, void* ws, hipStream_t stream);
size_t wb_assemble_nvf_adjoint_workspace_bytes(int L, int Ny, int Nx, int hu, int hv);
hipError_t wb_assemble_nvf_adjoint_launch(const cd* eps, int L, int Ny, int Nx, int hu, int hv,
                                          const double* metric, int cst, double ux, double uy,
                                          const cd* d_epx, const cd* d_epy, const cd* d_epz,
                                          const cd* d_off, cd* d_eps, void* ws, hipStream_t stream);
size_t wb_assemble_nvf_workspace_bytes(int L, int Ny, int Nx, int hu, int hv);
// metric (host): the cell's J^-T (row-major 2 x 2) in units of one u-step, then the v- and
// u-steps in those units; cst: a one-cell axis after the lift, whose constant normal is
// (ux, uy). info: 3 L ints.
// floor: floor the singular values of [[1/eps]] on the non-passive layers (one N x N Hermitian
// Schur decomposition per layer of the call).
hipError_t wb_assemble_nvf_launch(const cd* eps, int L, int Ny, int Nx, int hu, int hv,
                                  const double* metric, int cst, double ux, double uy, int floor,
                                  cd* epx, cd* epy, cd* epz_i, cd* epxy, void* ws, int* info,
                                  hipStream_t stream);
size_t wb_assemble_adjoint_workspace_bytes(int L, int Ny, int Nx, int hu, int hv);
hipError_t wb_assemble_adjoint_launch(const cd* eps, int L, int Ny, int Nx, int hu, int hv,
                                      const cd* epx, const cd* epz_i, const cd* d_epx,
                                      const cd* d_epy, const cd* d_epz, cd* d_eps, void* ws,
                                      hipStream_t stream);

struct wb_solve_args {
  int P, L, hu, hv, mode;
  const double* pts;        // P x PT_STRIDE
  const double* thick;      // P x L
  const int* conv_idx;      // P x L, index into the conv arrays
  const cd* epx;            // C x N x N
  const cd* epy;
  const cd* epz;
  // NVF layers (2D path): the off-diagonal slots epxy / epyx (C x N x N each, gathered by
  // conv_idx like the others; the same array twice for a symmetric projector), or nullptr for
  // the diagonal rules (Li, Laurent), whose off-diagonal slots are zero.
  const cd* epxy = nullptr;
  const cd* epyx = nullptr;
  double* eff;              // P x 3 x 4 x N
  cd* amp;                  // P x 3 x 4 x N
  double* pe;               // P x 3
  int* flags;               // P x L: eig non-convergence / singular factorizations
  void* ck = nullptr;       // optional checkpoint for the adjoint (2D path), or nullptr
  cd* smat = nullptr;       // optional P x 4 x n x n stack scattering matrix (S11, S12, S21, S22)
  void* graze = nullptr;    // 2D: scratch of the grazing-cluster rebuild (wb_graze_bytes), or
                            // nullptr for no rebuild
  // Split solve. modes_out: compute the layer modes only and write them raw (unrotated W, V and
  // q, pair b at b n^2 / b n) to modes_W / modes_V / modes_q, then return. modes_in: take the
  // modes of pair b from index modes_idx[b] (or b) of those arrays instead of computing them.
  int modes_out = 0, modes_in = 0;
  cd* modes_W = nullptr;
  cd* modes_V = nullptr;
  cd* modes_q = nullptr;
  const int* modes_idx = nullptr;
  // Field state: each pair's mode amplitudes for the public incidence (pair b at b 2n: the
  // down-going ones anchored at the layer's top face, then the up-going ones anchored at its
  // bottom face), from the cascade's gap waves; field_ws the scratch of wb_field_extra_bytes.
  // Not combined with smat.
  cd* field_amps = nullptr;
  void* field_ws = nullptr;
  // Ragged solve (with modes_in only; no ck, smat, field_amps or graze): the points' layer counts
  // differ. steps[l] (host memory, l < L) is the number of points with more than l layers: the
  // points are ordered by non-increasing layer count, steps[0] = P and L is the largest count
  // (at most WB_RAGGED_MAX). Pair (p, l) is numbered layer-major, base_l + p with base_l =
  // steps[0] + ... + steps[l-1], so each cascade step is a prefix of the points; thick,
  // modes_idx and flags follow that numbering. ragged_tab: B + P ints of device scratch.
  const int* steps = nullptr;
  int* ragged_tab = nullptr;
};
#define WB_RAGGED_MAX 64
size_t wb_solve_workspace_bytes(int P, int L, int hu, int hv, int mode);
// The solve workspace for B pairs in all (a ragged solve).
size_t wb_solve_workspace_bytes_b(int P, int B, int hu, int hv, int mode);
hipError_t wb_solve_launch(const wb_solve_args& a, void* ws, hipStream_t st);
