// Layout shared by the RCWA batch solve and its adjoint: the per-point parameter record, the
// per-order vectors and diagonal port blocks kept per point, the workspace carve of a solve
// and the checkpoint a solve keeps for its adjoint.
#pragma once
#include "wb_internal.h"
#include "../../include/wavebit_amd.h"   // WB_FLAG_* bits

// A layer mode's q from its eigenvalue w = q^2: the root with Re(q) >= 0, Im(q) <= 0 on the flat
// tie |Re(q)| <= 1e-12 (reference _common/conventions.py:186-250, mode_q; the SMM root).
__device__ inline cd mode_q(cd w) {
  const cd v = csqrt_p(w);
  const bool flip = (fabs(v.x) > 1e-12) ? (v.x < 0.0) : (v.y > 0.0);
  return flip ? cneg(v) : v;
}

// per-point parameter record (doubles), set by the host
enum {
  PT_WL = 0, PT_KX0 = 1, PT_KY0 = 2, PT_PHI = 3, PT_PX = 4, PT_PY = 5,
  PT_ETOP = 6, PT_EBOT = 8, PT_AS = 10, PT_AP = 12, PT_A1Y = 14, PT_A2X = 15, PT_STRIDE = 16
};
// The cell: PT_PX / PT_PY are a1x / a2y of the lattice vectors a1 = (a1x, a1y), a2 = (a2x, a2y)
// (the periods of a rectangular cell), PT_A1Y / PT_A2X its skew entries (zero when rectangular).

// The normalized in-plane wavevector of order r (u fastest, r = (n + hv) nu + m + hu) of a point:
// order (m, n) sits at k_inc + m b1 + n b2 (reference lattice.py:899-944), on a rectangular cell
// kx0 + m wl/Lx and ky0 + n wl/Ly, on a skewed one the closed form over det(a1, a2); the 1D
// paths have ky = 0.
__device__ inline void order_k(const double* pt, int hu, int hv, int mode, int r, double* kx,
                               double* ky) {
  const int nu = 2 * hu + 1;
  const int mi = r % nu - hu, ni = r / nu - hv;
  const double wl = pt[PT_WL];
  const double a1x = pt[PT_PX], a2y = pt[PT_PY], a1y = pt[PT_A1Y], a2x = pt[PT_A2X];
  if (a1y == 0.0 && a2x == 0.0) {
    *kx = pt[PT_KX0] + (double)mi * (wl / a1x);
    *ky = (mode == 0) ? pt[PT_KY0] + (double)ni * (wl / a2y) : 0.0;
  } else {
    const double lam = wl / (a1x * a2y - a2x * a1y);
    *kx = pt[PT_KX0] + lam * (a2y * (double)mi - a1y * (double)ni);
    *ky = pt[PT_KY0] + lam * (-a2x * (double)mi + a1x * (double)ni);
  }
}
// per-order vectors kept per point: kx, ky, kz_kt, kz_kb, gap y, Fc, Fs  (7 x N complex)
enum { OV_KX = 0, OV_KY, OV_KZT, OV_KZB, OV_Y, OV_FC, OV_FS, OV_COUNT };
// per-point diagonal blocks (n each): ref11, ref12, ref21, ref22, trn11, trn12, trn21, trn22
enum { DG_R11 = 0, DG_R12, DG_R21, DG_R22, DG_T11, DG_T12, DG_T21, DG_T22, DG_COUNT };

// Buffers of one solve's workspace. Pair (point p, layer l) is batch element p * L + l.
// After a solve: M0 holds LU(I - s11_l Gamma_l) per pair (pivots in ipiv[2], a P x n slice
// per layer), M1 the solved s12 columns Xs_l, M2 Gamma_l Xs_l, M3 X Y X, M4 A, M5 B, M6 per
// pair [Gamma_{l+1} | S12_l] (row stride 2n), G0 LU(I - t11 Gamma_L) per point (pivots in
// ipiv[3]), vec[2] / vec[3] the reflected / transmitted basis columns.
struct wb_solve_ws {
  cd* M[6];
  cd* M6;
  cd *w, *q, *tau;
  void* rotws;
  int* ipiv[4];
  int* info[4];
  cd *ov, *dg, *G0;
  cd* vec[4];
  double* sig;
  void* luws;              // scratch of the blocked many-right-hand-side solves
};
wb_solve_ws wb_solve_carve(void* ws, int P, int L, int hu, int hv, int mode);
// The same for B pairs in all (a ragged solve: B = the sum of the points' layer counts).
wb_solve_ws wb_solve_carve_b(void* ws, int P, int B, int hu, int hv, int mode);

// Checkpoint kept by a solve for its adjoint (2D path): per pair the unrotated eigenvectors W,
// Wi = (R W)^-1, Vi = (R' V)^-1 and the layer S11; per point the transmitted columns d_l,
// l = 0..L, of the back-sweep (d_L = u at the top interface).
struct wb_ck {
  cd *W, *Wi, *Vi, *S11, *d;
};
size_t wb_checkpoint_bytes(int P, int L, int hu, int hv, int mode);
wb_ck wb_ck_carve(void* ck, int P, int L, int hu, int hv, int mode);

// Grazing-cluster rebuild (wb_graze.hip): scratch of a batch of B pairs at mode order n, the
// stage run after the layer eigensolve and the patch run after the s/p rotation.
constexpr int WB_GRAZE_KMAX = 8;        // largest grazing cluster rebuilt (modes)
constexpr int WB_GRAZE_NMAX = 2048;     // largest mode order n = 2N rebuilt
enum { WB_GRAZE_NONE = 0, WB_GRAZE_KEPT, WB_GRAZE_ACTIVE, WB_GRAZE_REBUILT, WB_GRAZE_DEFECTIVE };
struct wb_graze_ws {
  int *jn, *status, *einfo, *sinfo, *jidx;
  double *wmax, *pad;
  cd *QE, *QH, *TE, *TH, *WC, *VC;          // n x K per pair
  cd *Rr, *Lr, *GE, *GH;                    // K x K per pair
  cd *Ms, *gv, *Zs, *Xs;                    // 2K x 2K per pair
  cd *mu, *taus, *qc;
  void* rot;
};
size_t wb_graze_bytes(int B, int n);
wb_graze_ws wb_graze_carve(void* ws, int B, int n);
// offd: the NVF off-diagonal slots epxy / epyx enter Omega_R (else they are zero and unread).
hipError_t wb_graze_stage(int N, int L, int B, const cd* w, const cd* W, BatchPtr epx,
                          BatchPtr epy, BatchPtr epz, BatchPtr epxy, BatchPtr epyx, int offd,
                          const cd* ov, cd* olr, cd* H, cd* Z, cd* X,
                          cd* Vh, cd* wh, cd* tau, void* rotws, const wb_graze_ws& g,
                          hipStream_t st);
hipError_t wb_graze_patch(int N, int L, int B, const cd* ov, cd* bigW, cd* bigV, cd* q,
                          const wb_graze_ws& g, int* flags, hipStream_t st);
hipError_t wb_graze_patch_raw(int n, int B, cd* W, cd* V, cd* q, const wb_graze_ws& g, int* flags,
                              hipStream_t st);

size_t wb_adjoint_workspace_bytes(int P, int L, int hu, int hv, int mode);
struct wb_adjoint_args {
  const wb_solve_args* fwd;   // the same arguments as the checkpointed solve
  int n_conv;                 // number of conv slots the gather index addresses
  const void* solve_ws;       // that solve's workspace, untouched since
  const void* ck;             // that solve's checkpoint
  const double* eff_bar;      // P x 3 x 4 x N cotangent of the efficiencies
  cd* d_epx;                  // n_conv x N x N (overwritten)
  cd* d_epy;
  cd* d_epz;
  double* d_thick;            // P x L (overwritten)
};
hipError_t wb_adjoint_launch(const wb_adjoint_args& a, void* ws, hipStream_t st);

// Field state (wb_fields.hip): scratch beyond the solve workspace that keeps every split plane's
// down-going gap columns and the per-pair amplitude partials, and the stage run after the
// cascade's back-sweep.
size_t wb_field_extra_bytes(int P, int L, int n, int nrhs);
struct wb_field_ws {
  cd *d, *u, *t;            // P x (L+1) planes of n x nrhs (down, up), B x 3 blocks of n x nrhs
};
wb_field_ws wb_field_carve(void* ws, int P, int L, int n, int nrhs);
// Gamma_0 at G0 (P x n x n), Gamma_l (l >= 1) the S11 half of M6's slot l - 1 (row stride 2n),
// A / B per pair in M4 / M5, q per pair; writes a.field_amps.
hipError_t wb_field_amps_stage(const wb_solve_args& a, int N, int n, int nrhs, const cd* G0,
                               const cd* M6, const cd* A, const cd* B, const cd* q,
                               const wb_field_ws& F, hipStream_t st);
// Fields of P points x L layers at nz local heights per layer (z: L x nz, device, within each
// layer's [0, d]) and nxy centre-origin points xy (device, x y pairs): out (P x L nz x nxy x 6)
// = scale[c] x component c. Pair b = p L + l takes its modes at modes_idx[b] (or b) of W, V
// (n x n, unrotated) and q (n), its amplitudes at b of amps (2n: down-going, then up-going), its
// inverse-rule matrix at conv_idx[b] (or b) of epz (N x N, unused on the TE path) and its
// thickness at thick[b]; scale is a host array of 6.
size_t wb_field_sample_bytes(int P, int L, int hu, int hv, int mode, int nz, int nxy);
hipError_t wb_field_sample_launch(int P, int L, int hu, int hv, int mode, const double* pts,
                                  const double* thick, const cd* W, const cd* V, const cd* q,
                                  const int* modes_idx, const cd* epz, const int* conv_idx,
                                  const cd* amps, int nz, const double* z, int nxy, const double* xy,
                                  const cd* scale, cd* out, void* ws, hipStream_t st);
