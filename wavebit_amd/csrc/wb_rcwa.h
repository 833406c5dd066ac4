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

constexpr double WB_PI = 3.141592653589793;
constexpr double WB_HALF_PI = 1.5707963267948966;

__device__ inline cd boundary_kz(cd eps, cd kpar2) {
  cd v = csqrt_p(csub(eps, kpar2));
  if (v.x + v.y < 0.0) v = cneg(v);
  return v;
}

// Real-angle order azimuth (kx, ky real): atan with the half-turn on Re(kx) < 0, the
// +-pi/2 - kx/ky column and the origin rule; returns (cos, sin).
__device__ inline void order_frame(double kx, double ky, double phi, double* c, double* s) {
  double varphi;
  if (kx == 0.0) {
    if (ky == 0.0) {
      double sp, cp;
      sincos(phi, &sp, &cp);
      varphi = atan(sp / cp) + (cp < 0.0 ? WB_PI : 0.0);
    } else {
      varphi = (ky < 0.0 ? -WB_HALF_PI : WB_HALF_PI) - kx / ky;
    }
  } else {
    varphi = atan(ky / kx) + (kx < 0.0 ? WB_PI : 0.0);
  }
  sincos(varphi, s, c);
}

__device__ inline cd pt_c(const double* p, int off) { return mkc(p[off], p[off + 1]); }

// The per-order quantities of one point (order r): its wavevector, the half-spaces' kz, the gap
// admittance y and the order frame into o (the point's OV_COUNT x N record), and the diagonal
// half-space -> gap interface blocks into d (the point's DG_COUNT x n record).
__device__ inline void point_order_setup(const double* pt, int hu, int hv, int mode, int r, cd* o,
                                         cd* d) {
  const int N = (2 * hu + 1) * (2 * hv + 1);
  const int n = (mode == 0) ? 2 * N : N;
  double kx, ky;
  order_k(pt, hu, hv, mode, r, &kx, &ky);
  const cd kpar2 = mkc(kx * kx + ky * ky, 0.0);
  // kernel top = scene bottom (substrate), kernel bottom = scene top (superstrate)
  const cd e_kt = pt_c(pt, PT_EBOT), e_kb = pt_c(pt, PT_ETOP);
  const cd n_kt = csqrt_p(e_kt), n_kb = csqrt_p(e_kb);
  const cd kz_kt = boundary_kz(cmul(n_kt, n_kt), kpar2);
  const cd kz_kb = boundary_kz(cmul(n_kb, n_kb), kpar2);
  const double y = sqrt(1.0 + (kx * kx + ky * ky));
  double fc = 1.0, fs = 0.0;
  if (mode == 0) order_frame(kx, ky, pt[PT_PHI], &fc, &fs);
  o[OV_KX * N + r] = mkc(kx, 0.0);
  o[OV_KY * N + r] = mkc(ky, 0.0);
  o[OV_KZT * N + r] = kz_kt;
  o[OV_KZB * N + r] = kz_kb;
  o[OV_Y * N + r] = mkc(y, 0.0);
  o[OV_FC * N + r] = mkc(fc, 0.0);
  o[OV_FS * N + r] = mkc(fs, 0.0);
  // half-space -> gap interface blocks, diagonal: P = Wg^-1 Wh, Q = Vg^-1 Vh
  for (int side = 0; side < 2; ++side) {
    const cd kz = side == 0 ? kz_kt : kz_kb;
    const cd e = side == 0 ? e_kt : e_kb;
    const int halves = (mode == 0) ? 2 : 1;
    for (int h = 0; h < halves; ++h) {
      cd Pd, Qd;
      if (mode == 0) {
        if (h == 0) { Pd = mkc(1.0, 0.0); Qd = cscale(kz, 1.0 / y); }
        else { Pd = cdiv(kz, cscale(e, y)); Qd = mkc(1.0, 0.0); }
      } else if (mode == 1) {
        Pd = mkc(1.0, 0.0); Qd = cscale(kz, 1.0 / y);
      } else {
        Pd = mkc(1.0, 0.0); Qd = cdiv(kz, cscale(e, y));
      }
      const cd A = cadd(Pd, Qd), B = csub(Pd, Qd);
      const cd BA = cdiv(B, A);
      const cd two_A = cdiv(mkc(2.0, 0.0), A);
      const cd half = cscale(csub(A, cmul(B, BA)), 0.5);
      const int idx = h * N + r;
      if (side == 0) {   // S_ref (start port, scene bottom)
        d[DG_R11 * n + idx] = cneg(BA);
        d[DG_R12 * n + idx] = two_A;
        d[DG_R21 * n + idx] = half;
        d[DG_R22 * n + idx] = BA;
      } else {           // S_trn (end port, scene top)
        d[DG_T11 * n + idx] = BA;
        d[DG_T12 * n + idx] = half;
        d[DG_T21 * n + idx] = two_A;
        d[DG_T22 * n + idx] = cneg(BA);
      }
    }
  }
}

// Efficiencies, amplitudes and power entering of point p from its reflected / transmitted basis
// columns cr, ct (n x nrhs), its per-order record o and its parameters; called by all NTB
// threads of a workgroup (red: 3 x NTB doubles of LDS).
// eff: [P][3][4][N] doubles (R_s, R_p, T_s, T_p); amp: [P][3][4][N] complex (r_s, r_p, t_s, t_p);
// pe: [P][3]. 1D fills variant 0 only.
template <int NTB>
__device__ inline void point_efficiency(int p, int hu, int hv, int mode, const double* pts,
                                        const cd* o, const cd* cr, const cd* ct, double* eff,
                                        cd* amp, double* pe, double* red) {
  const int nu = 2 * hu + 1, nv = 2 * hv + 1, N = nu * nv;
  const int n = (mode == 0) ? 2 * N : N;
  const int nrhs = (mode == 0) ? 2 : 1;
  const double* pt = pts + (long long)p * PT_STRIDE;
  // incidence side ("ref") is the scene top = kernel bottom
  const cd e_ref = pt_c(pt, PT_ETOP), e_trn = pt_c(pt, PT_EBOT);
  const cd n_ref = csqrt_p(pt_c(pt, PT_ETOP)), n_trn = csqrt_p(pt_c(pt, PT_EBOT));
  const int c0 = N / 2;
  const cd kz_c = o[OV_KZB * N + c0];
  const int t = threadIdx.x;
  double acc[3] = {0.0, 0.0, 0.0};
  if (mode == 0) {
    const cd as = pt_c(pt, PT_AS), ap = pt_c(pt, PT_AP);
    // the three incidence variants as (alpha, beta) weights of the (s, p) basis columns
    const cd mi_n = cmul(mkc(0.0, -1.0), n_ref);
    const double cos_half_pi = cos(WB_HALF_PI), sin_half_pi = sin(WB_HALF_PI);
    cd al[3] = {as, mkc(sin_half_pi, 0.0), mkc(0.0, 0.0)};
    cd be[3] = {cmul(mi_n, ap), cmul(mi_n, mkc(cos_half_pi, 0.0)), cmul(mi_n, mkc(1.0, 0.0))};
    const double s_power = kz_c.x;
    const double p_power = cnorm2(n_ref) * cdiv(kz_c, e_ref).x;
    double den[3] = {cnorm2(as) * s_power + cnorm2(ap) * p_power, s_power, p_power};
    for (int r = t; r < N; r += NTB) {
      const int mi = r % nu - hu, ni = r / nu - hv;
      const double par = ((mi + ni) & 1) ? -1.0 : 1.0;
      const cd kzr = o[OV_KZB * N + r], kzt = o[OV_KZT * N + r];
      const double adm_s_ref = kzr.x, adm_p_ref = cdiv(kzr, e_ref).x;
      const double adm_s_trn = kzt.x, adm_p_trn = cdiv(kzt, e_trn).x;
      for (int v = 0; v < 3; ++v) {
        const cd Rs = cadd(cmul(cr[r * 2 + 0], al[v]), cmul(cr[r * 2 + 1], be[v]));
        const cd Rp = cneg(cadd(cmul(cr[(N + r) * 2 + 0], al[v]), cmul(cr[(N + r) * 2 + 1], be[v])));
        const cd Ts = cadd(cmul(ct[r * 2 + 0], al[v]), cmul(ct[r * 2 + 1], be[v]));
        const cd Tp = cadd(cmul(ct[(N + r) * 2 + 0], al[v]), cmul(ct[(N + r) * 2 + 1], be[v]));
        double* ef = eff + (((long long)p * 3 + v) * 4) * N;
        ef[0 * N + r] = cnorm2(Rs) * adm_s_ref / den[v];
        ef[1 * N + r] = cnorm2(Rp) * adm_p_ref / den[v];
        ef[2 * N + r] = cnorm2(Ts) * adm_s_trn / den[v];
        ef[3 * N + r] = cnorm2(Tp) * adm_p_trn / den[v];
        cd* am = amp + (((long long)p * 3 + v) * 4) * N;
        am[0 * N + r] = cscale(Rs, par);
        am[1 * N + r] = cscale(cdiv(Rp, mi_n), par);
        am[2 * N + r] = cscale(Ts, par);
        am[3 * N + r] = cscale(cdiv(Tp, cmul(mkc(0.0, -1.0), n_trn)), par);
        // net power at the lit face: incident column entries are nonzero at the specular order
        const cd inc_s = (r == c0) ? al[v] : mkc(0.0, 0.0);
        const cd inc_p = (r == c0) ? be[v] : mkc(0.0, 0.0);
        const cd ps = cmul(cadd(inc_s, Rs), cconj(cmul(kzr, csub(inc_s, Rs))));
        const cd pp = cmul(cadd(inc_p, Rp), cconj(cmul(cdiv(kzr, e_ref), csub(inc_p, Rp))));
        acc[v] += (ps.x + pp.x) / den[v];
      }
    }
  } else {
    const int tm = (mode == 2);
    const double sig_c = (o[OV_KX * N + c0].x < 0.0) ? -1.0 : 1.0;
    const cd adm_c = tm ? cdiv(kz_c, e_ref) : kz_c;
    const double incoming = adm_c.x;
    for (int r = t; r < N; r += NTB) {
      const int mi = r % nu - hu;
      const double par = (mi & 1) ? -1.0 : 1.0;
      const double sig = (o[OV_KX * N + r].x < 0.0) ? -1.0 : 1.0;
      const cd kzr = o[OV_KZB * N + r], kzt = o[OV_KZT * N + r];
      const cd adm_ref = tm ? cdiv(kzr, e_ref) : kzr;
      const cd adm_trn = tm ? cdiv(kzt, e_trn) : kzt;
      const cd R = cr[r], T = ct[r];
      double* ef = eff + ((long long)p * 3 * 4) * N;
      cd* am = amp + ((long long)p * 3 * 4) * N;
      const double er = cnorm2(R) * adm_ref.x / incoming;
      const double et = cnorm2(T) * adm_trn.x / incoming;
      const cd zero = mkc(0.0, 0.0);
      if (!tm) {
        ef[0 * N + r] = er; ef[1 * N + r] = 0.0; ef[2 * N + r] = et; ef[3 * N + r] = 0.0;
        am[0 * N + r] = cscale(R, sig * par); am[1 * N + r] = zero;
        am[2 * N + r] = cscale(T, sig * par); am[3 * N + r] = zero;
      } else {
        ef[0 * N + r] = 0.0; ef[1 * N + r] = er; ef[2 * N + r] = 0.0; ef[3 * N + r] = et;
        am[0 * N + r] = zero; am[1 * N + r] = cscale(R, sig * par);
        am[2 * N + r] = zero; am[3 * N + r] = cscale(cmul(T, cdiv(n_ref, n_trn)), sig * par);
      }
      const cd inc = (r == c0) ? mkc(sig_c, 0.0) : mkc(0.0, 0.0);
      const cd pw = cmul(cadd(inc, R), cconj(cmul(adm_ref, csub(inc, R))));
      acc[0] += pw.x / incoming;
    }
  }
  for (int v = 0; v < 3; ++v) red[v * NTB + t] = acc[v];
  __syncthreads();
  for (int s = NTB / 2; s > 0; s >>= 1) {
    if (t < s)
      for (int v = 0; v < 3; ++v) red[v * NTB + t] += red[v * NTB + t + s];
    __syncthreads();
  }
  if (t == 0)
    for (int v = 0; v < 3; ++v) pe[p * 3 + v] = red[v * NTB];
}

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

// Fused plain solve of 1D points with at most 32 harmonics (wb_small.hip): one workgroup per
// point, the factors it keeps for its back-sweep in the solve workspace W.
int wb_small_1d_eligible(const wb_solve_args& a);
hipError_t wb_small_1d_launch(const wb_solve_args& a, const wb_solve_ws& W, BatchPtr epx, BatchPtr epy,
                              BatchPtr epz, hipStream_t st);

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
  cd* d_epxy = nullptr;       // the NVF off-diagonal slots' cotangents (with fwd->epxy / epyx)
  cd* d_epyx = nullptr;
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
