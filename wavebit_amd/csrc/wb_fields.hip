// In-layer field reconstruction on one MI355X: the per-layer mode amplitudes a solve leaves in
// its cascade, and the six Cartesian components they give at sample points of one layer.
//
// Reference arithmetic restated here (paths under the reference's src/wavebit/):
//   amplitudes  rcwa/_engine/scattering.py:221-333 (_below_reflections, _field_state): with
//               d_l the down-going gap wave at split plane l (plane 0 above the substrate, plane
//               L below the superstrate, the cascade's own back-sweep) and u_l = Gamma_l d_l the
//               up-going one, layer l's down-going modes are (A d_{l+1} + B u_{l+1}) / 2 at its
//               top face and its up-going ones (B d_l + A u_l) / 2 at its bottom face, or, where
//               |k0 q d| < 1, (B d_{l+1} + A u_{l+1}) / 2 at the top face carried down by
//               exp(k0 q d) (modal_gradient.py:1115-1175 two_face, scattering.py:233 _TOP_FACE);
//               the basis columns combined into the public incidence: a_s s + a_p (-i n) p on the
//               coupled path, n times the unit port variable on the pure TM one;
//   components  rcwa/_engine/fields.py:503-630 (field_1d), :630-814 (_field_coupled), in the
//               frame the cascade runs in (+z up, the public one): M = W (a_up + a_down),
//               U = V (a_down - a_up);
//                 2D  (Mx, My, -i epz_i (Kx Uy - Ky Ux), i Ux, i Uy, Kx My - Ky Mx)
//                 TE  (0, M, 0, i U, 0, Kx M)     TM  (-i U, 0, -epz_i Kx M, 0, M, 0)
//               H in units of E; the caller's per-component factors (1/Z0 for H, the source
//               amplitude) are applied with the centre parity (-1)^(m+n) of the coefficients
//               (field_result.py:673-679);
//   expansion   rcwa/_engine/field_result.py:109-114 (expand_spectrum): a GEMM of the
//               exp(i k0 (x kx + y ky)) matrix of the sample points against the coefficients.
#include "wb_rcwa.h"

namespace {

constexpr double PI = 3.141592653589793;

inline size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }
inline dim3 grid2(long long elems, int batch) {
  return dim3((unsigned)((elems + 255) / 256), (unsigned)batch);
}


// Combine the basis columns of the three partials of pair b into the public incidence and pick
// the face each up-going mode is read from. One thread per (row, pair).
__global__ void k_field_amps(int n, int L, int nrhs, int mode, const double* pts,
                             const double* thick, const cd* q, const cd* t, cd* amps) {
  const long long b = blockIdx.y;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const long long p = b / L;
  const double* pt = pts + p * PT_STRIDE;
  const cd n_inc = csqrt_p(pt_c(pt, PT_ETOP));
  cd w[2];
  if (nrhs == 2) {
    w[0] = pt_c(pt, PT_AS);
    w[1] = cmul(mkc(n_inc.y, -n_inc.x), pt_c(pt, PT_AP));     // -i n a_p
  } else {
    w[0] = (mode == 2) ? n_inc : mkc(1.0, 0.0);
    w[1] = mkc(0.0, 0.0);
  }
  const long long vs = (long long)n * nrhs;
  cd part[3];
  for (int k = 0; k < 3; ++k) {
    const cd* tk = t + (b * 3 + k) * vs + (long long)r * nrhs;
    cd acc = cmul(w[0], tk[0]);
    if (nrhs == 2) acc = cadd(acc, cmul(w[1], tk[1]));
    part[k] = acc;
  }
  const double k0 = 2.0 * PI / pt[PT_WL];
  const cd kqd = cscale(q[b * n + r], k0 * thick[b]);
  const cd up = (cabs(kqd) < 1.0) ? cmul(cexp_(kqd), part[2]) : part[1];
  amps[b * 2 * n + r] = part[0];
  amps[b * 2 * n + n + r] = up;
}

// Propagated amplitudes at the samples of every pair: Pp = a_up + a_down, Pm = a_down - a_up
// (pair b: n x nz, row = mode, column = sample j of layer l = b % L at local height z[l nz + j]).
__global__ void k_field_prop(int n, int L, int nz, const double* pts, const double* thick, const cd* q,
                             const int* modes_idx, const cd* amps, const double* z, cd* Pp, cd* Pm) {
  const long long b = blockIdx.y;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n * nz) return;
  const int i = (int)(e / nz), j = (int)(e % nz);
  const long long p = b / L;
  const int l = (int)(b % L);
  const double k0 = 2.0 * PI / pts[p * PT_STRIDE + PT_WL];
  const long long mb = modes_idx ? modes_idx[b] : b;
  const cd qi = q[mb * n + i];
  const double zj = z[(long long)l * nz + j];
  const double d = thick[b];
  const cd up = cmul(cexp_(cscale(qi, -k0 * zj)), amps[b * 2 * n + n + i]);
  const cd down = cmul(cexp_(cscale(qi, k0 * (zj - d))), amps[b * 2 * n + i]);
  const long long o = b * (long long)n * nz + e;
  Pp[o] = cadd(up, down);
  Pm[o] = csub(down, up);
}

// The longitudinal partial that epz_i multiplies: 2D Kx Uy - Ky Ux, TM Kx M (pair b: N x nz).
__global__ void k_field_tz(int hu, int hv, int mode, int L, int nz, const double* pts, const cd* S,
                           const cd* U, cd* T) {
  const int N = (2 * hu + 1) * (2 * hv + 1);
  const int n = (mode == 0) ? 2 * N : N;
  const long long b = blockIdx.y;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)N * nz) return;
  const int o = (int)(e / nz), j = (int)(e % nz);
  double kx, ky;
  order_k(pts + (b / L) * PT_STRIDE, hu, hv, mode, o, &kx, &ky);
  const cd* Sb = S + b * (long long)n * nz;
  const cd* Ub = U + b * (long long)n * nz;
  cd v;
  if (mode == 0) v = csub(cscale(Ub[(long long)(N + o) * nz + j], kx), cscale(Ub[(long long)o * nz + j], ky));
  else v = cscale(Sb[(long long)o * nz + j], kx);
  T[b * (long long)N * nz + e] = v;
}

struct Scale6 { cd c[6]; };

// Coefficients of the six components with the centre parity of each order and the caller's
// component factors: point p's block is N x 6 L nz, column (l nz + j) 6 + component.
__global__ void k_field_coef(int hu, int hv, int mode, int L, int nz, const double* pts, const cd* S,
                             const cd* U, const cd* Ez, Scale6 f, cd* C) {
  const int nu = 2 * hu + 1;
  const int N = nu * (2 * hv + 1);
  const int n = (mode == 0) ? 2 * N : N;
  const long long b = blockIdx.y;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)N * nz) return;
  const int o = (int)(e / nz), j = (int)(e % nz);
  const long long p = b / L;
  const int l = (int)(b % L);
  double kx, ky;
  order_k(pts + p * PT_STRIDE, hu, hv, mode, o, &kx, &ky);
  const int m = o % nu - hu, nn = o / nu - hv;
  const double par = ((m + nn) & 1) ? -1.0 : 1.0;
  const cd* Sb = S + b * (long long)n * nz;
  const cd* Ub = U + b * (long long)n * nz;
  const cd zero = mkc(0.0, 0.0);
  cd v[6];
  if (mode == 0) {
    const cd mx = Sb[(long long)o * nz + j], my = Sb[(long long)(N + o) * nz + j];
    const cd ux = Ub[(long long)o * nz + j], uy = Ub[(long long)(N + o) * nz + j];
    const cd ez = Ez[b * (long long)N * nz + e];
    v[0] = mx;
    v[1] = my;
    v[2] = mkc(ez.y, -ez.x);                        // -i epz_i (Kx Uy - Ky Ux)
    v[3] = mkc(-ux.y, ux.x);                        // i Ux
    v[4] = mkc(-uy.y, uy.x);                        // i Uy
    v[5] = csub(cscale(my, kx), cscale(mx, ky));
  } else if (mode == 1) {
    const cd mm = Sb[(long long)o * nz + j], u = Ub[(long long)o * nz + j];
    v[0] = zero; v[1] = mm; v[2] = zero;
    v[3] = mkc(-u.y, u.x);                          // i U
    v[4] = zero;
    v[5] = cscale(mm, kx);
  } else {
    const cd mm = Sb[(long long)o * nz + j], u = Ub[(long long)o * nz + j];
    const cd ez = Ez[b * (long long)N * nz + e];
    v[0] = mkc(u.y, -u.x);                          // -i U
    v[1] = zero;
    v[2] = cneg(ez);                                // -epz_i Kx M
    v[3] = zero; v[4] = mm; v[5] = zero;
  }
  const long long cols = 6LL * L * nz;
  cd* c = C + p * (long long)N * cols + (long long)o * cols + ((long long)l * nz + j) * 6;
  for (int k = 0; k < 6; ++k) c[k] = cscale(cmul(f.c[k], v[k]), par);
}

// E_p (nxy x N): exp(i k0 (x kx + y ky)) at centre-origin sample points, per point p.
__global__ void k_field_phase(int hu, int hv, int mode, int nxy, const double* pts, const double* xy,
                              cd* E) {
  const int N = (2 * hu + 1) * (2 * hv + 1);
  const long long p = blockIdx.y;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)nxy * N) return;
  const int s = (int)(e / N), o = (int)(e % N);
  const double* pt = pts + p * PT_STRIDE;
  double kx, ky;
  order_k(pt, hu, hv, mode, o, &kx, &ky);
  const double k0 = 2.0 * PI / pt[PT_WL];
  const double ph = k0 * (xy[2 * s] * kx + xy[2 * s + 1] * ky);
  double sn, cs;
  sincos(ph, &sn, &cs);
  E[p * (long long)nxy * N + e] = mkc(cs, sn);
}

// out_p (L nz x nxy x 6) from F_p (nxy x 6 L nz).
__global__ void k_field_permute(int rows, int nxy, const cd* F, cd* out) {
  const long long p = blockIdx.y;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long per = (long long)rows * nxy * 6;
  if (e >= per) return;
  const int c = (int)(e % 6);
  const long long rest = e / 6;
  const int s = (int)(rest % nxy), j = (int)(rest / nxy);
  out[p * per + e] = F[p * per + (long long)s * 6 * rows + (long long)j * 6 + c];
}

struct SampleWs {
  cd *Pp, *Pm, *S, *U, *T, *Ez, *C, *E, *F;
};

SampleWs sample_carve(void* ws, int P, int L, int N, int n, int nz, int nxy) {
  char* p = (char*)ws;
  auto take = [&](size_t elems) { char* r = p; p += align_up(elems * sizeof(cd)); return (cd*)r; };
  const size_t B = (size_t)P * L;
  SampleWs w;
  w.Pp = take(B * n * nz);
  w.Pm = take(B * n * nz);
  w.S = take(B * n * nz);
  w.U = take(B * n * nz);
  w.T = take(B * N * nz);
  w.Ez = take(B * N * nz);
  w.C = take(B * N * 6 * nz);
  w.E = take((size_t)P * nxy * N);
  w.F = take(B * (size_t)nxy * 6 * nz);
  return w;
}

}  // namespace

size_t wb_field_extra_bytes(int P, int L, int n, int nrhs) {
  const size_t planes = (size_t)P * (L + 1) * n * nrhs * sizeof(cd);
  const size_t parts = (size_t)P * L * 3 * n * nrhs * sizeof(cd);
  return 2 * align_up(planes) + align_up(parts) + 256;
}

wb_field_ws wb_field_carve(void* ws, int P, int L, int n, int nrhs) {
  char* p = (char*)ws;
  const size_t planes = (size_t)P * (L + 1) * n * nrhs * sizeof(cd);
  wb_field_ws f;
  f.d = (cd*)p; p += align_up(planes);
  f.u = (cd*)p; p += align_up(planes);
  f.t = (cd*)p;
  return f;
}

hipError_t wb_field_amps_stage(const wb_solve_args& a, int N, int n, int nrhs, const cd* G0,
                               const cd* M6, const cd* A, const cd* B, const cd* q,
                               const wb_field_ws& F, hipStream_t st) {
  (void)N;
  const int P = a.P, L = a.L;
  const long long vs = (long long)n * nrhs;
  const long long nn = (long long)n * n;
  const long long n2 = 2 * nn;
  const long long plane_stride = (long long)(L + 1) * vs;
  const cd one = mkc(1.0, 0.0), zero = mkc(0.0, 0.0), half = mkc(0.5, 0.0);
  hipError_t e;
  // u_l = Gamma_l d_l on every split plane
  for (int l = 0; l <= L; ++l) {
    const BatchPtr gam = (l == 0) ? bp_strided(G0, nn) : bp_strided(M6 + (long long)(l - 1) * n2, (long long)L * n2);
    const int ldg = (l == 0) ? n : 2 * n;
    if ((e = wb_zgemm_launch(n, nrhs, n, one, gam, ldg, bp_strided(F.d + l * vs, plane_stride), nrhs,
                             zero, bp_strided(F.u + l * vs, plane_stride), nrhs, P, st))) return e;
  }
  // per layer: (A d_top + B u_top)/2, (B d_bot + A u_bot)/2, (B d_top + A u_top)/2
  for (int l = 0; l < L; ++l) {
    const BatchPtr Al = bp_strided(A + (long long)l * nn, (long long)L * nn);
    const BatchPtr Bl = bp_strided(B + (long long)l * nn, (long long)L * nn);
    const BatchPtr d_bot = bp_strided(F.d + l * vs, plane_stride);
    const BatchPtr d_top = bp_strided(F.d + (l + 1) * vs, plane_stride);
    const BatchPtr u_bot = bp_strided(F.u + l * vs, plane_stride);
    const BatchPtr u_top = bp_strided(F.u + (l + 1) * vs, plane_stride);
    BatchPtr T[3];
    for (int k = 0; k < 3; ++k) T[k] = bp_strided(F.t + (long long)(l * 3 + k) * vs, (long long)L * 3 * vs);
    const BatchPtr first[3] = {Al, Bl, Bl}, second[3] = {Bl, Al, Al};
    const BatchPtr x1[3] = {d_top, d_bot, d_top}, x2[3] = {u_top, u_bot, u_top};
    for (int k = 0; k < 3; ++k) {
      if ((e = wb_zgemm_launch(n, nrhs, n, half, first[k], n, x1[k], nrhs, zero, T[k], nrhs, P, st))) return e;
      if ((e = wb_zgemm_launch(n, nrhs, n, half, second[k], n, x2[k], nrhs, one, T[k], nrhs, P, st))) return e;
    }
  }
  hipLaunchKernelGGL(k_field_amps, grid2(n, P * L), dim3(256), 0, st, n, L, nrhs, a.mode, a.pts,
                     a.thick, q, F.t, a.field_amps);
  return hipGetLastError();
}

size_t wb_field_sample_bytes(int P, int L, int hu, int hv, int mode, int nz, int nxy) {
  const size_t N = (size_t)(2 * hu + 1) * (2 * hv + 1);
  const size_t n = (mode == 0) ? 2 * N : N;
  const size_t B = (size_t)P * L;
  return 4 * align_up(B * n * nz * sizeof(cd)) + 2 * align_up(B * N * nz * sizeof(cd)) +
         align_up(B * N * 6 * nz * sizeof(cd)) + align_up((size_t)P * nxy * N * sizeof(cd)) +
         align_up(B * (size_t)nxy * 6 * nz * sizeof(cd)) + 256;
}

hipError_t wb_field_sample_launch(int P, int L, int hu, int hv, int mode, const double* pts,
                                  const double* thick, const cd* W, const cd* V, const cd* q,
                                  const int* modes_idx, const cd* epz, const int* conv_idx,
                                  const cd* amps, int nz, const double* z, int nxy, const double* xy,
                                  const cd* scale, cd* out, void* ws, hipStream_t st) {
  const int N = (2 * hu + 1) * (2 * hv + 1);
  const int n = (mode == 0) ? 2 * N : N;
  const int B = P * L;
  const long long nn = (long long)n * n;
  const SampleWs w = sample_carve(ws, P, L, N, n, nz, nxy);
  const cd one = mkc(1.0, 0.0), zero = mkc(0.0, 0.0);
  hipError_t e;
  hipLaunchKernelGGL(k_field_prop, grid2((long long)n * nz, B), dim3(256), 0, st, n, L, nz, pts, thick, q,
                     modes_idx, amps, z, w.Pp, w.Pm);
  if ((e = hipGetLastError())) return e;
  const BatchPtr Wb = modes_idx ? bp_gather(W, nn, modes_idx) : bp_strided(W, nn);
  const BatchPtr Vb = modes_idx ? bp_gather(V, nn, modes_idx) : bp_strided(V, nn);
  const long long vz = (long long)n * nz;
  if ((e = wb_zgemm_launch(n, nz, n, one, Wb, n, bp_strided(w.Pp, vz), nz, zero, bp_strided(w.S, vz),
                           nz, B, st))) return e;
  if ((e = wb_zgemm_launch(n, nz, n, one, Vb, n, bp_strided(w.Pm, vz), nz, zero, bp_strided(w.U, vz),
                           nz, B, st))) return e;
  const long long tz = (long long)N * nz;
  if (mode != 1) {
    hipLaunchKernelGGL(k_field_tz, grid2(tz, B), dim3(256), 0, st, hu, hv, mode, L, nz, pts, w.S, w.U, w.T);
    if ((e = hipGetLastError())) return e;
    const BatchPtr Zb = conv_idx ? bp_gather(epz, (long long)N * N, conv_idx) : bp_strided(epz, (long long)N * N);
    if ((e = wb_zgemm_launch(N, nz, N, one, Zb, N, bp_strided(w.T, tz), nz, zero, bp_strided(w.Ez, tz), nz,
                             B, st))) return e;
  }
  Scale6 f;
  for (int k = 0; k < 6; ++k) f.c[k] = scale[k];
  hipLaunchKernelGGL(k_field_coef, grid2(tz, B), dim3(256), 0, st, hu, hv, mode, L, nz, pts, w.S, w.U,
                     w.Ez, f, w.C);
  if ((e = hipGetLastError())) return e;
  hipLaunchKernelGGL(k_field_phase, grid2((long long)nxy * N, P), dim3(256), 0, st, hu, hv, mode, nxy, pts,
                     xy, w.E);
  if ((e = hipGetLastError())) return e;
  const int cols = 6 * L * nz;
  if ((e = wb_zgemm_launch(nxy, cols, N, one, bp_strided(w.E, (long long)nxy * N), N,
                           bp_strided(w.C, (long long)N * cols), cols, zero,
                           bp_strided(w.F, (long long)nxy * cols), cols, P, st))) return e;
  hipLaunchKernelGGL(k_field_permute, grid2((long long)L * nz * nxy * 6, P), dim3(256), 0, st, L * nz, nxy,
                     w.F, out);
  return hipGetLastError();
}
