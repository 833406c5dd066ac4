// Hand-written adjoint of the batched RCWA solve (2D coupled path) on one MI355X.
//
// Differentiates the efficiencies of a checkpointed wb_solve_launch with respect to the layer
// convolution matrices (epx, epy, epz_i, summed per conv slot) and the layer thicknesses. The
// eigenbasis W of each layer is held fixed and the modal constants enter as a matrix Q in that
// basis (value diag(q)): a perturbation dOmega2 moves Q by (W^-1 dOmega2 W) / (q_i + q_j), the
// generator the reference uses (rcwa/_engine/modal_gradient.py:161-251), so
//   Qbar -> Omega2bar = W^-H (Qbar / conj(q_i + q_j)) W^H,
// and X = exp(-k0 d Q) is differentiated through divided differences of exp on the pairs
// (q_i, q_j) (modal_gradient.py:446-470). V = Omega_R W Q^-1 (modal_gradient.py:1239-1244)
// carries the Omega_R = [[-KxKy, Kx^2 - epy], [epx - Ky^2, KyKx]] dependence.
//
// The reverse sweep mirrors the forward cascade of wb_rcwa.hip (pair (p, l) = p * L + l):
//   efficiencies -> reflected / transmitted columns -> the back-sweep d_l = Mx_l^-1 s12_l d_{l+1}
//   -> the top interface u = (I - t11 G_L)^-1 t12 e -> the reflection recursion
//   G_{l+1} = s11_l + s12_l G_l Xs_l, Xs_l = Mx_l^-1 s12_l, Mx_l = I - s11_l G_l, top layer first.
// Each layer's cotangents [S11bar | S12bar] are taken through the layer S-matrix
// (S11 = M^-1 (XYX A - B), S12 = M^-1 X (A - Y B), Y = B A^-1, M = A - XYX B) as soon as they
// are complete, one layer of P pairs at a time: the per-layer S-matrices and eigensystems are the
// checkpoint, every other intermediate of the layer is recomputed from them.
// All n x n work is batched zgemm (with conjugate-transposed operands read in place) and LU
// solves with A or A^H; the rest are elementwise kernels below.
#include "wb_rcwa.h"

namespace {

constexpr double PI = 3.141592653589793;
constexpr double HALF_PI = 1.5707963267948966;

inline dim3 grid2(long long elems, int batch) {
  return dim3((unsigned)((elems + 255) / 256), (unsigned)batch);
}


__device__ inline cd x_of(cd q, double k0d) { return cexp_(cscale(q, -k0d)); }

// Cotangents of the reflected / transmitted basis columns (n x 2 per point) from the cotangent
// of the efficiencies of all three incidence variants.
__global__ __launch_bounds__(256) void k_eff_bar(int hu, int hv, const double* pts, const cd* ov,
                                                 const cd* cref, const cd* ctrn, const double* eb,
                                                 cd* crb, cd* ctb) {
  const int p = blockIdx.x;
  const int nu = 2 * hu + 1, nv = 2 * hv + 1, N = nu * nv, n = 2 * N;
  const double* pt = pts + (long long)p * PT_STRIDE;
  const cd* o = ov + (long long)p * OV_COUNT * N;
  const cd* cr = cref + (long long)p * n * 2;
  const cd* ct = ctrn + (long long)p * n * 2;
  const cd e_ref = pt_c(pt, PT_ETOP), e_trn = pt_c(pt, PT_EBOT);
  const cd n_ref = csqrt_p(e_ref);
  const cd kz_c = o[OV_KZB * N + N / 2];
  const cd as = pt_c(pt, PT_AS), ap = pt_c(pt, PT_AP);
  const cd mi_n = cmul(mkc(0.0, -1.0), n_ref);
  const double cos_half_pi = cos(HALF_PI), sin_half_pi = sin(HALF_PI);
  const cd al[3] = {as, mkc(sin_half_pi, 0.0), mkc(0.0, 0.0)};
  const cd be[3] = {cmul(mi_n, ap), cmul(mi_n, mkc(cos_half_pi, 0.0)), cmul(mi_n, mkc(1.0, 0.0))};
  const double s_power = kz_c.x;
  const double p_power = cnorm2(n_ref) * cdiv(kz_c, e_ref).x;
  const double den[3] = {cnorm2(as) * s_power + cnorm2(ap) * p_power, s_power, p_power};
  for (int r = threadIdx.x; r < N; r += blockDim.x) {
    const cd kzr = o[OV_KZB * N + r], kzt = o[OV_KZT * N + r];
    const double adm[4] = {kzr.x, cdiv(kzr, e_ref).x, kzt.x, cdiv(kzt, e_trn).x};
    cd g[4][2];   // rows r (s part) and N + r (p part) of the reflected / transmitted columns
    for (int i = 0; i < 4; ++i) { g[i][0] = mkc(0.0, 0.0); g[i][1] = mkc(0.0, 0.0); }
    for (int v = 0; v < 3; ++v) {
      const double* e = eb + (((long long)p * 3 + v) * 4) * N;
      const cd rows[4][2] = {{cr[r * 2 + 0], cr[r * 2 + 1]},
                             {cr[(N + r) * 2 + 0], cr[(N + r) * 2 + 1]},
                             {ct[r * 2 + 0], ct[r * 2 + 1]},
                             {ct[(N + r) * 2 + 0], ct[(N + r) * 2 + 1]}};
      for (int i = 0; i < 4; ++i) {
        const double w = e[i * N + r];
        if (w == 0.0) continue;
        // efficiency = |c0 al + c1 be|^2 adm / den (the sign of R_p drops out of |.|^2)
        const cd c = cadd(cmul(rows[i][0], al[v]), cmul(rows[i][1], be[v]));
        const cd gc = cscale(c, 2.0 * w * adm[i] / den[v]);
        g[i][0] = cadd(g[i][0], cmul(gc, cconj(al[v])));
        g[i][1] = cadd(g[i][1], cmul(gc, cconj(be[v])));
      }
    }
    cd* rb = crb + (long long)p * n * 2;
    cd* tb = ctb + (long long)p * n * 2;
    rb[r * 2 + 0] = g[0][0]; rb[r * 2 + 1] = g[0][1];
    rb[(N + r) * 2 + 0] = g[1][0]; rb[(N + r) * 2 + 1] = g[1][1];
    tb[r * 2 + 0] = g[2][0]; tb[r * 2 + 1] = g[2][1];
    tb[(N + r) * 2 + 0] = g[3][0]; tb[(N + r) * 2 + 1] = g[3][1];
  }
}

// out = conj(dg[block]) (.) in (+ out when acc), n x 2 columns per point.
__global__ void k_vec_diag(int n, const cd* dg, int block, const cd* in, cd* out, int acc) {
  const long long p = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * 2) return;
  const int r = e / 2;
  cd v = cmulc(dg[p * DG_COUNT * n + block * n + r], in[p * (long long)n * 2 + e]);
  if (acc) v = cadd(v, out[p * (long long)n * 2 + e]);
  out[p * (long long)n * 2 + e] = v;
}

// dst = src, rows x cols with leading dimensions
__global__ void k_copy2(int rows, int cols, BatchPtr src, int lds, BatchPtr dst, int ldd) {
  const long long b = blockIdx.y;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)rows * cols) return;
  const int r = (int)(e / cols), c = (int)(e % cols);
  dst.ptr(b)[(long long)r * ldd + c] = src.ptr(b)[(long long)r * lds + c];
}

// Y += X (n x n with leading dimensions)
__global__ void k_add(int n, BatchPtr X, int ldx, BatchPtr Y, int ldy) {
  const long long b = blockIdx.y;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n * n) return;
  const int r = (int)(e / n), c = (int)(e % n);
  cd* y = Y.ptr(b) + (long long)r * ldy + c;
  *y = cadd(*y, X.ptr(b)[(long long)r * ldx + c]);
}

// Gamma_0 = diag(ref22) as a full matrix, per point
__global__ void k_gamma0_full(int n, const cd* dg, cd* G) {
  const long long p = blockIdx.y;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n * n) return;
  const int r = (int)(e / n), c = (int)(e % n);
  G[p * (long long)n * n + e] = (r == c) ? dg[p * DG_COUNT * n + DG_R22 * n + r] : mkc(0.0, 0.0);
}

// Layer l, pair p: Fb = conj(X) E2b, Ab = Fb + Mb, Bb = -E1b. SB = [E1b | E2b] (row stride 2n).
__global__ void k_adj_a(int n, int L, int l, const cd* SB, const cd* Mb, const cd* q,
                        const double* pts, const double* thick, cd* Fb, cd* Ab, cd* Bb) {
  const long long p = blockIdx.y;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n * n) return;
  const int r = (int)(e / n), c = (int)(e % n);
  const long long pair = p * L + l;
  const double k0d = 2.0 * PI / pts[p * PT_STRIDE + PT_WL] * thick[pair];
  const cd xr = x_of(q[pair * n + r], k0d);
  const long long o = p * (long long)n * n + e;
  const cd* sb = SB + p * 2LL * n * n + (long long)r * 2 * n;
  const cd f = cmulc(xr, sb[n + c]);
  Fb[o] = f;
  Ab[o] = cadd(f, Mb[o]);
  Bb[o] = cneg(sb[c]);
}

// Yb += conj(x_r) XYXb conj(x_c); T1 = XYXb conj(x_c) (columns); T2 = conj(x_r) XYXb (rows).
__global__ void k_adj_xyx(int n, int L, int l, const cd* XYXb, const cd* q, const double* pts,
                          const double* thick, cd* Yb, cd* T1, cd* T2) {
  const long long p = blockIdx.y;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n * n) return;
  const int r = (int)(e / n), c = (int)(e % n);
  const long long pair = p * L + l;
  const double k0d = 2.0 * PI / pts[p * PT_STRIDE + PT_WL] * thick[pair];
  const cd xr = cconj(x_of(q[pair * n + r], k0d)), xc = cconj(x_of(q[pair * n + c], k0d));
  const long long o = p * (long long)n * n + e;
  const cd v = XYXb[o];
  Yb[o] = cadd(Yb[o], cmul(cmul(xr, v), xc));
  T1[o] = cmul(v, xc);
  T2[o] = cmul(xr, v);
}

// Vib = (Ab - Bb - T3) conj(Dv) by columns; Gm = Vi / q by rows; Gbm = conj(q) Vib by rows.
// 2D gap: Dv = [-i y, 1], so conj(Dv) = [i y, 1].
__global__ void k_adj_vi(int N, int L, int l, const cd* Ab, const cd* Bb, const cd* T3,
                         const cd* Vi_ck, const cd* q, const cd* ov, cd* Vib, cd* Gm, cd* Gbm) {
  const long long p = blockIdx.y;
  const int n = 2 * N;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n * n) return;
  const int r = (int)(e / n), c = (int)(e % n);
  const long long pair = p * L + l;
  const long long o = p * (long long)n * n + e;
  cd v = csub(csub(Ab[o], Bb[o]), T3[o]);
  if (c < N) {
    const double y = ov[p * OV_COUNT * N + OV_Y * N + c].x;
    v = cmul(v, mkc(0.0, y));
  }
  Vib[o] = v;
  const cd qr = q[pair * n + r];
  Gm[o] = cdiv(Vi_ck[pair * (long long)n * n + e], qr);
  Gbm[o] = cmulc(qr, v);
}

// Qb += Xb (.) conj(divided differences of exp(s z) on (q_r, q_c)), s = -k0 d; then
// Qb /= conj(q_r + q_c) (the generator's Sylvester divide), in place.
__global__ void k_adj_expm(int n, int L, int l, const cd* Xb, const cd* q, const double* pts,
                           const double* thick, cd* Qb) {
  const long long p = blockIdx.y;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n * n) return;
  const int r = (int)(e / n), c = (int)(e % n);
  const long long pair = p * L + l;
  const double s = -2.0 * PI / pts[p * PT_STRIDE + PT_WL] * thick[pair];
  const cd qi = q[pair * n + r], qj = q[pair * n + c];
  const cd fi = cexp_(cscale(qi, s)), fj = cexp_(cscale(qj, s));
  const cd dq = csub(qi, qj);
  const cd z = cscale(dq, s);
  cd dd;
  if (cabs(z) < 1e-3) {
    // f(q_j) s (e^z - 1) / z by its series to z^4
    const cd z2 = cmul(z, z), z3 = cmul(z2, z), z4 = cmul(z3, z);
    cd ser = cadd(mkc(1.0, 0.0), cscale(z, 0.5));
    ser = cadd(ser, cscale(z2, 1.0 / 6.0));
    ser = cadd(ser, cscale(z3, 1.0 / 24.0));
    ser = cadd(ser, cscale(z4, 1.0 / 120.0));
    dd = cscale(cmul(fj, ser), s);
  } else {
    dd = cdiv(csub(fi, fj), dq);
  }
  const long long o = p * (long long)n * n + e;
  const cd qb = cadd(Qb[o], cmulc(dd, Xb[o]));
  Qb[o] = cdiv(qb, cconj(cadd(qi, qj)));
}

// d(thickness): Re sum_i conj(-k0 q_i x_i) Xb_ii, one workgroup per point.
__global__ __launch_bounds__(256) void k_thick_bar(int n, int L, int l, const cd* Xb, const cd* q,
                                                   const double* pts, const double* thick,
                                                   double* d_thick) {
  const long long p = blockIdx.x;
  const long long pair = p * L + l;
  const double k0 = 2.0 * PI / pts[p * PT_STRIDE + PT_WL];
  const double d = thick[pair];
  __shared__ double red[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const cd qi = q[pair * n + i];
    const cd dx = cscale(cmul(qi, x_of(qi, k0 * d)), -k0);
    acc += cmulc(dx, Xb[p * (long long)n * n + (long long)i * n + i]).x;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) d_thick[pair] = red[0];
}

// O2b = R T1 with R = [[-Fs, Fc], [Fc, Fs]] (rows; R is symmetric and orthogonal).
__global__ void k_rot_rows_R(int N, const cd* T1, const cd* ov, cd* O2b) {
  const long long p = blockIdx.y;
  const int n = 2 * N;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)N * n) return;
  const int r = (int)(e / n), c = (int)(e % n);
  const double fc = ov[p * OV_COUNT * N + OV_FC * N + r].x;
  const double fs = ov[p * OV_COUNT * N + OV_FS * N + r].x;
  const long long base = p * (long long)n * n;
  const cd top = T1[base + (long long)r * n + c], bot = T1[base + (long long)(N + r) * n + c];
  O2b[base + (long long)r * n + c] = cadd(cscale(top, -fs), cscale(bot, fc));
  O2b[base + (long long)(N + r) * n + c] = cadd(cscale(top, fc), cscale(bot, fs));
}

// BE = [[Kx Z Kx - I], [Ky Z Kx]], CD = [[Kx Z Ky], [Ky Z Ky - I]] (n x N each, BE then CD).
__global__ void k_be_cd(int N, BatchPtr epz, const cd* ov, cd* becd) {
  const long long p = blockIdx.y;
  const int n = 2 * N;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)N * N) return;
  const int i = (int)(e / N), j = (int)(e % N);
  const cd* o = ov + p * OV_COUNT * N;
  const cd kxi = o[OV_KX * N + i], kyi = o[OV_KY * N + i];
  const cd kxj = o[OV_KX * N + j], kyj = o[OV_KY * N + j];
  const cd z = epz.ptr(p)[e];
  cd* be = becd + p * (long long)n * n;
  cd* cdm = be + (long long)n * N;
  cd Bm = cmul(cmul(kxi, z), kxj), Dm = cmul(cmul(kyi, z), kyj);
  if (i == j) { Bm = csub(Bm, mkc(1.0, 0.0)); Dm = csub(Dm, mkc(1.0, 0.0)); }
  be[(long long)i * N + j] = Bm;
  be[(long long)(N + i) * N + j] = cmul(cmul(kyi, z), kxj);
  cdm[(long long)i * N + j] = cmul(cmul(kxi, z), kyj);
  cdm[(long long)(N + i) * N + j] = Dm;
}

// Per pair: epz-bar from BEb / CDb (n x N each, in T2), and the Omega_R blocks of ORt (= Ub W^H
// before the R'^T row rotation) added into epx-bar / epy-bar (and, with the off-diagonal slots,
// Omega_R[:N, :N] = -KxKy + epyx into epyx-bar, Omega_R[N:, N:] = KyKx - epxy into epxy-bar).
// E3 holds [epx | epy | epz | epxy | epyx] bars.
__global__ void k_adj_conv(int N, const cd* BECDb, const cd* ORt, const cd* ov, int offd, cd* E3) {
  const long long p = blockIdx.y;
  const int n = 2 * N;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long NN = (long long)N * N;
  if (e >= NN) return;
  const int i = (int)(e / N), j = (int)(e % N);
  const cd* o = ov + p * OV_COUNT * N;
  const cd kxi = cconj(o[OV_KX * N + i]), kyi = cconj(o[OV_KY * N + i]);
  const cd kxj = cconj(o[OV_KX * N + j]), kyj = cconj(o[OV_KY * N + j]);
  const cd* be = BECDb + p * (long long)n * n;
  const cd* cdm = be + (long long)n * N;
  cd z = cmul(cmul(kxi, be[(long long)i * N + j]), kxj);
  z = cadd(z, cmul(cmul(kyi, be[(long long)(N + i) * N + j]), kxj));
  z = cadd(z, cmul(cmul(kxi, cdm[(long long)i * N + j]), kyj));
  z = cadd(z, cmul(cmul(kyi, cdm[(long long)(N + i) * N + j]), kyj));
  cd* E = E3 + p * 5 * NN;
  E[2 * NN + e] = z;
  // R'^T U: top = c U_top - s U_bot, bottom = s U_top + c U_bot (c, s of the row's order)
  const double ci = o[OV_FC * N + i].x, si = o[OV_FS * N + i].x;
  const cd* U = ORt + p * (long long)n * n;
  const cd top_r = U[(long long)i * n + N + j], bot_r = U[(long long)(N + i) * n + N + j];
  const cd top_l = U[(long long)i * n + j], bot_l = U[(long long)(N + i) * n + j];
  const cd or_tr = csub(cscale(top_r, ci), cscale(bot_r, si));   // ORb[:N, N:]
  const cd or_bl = cadd(cscale(top_l, si), cscale(bot_l, ci));   // ORb[N:, :N]
  E[e] = cadd(E[e], or_bl);                 // Omega_R[N:, :N] = epx - Ky^2
  E[NN + e] = csub(E[NN + e], or_tr);       // Omega_R[:N, N:] = Kx^2 - epy
  if (offd) {
    const cd or_tl = csub(cscale(top_l, ci), cscale(bot_l, si));   // ORb[:N, :N]
    const cd or_br = cadd(cscale(top_r, si), cscale(bot_r, ci));   // ORb[N:, N:]
    E[4 * NN + e] = cadd(E[4 * NN + e], or_tl);
    E[3 * NN + e] = csub(E[3 * NN + e], or_br);
  }
}

// out[c] += sum over points p with conv_idx[p * L + l] == c of E3[p][slot] (slot grid z).
__global__ void k_reduce_conv(int P, int L, int l, long long NN, const int* conv_idx, const cd* E3,
                              cd* d_epx, cd* d_epy, cd* d_epz, cd* d_epxy, cd* d_epyx) {
  const int c = blockIdx.y, slot = blockIdx.z;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= NN) return;
  cd acc = mkc(0.0, 0.0);
  bool any = false;
  for (int p = 0; p < P; ++p) {
    if (conv_idx[(long long)p * L + l] != c) continue;
    acc = cadd(acc, E3[((long long)p * 5 + slot) * NN + e]);
    any = true;
  }
  if (!any) return;
  cd* const outs[5] = {d_epx, d_epy, d_epz, d_epxy, d_epyx};
  cd* out = outs[slot] + (long long)c * NN + e;
  *out = cadd(*out, acc);
}

__global__ void k_zero_c(long long count, cd* x) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < count) x[e] = mkc(0.0, 0.0);
}

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

struct AdjWs {
  cd *Gb, *Gn, *Hb, *Xsb, *Mxb, *G0f, *SB;
  cd *Ai, *Atmp, *Y, *Mlu, *F, *Mb, *Xb, *Fb, *Ab, *Bb, *Yb, *XYXb, *T1, *T2, *T3, *BECD;
  cd *E3;
  cd *crb, *ctb, *db, *gub, *ub, *tv, *Zs;
  int *pivA, *pivM, *infoA, *infoM;
  void* luws;
};

AdjWs adj_carve(void* ws, int P, int L, int n, int N, bool count_only, size_t* bytes) {
  char* p = (char*)ws;
  size_t used = 0;
  auto take = [&](size_t b) { char* r = p; p += align_up(b); used += align_up(b); return r; };
  const size_t nn = (size_t)n * n * sizeof(cd);
  AdjWs a;
  cd** mats[] = {&a.Gb, &a.Gn, &a.Hb, &a.Xsb, &a.Mxb, &a.G0f, &a.Ai, &a.Atmp, &a.Y, &a.Mlu,
                 &a.F, &a.Mb, &a.Xb, &a.Fb, &a.Ab, &a.Bb, &a.Yb, &a.XYXb, &a.T1, &a.T2, &a.T3,
                 &a.BECD};
  for (cd** m : mats) *m = (cd*)take((size_t)P * nn);
  a.SB = (cd*)take((size_t)P * nn * 2);
  a.E3 = (cd*)take((size_t)P * 5 * N * N * sizeof(cd));
  cd** vecs[] = {&a.crb, &a.ctb, &a.db, &a.gub, &a.ub, &a.tv};
  for (cd** v : vecs) *v = (cd*)take((size_t)P * n * 2 * sizeof(cd));
  a.Zs = (cd*)take((size_t)P * L * n * 2 * sizeof(cd));
  a.pivA = (int*)take((size_t)P * n * sizeof(int));
  a.pivM = (int*)take((size_t)P * n * sizeof(int));
  a.infoA = (int*)take((size_t)P * sizeof(int));
  a.infoM = (int*)take((size_t)P * sizeof(int));
  a.luws = take(wb_getrs_ws_bytes(n, 2 * n, P));
  (void)count_only;
  if (bytes) *bytes = used + 256;
  return a;
}

}  // namespace

size_t wb_adjoint_workspace_bytes(int P, int L, int hu, int hv, int mode) {
  const int N = (2 * hu + 1) * (2 * hv + 1);
  const int n = (mode == 0) ? 2 * N : N;
  size_t b = 0;
  adj_carve(nullptr, P, L, n, N, true, &b);
  return b;
}

hipError_t wb_adjoint_launch(const wb_adjoint_args& a, void* ws, hipStream_t st) {
  const wb_solve_args& f = *a.fwd;
  const int P = f.P, L = f.L, mode = f.mode;
  if (mode != 0) return hipErrorNotSupported;
  const int N = (2 * f.hu + 1) * (2 * f.hv + 1);
  const int n = 2 * N;
  const long long nn = (long long)n * n, NN = (long long)N * N, n2 = 2 * nn;
  const long long vs = (long long)n * 2;
  const wb_solve_ws W = wb_solve_carve(const_cast<void*>(a.solve_ws), P, L, f.hu, f.hv, mode);
  const wb_ck CK = wb_ck_carve(const_cast<void*>(a.ck), P, L, f.hu, f.hv, mode);
  AdjWs T = adj_carve(ws, P, L, n, N, false, nullptr);
  const cd one = mkc(1.0, 0.0), zero = mkc(0.0, 0.0), mone = mkc(-1.0, 0.0);
  hipError_t e;
  auto S = [&](const cd* base, long long stride) { return bp_strided(base, stride); };
  // one layer's column of a per-pair buffer
  auto LY = [&](const cd* base, long long per, int l) { return bp_strided(base + (long long)l * per, (long long)L * per); };
  const long long dstride = (long long)(L + 1) * vs;

  // the NVF off-diagonal slots (epxy in Omega2's second column block and Omega_R's bottom-right
  // block, epyx in the first and the top-left), both or neither
  const int offd = (f.epxy != nullptr && f.epyx != nullptr && a.d_epxy != nullptr && a.d_epyx != nullptr) ? 1 : 0;
  if ((f.epxy != nullptr) != (a.d_epxy != nullptr)) return hipErrorInvalidValue;
  // outputs start at zero (fill kernels: no memset nodes in a capturable sequence)
  const long long conv_elems = (long long)a.n_conv * NN;
  for (cd* out : {a.d_epx, a.d_epy, a.d_epz, a.d_epxy, a.d_epyx}) {
    if (!out) continue;
    hipLaunchKernelGGL(k_zero_c, dim3((unsigned)((conv_elems + 255) / 256)), dim3(256), 0, st,
                       conv_elems, out);
    WB_CHECK_LAUNCH();
  }

  // ------------------------------------------------ efficiencies -> basis columns
  hipLaunchKernelGGL(k_eff_bar, dim3(P), dim3(256), 0, st, f.hu, f.hv, f.pts, W.ov, W.vec[2],
                     W.vec[3], a.eff_bar, T.crb, T.ctb);
  WB_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_vec_diag, grid2(vs, P), dim3(256), 0, st, n, W.dg, (int)DG_T21, T.crb, T.gub, 0);
  WB_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_vec_diag, grid2(vs, P), dim3(256), 0, st, n, W.dg, (int)DG_R12, T.ctb, T.db, 0);
  WB_CHECK_LAUNCH();

  // ------------------------------------------------ back-sweep, bottom layer first
  for (int l = 0; l < L; ++l) {
    BatchPtr Zl = bp_strided(T.Zs + (long long)l * vs, (long long)L * vs);
    BatchPtr LUMx = LY(W.M[0], nn, l);
    const int* piv = W.ipiv[2] + (long long)l * P * n;
    BatchPtr S12 = bp_offset(LY(W.M6, n2, l), n);
    hipLaunchKernelGGL(k_copy2, grid2(vs, P), dim3(256), 0, st, n, 2, S(T.db, vs), 2, Zl, 2);
    WB_CHECK_LAUNCH();
    if ((e = wb_getrs_conj_launch(n, 2, LUMx, n, piv, Zl, 2, P, st))) return e;
    // d-bar_{l+1} = s12^H z_l
    if ((e = wb_zgemm_ops_launch(1, 0, n, 2, n, one, S12, 2 * n, Zl, 2, zero, S(T.db, vs), 2, P, st)))
      return e;
  }
  // top interface: u-bar = G_L^H gu-bar + d-bar_L ; z = Tf^-H u-bar
  BatchPtr GamL = LY(W.M6, n2, L - 1);
  hipLaunchKernelGGL(k_copy2, grid2(vs, P), dim3(256), 0, st, n, 2, S(T.db, vs), 2, S(T.ub, vs), 2);
  WB_CHECK_LAUNCH();
  if ((e = wb_zgemm_ops_launch(1, 0, n, 2, n, one, GamL, 2 * n, S(T.gub, vs), 2, one, S(T.ub, vs), 2,
                               P, st)))
    return e;
  if ((e = wb_getrs_conj_launch(n, 2, S(W.G0, nn), n, W.ipiv[3], S(T.ub, vs), 2, P, st))) return e;
  // Gamma_L-bar = (gu-bar + conj(t11) z) u^H
  hipLaunchKernelGGL(k_copy2, grid2(vs, P), dim3(256), 0, st, n, 2, S(T.gub, vs), 2, S(T.tv, vs), 2);
  WB_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_vec_diag, grid2(vs, P), dim3(256), 0, st, n, W.dg, (int)DG_T11, T.ub, T.tv, 1);
  WB_CHECK_LAUNCH();
  if ((e = wb_zgemm_ops_launch(0, 1, n, n, 2, one, S(T.tv, vs), 2, S(CK.d + (long long)L * vs, dstride),
                               2, zero, S(T.Gb, nn), n, P, st)))
    return e;
  hipLaunchKernelGGL(k_gamma0_full, grid2(nn, P), dim3(256), 0, st, n, W.dg, T.G0f);
  WB_CHECK_LAUNCH();

  // ------------------------------------------------ layers, top first
  for (int l = L - 1; l >= 0; --l) {
    BatchPtr LUMx = LY(W.M[0], nn, l);
    const int* piv = W.ipiv[2] + (long long)l * P * n;
    BatchPtr Xs = LY(W.M[1], nn, l);
    BatchPtr Hl = LY(W.M[2], nn, l);
    BatchPtr XYX = LY(W.M[3], nn, l);
    BatchPtr Am = LY(W.M[4], nn, l);
    BatchPtr Bm = LY(W.M[5], nn, l);
    BatchPtr S12 = bp_offset(LY(W.M6, n2, l), n);
    BatchPtr S11 = LY(CK.S11, nn, l);
    BatchPtr Wck = LY(CK.W, nn, l);
    BatchPtr Wick = LY(CK.Wi, nn, l);
    BatchPtr Gam = (l == 0) ? S(T.G0f, nn) : LY(W.M6, n2, l - 1);
    const int ldgam = (l == 0) ? n : 2 * n;
    BatchPtr Zl = bp_strided(T.Zs + (long long)l * vs, (long long)L * vs);
    BatchPtr dl = S(CK.d + (long long)l * vs, dstride);
    BatchPtr dl1 = S(CK.d + (long long)(l + 1) * vs, dstride);
    BatchPtr SB = S(T.SB, n2), SBr = bp_offset(S(T.SB, n2), n);
    auto TT = [&](cd* x) { return S(x, nn); };

    // ---- cascade: Gamma_{l+1}-bar (in Gb) -> S11-bar, S12-bar, Gamma_l-bar
    hipLaunchKernelGGL(k_copy2, grid2(nn, P), dim3(256), 0, st, n, n, TT(T.Gb), n, SB, 2 * n);
    WB_CHECK_LAUNCH();
    if ((e = wb_zgemm_ops_launch(0, 1, n, n, n, one, TT(T.Gb), n, Hl, n, zero, SBr, 2 * n, P, st))) return e;
    if ((e = wb_zgemm_ops_launch(0, 1, n, n, 2, one, Zl, 2, dl1, 2, one, SBr, 2 * n, P, st))) return e;
    if ((e = wb_zgemm_ops_launch(1, 0, n, n, n, one, S12, 2 * n, TT(T.Gb), n, zero, TT(T.Hb), n, P, st))) return e;
    if ((e = wb_zgemm_ops_launch(0, 1, n, n, n, one, TT(T.Hb), n, Xs, n, zero, TT(T.Gn), n, P, st))) return e;
    if ((e = wb_zgemm_ops_launch(1, 0, n, n, n, one, Gam, ldgam, TT(T.Hb), n, zero, TT(T.Xsb), n, P, st))) return e;
    if ((e = wb_getrs_blocked(n, n, LUMx, n, piv, TT(T.Xsb), n, P, 1, T.luws, st))) return e;
    hipLaunchKernelGGL(k_add, grid2(nn, P), dim3(256), 0, st, n, TT(T.Xsb), n, SBr, 2 * n);
    WB_CHECK_LAUNCH();
    if ((e = wb_zgemm_ops_launch(0, 1, n, n, n, mone, TT(T.Xsb), n, Xs, n, zero, TT(T.Mxb), n, P, st))) return e;
    if ((e = wb_zgemm_ops_launch(0, 1, n, n, 2, mone, Zl, 2, dl, 2, one, TT(T.Mxb), n, P, st))) return e;
    if ((e = wb_zgemm_ops_launch(0, 1, n, n, n, mone, TT(T.Mxb), n, Gam, ldgam, one, SB, 2 * n, P, st))) return e;
    if ((e = wb_zgemm_ops_launch(1, 0, n, n, n, mone, S11, n, TT(T.Mxb), n, one, TT(T.Gn), n, P, st))) return e;
    { cd* t = T.Gb; T.Gb = T.Gn; T.Gn = t; }

    // ---- layer S-matrix: [S11-bar | S12-bar] -> A-bar, B-bar, X-bar
    hipLaunchKernelGGL(k_copy2, grid2(nn, P), dim3(256), 0, st, n, n, Am, n, TT(T.Atmp), n);
    WB_CHECK_LAUNCH();
    if ((e = wb_inverse_blocked(n, TT(T.Atmp), n, TT(T.Ai), n, T.pivA, T.infoA, P, T.luws, st))) return e;
    if ((e = wb_zgemm_ops_launch(0, 0, n, n, n, one, Bm, n, TT(T.Ai), n, zero, TT(T.Y), n, P, st))) return e;
    hipLaunchKernelGGL(k_copy2, grid2(nn, P), dim3(256), 0, st, n, n, Am, n, TT(T.Mlu), n);
    WB_CHECK_LAUNCH();
    if ((e = wb_zgemm_ops_launch(0, 0, n, n, n, mone, XYX, n, Bm, n, one, TT(T.Mlu), n, P, st))) return e;
    if ((e = wb_getrf_launch(n, TT(T.Mlu), n, T.pivM, T.infoM, P, st))) return e;
    hipLaunchKernelGGL(k_copy2, grid2(nn, P), dim3(256), 0, st, n, n, Am, n, TT(T.F), n);
    WB_CHECK_LAUNCH();
    if ((e = wb_zgemm_ops_launch(0, 0, n, n, n, mone, TT(T.Y), n, Bm, n, one, TT(T.F), n, P, st))) return e;
    // [E1-bar | E2-bar] = M^-H [S11-bar | S12-bar]
    if ((e = wb_getrs_blocked(n, 2 * n, TT(T.Mlu), n, T.pivM, SB, 2 * n, P, 1, T.luws, st))) return e;
    // M-bar = -(E1b S11^H + E2b S12^H)
    if ((e = wb_zgemm_ops_launch(0, 1, n, n, n, mone, SB, 2 * n, S11, n, zero, TT(T.Mb), n, P, st))) return e;
    if ((e = wb_zgemm_ops_launch(0, 1, n, n, n, mone, SBr, 2 * n, S12, 2 * n, one, TT(T.Mb), n, P, st))) return e;
    // X-bar = E2b F^H
    if ((e = wb_zgemm_ops_launch(0, 1, n, n, n, one, SBr, 2 * n, TT(T.F), n, zero, TT(T.Xb), n, P, st))) return e;
    hipLaunchKernelGGL(k_adj_a, grid2(nn, P), dim3(256), 0, st, n, L, l, T.SB, T.Mb, W.q, f.pts,
                       f.thick, T.Fb, T.Ab, T.Bb);
    WB_CHECK_LAUNCH();
    // Y-bar = -Fb B^H ; B-bar -= Y^H Fb
    if ((e = wb_zgemm_ops_launch(0, 1, n, n, n, mone, TT(T.Fb), n, Bm, n, zero, TT(T.Yb), n, P, st))) return e;
    if ((e = wb_zgemm_ops_launch(1, 0, n, n, n, mone, TT(T.Y), n, TT(T.Fb), n, one, TT(T.Bb), n, P, st))) return e;
    // XYX-bar = E1b A^H - Mb B^H
    if ((e = wb_zgemm_ops_launch(0, 1, n, n, n, one, SB, 2 * n, Am, n, zero, TT(T.XYXb), n, P, st))) return e;
    if ((e = wb_zgemm_ops_launch(0, 1, n, n, n, mone, TT(T.Mb), n, Bm, n, one, TT(T.XYXb), n, P, st))) return e;
    // A-bar += XYX^H E1b ; B-bar -= XYX^H Mb
    if ((e = wb_zgemm_ops_launch(1, 0, n, n, n, one, XYX, n, SB, 2 * n, one, TT(T.Ab), n, P, st))) return e;
    if ((e = wb_zgemm_ops_launch(1, 0, n, n, n, mone, XYX, n, TT(T.Mb), n, one, TT(T.Bb), n, P, st))) return e;
    // XYX = X Y X
    hipLaunchKernelGGL(k_adj_xyx, grid2(nn, P), dim3(256), 0, st, n, L, l, T.XYXb, W.q, f.pts,
                       f.thick, T.Yb, T.T1, T.T2);
    WB_CHECK_LAUNCH();
    if ((e = wb_zgemm_ops_launch(0, 1, n, n, n, one, TT(T.T1), n, TT(T.Y), n, one, TT(T.Xb), n, P, st))) return e;
    if ((e = wb_zgemm_ops_launch(1, 0, n, n, n, one, TT(T.Y), n, TT(T.T2), n, one, TT(T.Xb), n, P, st))) return e;
    // Y = B A^-1: T3 = Yb A^-H ; B-bar += T3 (folded below) ; A-bar -= Y^H T3
    if ((e = wb_zgemm_ops_launch(0, 1, n, n, n, one, TT(T.Yb), n, TT(T.Ai), n, zero, TT(T.T3), n, P, st))) return e;
    if ((e = wb_zgemm_ops_launch(1, 0, n, n, n, mone, TT(T.Y), n, TT(T.T3), n, one, TT(T.Ab), n, P, st))) return e;
    // A, B -> Vi (the eigenbasis side Wi is held fixed); Vi = Q G
    hipLaunchKernelGGL(k_adj_vi, grid2(nn, P), dim3(256), 0, st, N, L, l, T.Ab, T.Bb, T.T3,
                       CK.Vi, W.q, W.ov, T.Fb, T.T1, T.T2);
    WB_CHECK_LAUNCH();
    // Q-bar = Vib G^H (into Mb)
    if ((e = wb_zgemm_ops_launch(0, 1, n, n, n, one, TT(T.Fb), n, TT(T.T1), n, zero, TT(T.Mb), n, P, st))) return e;
    // U-bar = -G^H (Gb G^H) (into Yb), G = U^-1
    if ((e = wb_zgemm_ops_launch(0, 1, n, n, n, one, TT(T.T2), n, TT(T.T1), n, zero, TT(T.T3), n, P, st))) return e;
    if ((e = wb_zgemm_ops_launch(1, 0, n, n, n, mone, TT(T.T1), n, TT(T.T3), n, zero, TT(T.Yb), n, P, st))) return e;
    // Omega_R-bar before the R'^T rotation: U-bar W^H (into Ab)
    if ((e = wb_zgemm_ops_launch(0, 1, n, n, n, one, TT(T.Yb), n, Wck, n, zero, TT(T.Ab), n, P, st))) return e;
    // X = exp(-k0 d Q): divided differences, then the generator divide; thickness
    hipLaunchKernelGGL(k_thick_bar, dim3(P), dim3(256), 0, st, n, L, l, T.Xb, W.q, f.pts, f.thick,
                       a.d_thick);
    WB_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_adj_expm, grid2(nn, P), dim3(256), 0, st, n, L, l, T.Xb, W.q, f.pts, f.thick,
                       T.Mb);
    WB_CHECK_LAUNCH();
    // Omega2-bar = W^-H Qt W^H = R Wi^H Qt W^H (into Bb)
    if ((e = wb_zgemm_ops_launch(0, 1, n, n, n, one, TT(T.Mb), n, Wck, n, zero, TT(T.T3), n, P, st))) return e;
    if ((e = wb_zgemm_ops_launch(1, 0, n, n, n, one, Wick, n, TT(T.T3), n, zero, TT(T.T1), n, P, st))) return e;
    hipLaunchKernelGGL(k_rot_rows_R, grid2((long long)N * n, P), dim3(256), 0, st, N, T.T1, W.ov, T.Bb);
    WB_CHECK_LAUNCH();
    // Omega2 = diag + [BE epx | CD epy], BE / CD from epz_i
    BatchPtr EPX = bp_gather(f.epx, NN, f.conv_idx + l, L);
    BatchPtr EPY = bp_gather(f.epy, NN, f.conv_idx + l, L);
    BatchPtr EPZ = bp_gather(f.epz, NN, f.conv_idx + l, L);
    hipLaunchKernelGGL(k_be_cd, grid2(NN, P), dim3(256), 0, st, N, EPZ, W.ov, T.BECD);
    WB_CHECK_LAUNCH();
    BatchPtr E3x = S(T.E3, 5 * NN), E3y = S(T.E3 + NN, 5 * NN);
    if ((e = wb_zgemm_ops_launch(1, 0, N, N, n, one, TT(T.BECD), N, TT(T.Bb), n, zero, E3x, N, P, st))) return e;
    if ((e = wb_zgemm_ops_launch(1, 0, N, N, n, one, S(T.BECD + (long long)n * N, nn), N,
                                 S(T.Bb + N, nn), n, zero, E3y, N, P, st))) return e;
    if ((e = wb_zgemm_ops_launch(0, 1, n, N, N, one, TT(T.Bb), n, EPX, N, zero, TT(T.T2), N, P, st))) return e;
    if ((e = wb_zgemm_ops_launch(0, 1, n, N, N, one, S(T.Bb + N, nn), n, EPY, N, zero,
                                 S(T.T2 + (long long)n * N, nn), N, P, st))) return e;
    if (offd) {
      // Omega2[:, :N] -= CD epyx, Omega2[:, N:] -= BE epxy: the slots' bars, and BE / CD's
      BatchPtr EPXY = bp_gather(f.epxy, NN, f.conv_idx + l, L);
      BatchPtr EPYX = bp_gather(f.epyx, NN, f.conv_idx + l, L);
      BatchPtr E3xy = S(T.E3 + 3 * NN, 5 * NN), E3yx = S(T.E3 + 4 * NN, 5 * NN);
      if ((e = wb_zgemm_ops_launch(1, 0, N, N, n, mone, TT(T.BECD), N, S(T.Bb + N, nn), n, zero, E3xy, N,
                                   P, st))) return e;
      if ((e = wb_zgemm_ops_launch(1, 0, N, N, n, mone, S(T.BECD + (long long)n * N, nn), N, TT(T.Bb), n,
                                   zero, E3yx, N, P, st))) return e;
      if ((e = wb_zgemm_ops_launch(0, 1, n, N, N, mone, S(T.Bb + N, nn), n, EPXY, N, one, TT(T.T2), N, P,
                                   st))) return e;
      if ((e = wb_zgemm_ops_launch(0, 1, n, N, N, mone, TT(T.Bb), n, EPYX, N, one,
                                   S(T.T2 + (long long)n * N, nn), N, P, st))) return e;
    }
    hipLaunchKernelGGL(k_adj_conv, grid2(NN, P), dim3(256), 0, st, N, T.T2, T.Ab, W.ov, offd, T.E3);
    WB_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_reduce_conv, dim3((unsigned)((NN + 255) / 256), a.n_conv, 3 + 2 * offd), dim3(256),
                       0, st, P, L, l, NN, f.conv_idx, T.E3, a.d_epx, a.d_epy, a.d_epz, a.d_epxy, a.d_epyx);
    WB_CHECK_LAUNCH();
  }
  return hipSuccess;
}
