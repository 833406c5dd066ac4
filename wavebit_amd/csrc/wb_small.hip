// Fused plain solve of 1D (TE / TM) points with at most 32 harmonics: one 256-thread workgroup per
// point carries the whole chain in LDS. The batched path runs a 21-harmonic grating as some
// seventy dependent launches, each a few microseconds of launch and drain around a few
// microseconds of work; here the same arithmetic runs with no launch boundary:
//
//   per order     point_order_setup (wb_rcwa.h): wavevectors, half-space kz, gap admittance and
//                 the diagonal half-space -> gap interface blocks;
//   per layer     the operator (TE: Kx^2 - epy; TM: epx (Kx epz_i Kx - I), solve_1d.py:203-253),
//                 Householder reduction to Hessenberg form with the unitary accumulated, the
//                 single-shift Schur iteration of wb_waveqr.h on one wave, the triangular
//                 eigenvectors (ztrevc's back substitution) and W = Z X with unit columns (the
//                 batched eigensolver's steps, wb_eig.hip), q = mode_q(w), V = W q (TE) or
//                 epx^-1 W q (TM);
//                 the gap-embedded interface matrices A, B = W^-1 Wg +- V^-1 Vg and
//                 A', B' = Wg^-1 W +- Vg^-1 V, stored transposed, and one step of the transposed
//                 reflection-matrix recursion of wb_rcwa.hip: P^T = A^T + G^T B^T,
//                 R^T = X P^-T (B^T + G^T A^T) X, G'^T = P'^-T (B'^T + R^T A'^T) with
//                 P'^T = A'^T + R^T B'^T (the factors of P^T and P'^T kept for the back-sweep);
//   final         M^T = I - G^T diag(t11), u = M^-1 t12 e, R = t22 e + t21 G u, the transmitted
//                 wave d_l = 4 P_l^-1 X_l P'_l^-1 d_{l+1} down to d_0 and T = r12 d_0, then the
//                 efficiencies (point_efficiency, wb_rcwa.h).
//
// Matrices sit in LDS with an odd row stride (33), eight of them; the factors of every layer's
// P^T and P'^T, their pivots, the layer's propagation diagonal and G^T between layers go to the
// solve workspace (global memory, read back once in the back-sweep). Dense steps use all 256
// threads: products by entry, LU with partial pivoting one column per step (every wave finds
// the pivot, each thread reads its entries' new values, then all write), many-right-hand-side
// solves by eight-thread groups per column; the vector solves of the final interface and the
// back-sweep run on one wave. Flags follow the batched path: NOCONVERGE and SINGULAR (W, V) per
// pair, SINGULAR of the recursion's and the final interface's factors on every pair of the point.
#include "wb_rcwa.h"
#include "wb_waveqr.h"

namespace {

constexpr int SN = 32;              // largest harmonic count
constexpr int SLD = 33;             // LDS row stride
constexpr int SMAT = SN * SLD;      // complex entries per LDS matrix
constexpr int ST = 64;              // one wave per point

#define AT(M, r, c) (M)[(r) * SLD + (c)]

struct SmallLds {
  cd* M[8];
  cd *ov, *dg, *w, *q, *xv, *va, *vb, *hv, *rd;
  double* red;
  int *piv, *piv2;
};

constexpr size_t small_lds_bytes() {
  return sizeof(cd) * ((size_t)8 * SMAT + (size_t)(OV_COUNT + DG_COUNT + 7) * SN) +
         sizeof(double) * 3 * ST + sizeof(int) * 2 * SN;
}

__device__ inline SmallLds small_carve(unsigned char* base) {
  SmallLds s;
  cd* c = (cd*)base;
  for (int i = 0; i < 8; ++i) { s.M[i] = c; c += SMAT; }
  s.ov = c; c += OV_COUNT * SN;
  s.dg = c; c += DG_COUNT * SN;
  s.w = c; c += SN;
  s.q = c; c += SN;
  s.xv = c; c += SN;
  s.va = c; c += SN;
  s.vb = c; c += SN;
  s.hv = c; c += SN;
  s.rd = c; c += SN;
  s.red = (double*)c;
  int* ip = (int*)(s.red + 3 * ST);
  s.piv = ip; ip += SN;
  s.piv2 = ip;
  return s;
}

// Everything below runs on the point's one wave: the LDS accesses of a wave complete in program
// order, so a step reads what the previous steps wrote with no barrier.

// C = D + A B (D may be null or C itself; C must not be A or B).
__device__ void w_mm(int n, const cd* A, const cd* B, const cd* D, cd* C) {
  for (int e = threadIdx.x; e < n * n; e += ST) {
    const int i = e / n, j = e % n;
    cd acc = D ? AT(D, i, j) : mkc(0.0, 0.0);
    cd a1 = mkc(0.0, 0.0), a2 = mkc(0.0, 0.0), a3 = mkc(0.0, 0.0);
    int k = 0;
    for (; k + 4 <= n; k += 4) {
      acc = cadd(acc, cmul(AT(A, i, k), AT(B, k, j)));
      a1 = cadd(a1, cmul(AT(A, i, k + 1), AT(B, k + 1, j)));
      a2 = cadd(a2, cmul(AT(A, i, k + 2), AT(B, k + 2, j)));
      a3 = cadd(a3, cmul(AT(A, i, k + 3), AT(B, k + 3, j)));
    }
    for (; k < n; ++k) acc = cadd(acc, cmul(AT(A, i, k), AT(B, k, j)));
    AT(C, i, j) = cadd(cadd(acc, a1), cadd(a2, a3));
  }
}

__device__ void w_identity(int n, cd* A) {
  for (int e = threadIdx.x; e < n * n; e += ST) {
    const int i = e / n, j = e % n;
    AT(A, i, j) = mkc(i == j ? 1.0 : 0.0, 0.0);
  }
}

// C = D + op(A) op(B), op the conjugate transpose where ha / hb is set (D may be null or C).
__device__ void w_mm_op(int n, int ha, const cd* A, int hb, const cd* B, const cd* D, cd* C) {
  for (int e = threadIdx.x; e < n * n; e += ST) {
    const int i = e / n, j = e % n;
    cd acc = D ? AT(D, i, j) : mkc(0.0, 0.0);
    for (int k = 0; k < n; ++k) {
      const cd a = ha ? cconj(AT(A, k, i)) : AT(A, i, k);
      const cd b = hb ? cconj(AT(B, j, k)) : AT(B, k, j);
      acc = cadd(acc, cmul(a, b));
    }
    AT(C, i, j) = acc;
  }
}

// Whether A equals A^H to 64 ulp of its largest entry.
__device__ bool w_is_herm(int n, const cd* A) {
  double dmax = 0.0, amax = 0.0;
  for (int e = threadIdx.x; e < n * n; e += ST) {
    const int i = e / n, j = e % n;
    dmax = fmax(dmax, cabs1(csub(AT(A, i, j), cconj(AT(A, j, i)))));
    amax = fmax(amax, cabs1(AT(A, i, j)));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    dmax = fmax(dmax, __shfl_xor(dmax, off));
    amax = fmax(amax, __shfl_xor(amax, off));
  }
  return dmax <= 64.0 * ULP * amax;
}

// A := (A + A^H) / 2.
__device__ void w_symmetrize(int n, cd* A) {
  for (int e = threadIdx.x; e < n * n; e += ST) {
    const int i = e / n, j = e % n;
    if (i <= j) {
      const cd h = cscale(cadd(AT(A, i, j), cconj(AT(A, j, i))), 0.5);
      AT(A, i, j) = h;
      AT(A, j, i) = cconj(h);
    }
  }
}

// Cholesky A = C C^H in place (C lower triangular with a positive real diagonal, the upper part
// cleared): lane i updates row i. Returns 1 when a pivot is not positive and finite (A is not
// positive definite), else 0.
__device__ int w_chol(int n, cd* A) {
  const int i = threadIdx.x & 63;
  for (int k = 0; k < n; ++k) {
    const double akk = AT(A, k, k).x;
    if (!(akk > 0.0) || !isfinite(akk)) return 1;
    const double c = sqrt(akk), ic = 1.0 / c;
    if (i == k) AT(A, k, k) = mkc(c, 0.0);
    if (i > k && i < n) AT(A, i, k) = cscale(AT(A, i, k), ic);
    if (i > k && i < n) {
      const cd lik = AT(A, i, k);
      for (int j = k + 1; j <= i; ++j) AT(A, i, j) = csub(AT(A, i, j), cmul(lik, cconj(AT(A, j, k))));
    }
  }
  for (int e = threadIdx.x; e < n * n; e += ST) {
    const int r = e / n, c = e % n;
    if (c > r) AT(A, r, c) = mkc(0.0, 0.0);
  }
  return 0;
}

// X = C^-1 for lower triangular C: lane j solves C x = e_j by forward substitution.
__device__ void w_trinv_lower(int n, const cd* C, cd* X) {
  const int j = threadIdx.x & 63;
  if (j < n) {
    for (int r = 0; r < j; ++r) AT(X, r, j) = mkc(0.0, 0.0);
    for (int r = j; r < n; ++r) {
      cd acc = mkc(r == j ? 1.0 : 0.0, 0.0);
      for (int k = j; k < r; ++k) acc = csub(acc, cmul(AT(C, r, k), AT(X, k, j)));
      AT(X, r, j) = cdiv(acc, AT(C, r, r));
    }
  }
}

// A = P L U in place (LAPACK getrf: piv[k] is the row swapped with row k, whole rows swapped;
// the pivot is the largest |re| + |im|, the first on ties). Lane (j, h) = (lane & 31, lane >> 5)
// updates column j's rows of parity h. Returns the first zero pivot's column + 1, or 0.
__device__ int w_getrf(int n, cd* A, int* piv) {
  const int lane = threadIdx.x & 63, j = lane & 31, h = lane >> 5;
  int info = 0;
  for (int k = 0; k < n; ++k) {
    const int r = k + lane;
    double v = (lane < 32 && r < n) ? cabs1(AT(A, r, k)) : -1.0;
    double m = v;
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    m = fmax(m, __shfl_xor(m, 32));
    const unsigned long long hit = __ballot(v == m && lane < 32 && r < n);
    const int p = k + (int)__ffsll((long long)hit) - 1;
    if (lane == 0) piv[k] = p;
    const cd pv = AT(A, p, k);
    const bool zero = (pv.x == 0.0 && pv.y == 0.0);
    if (zero && info == 0) info = k + 1;
    if (p != k && h == 0 && j < n) {
      const cd a = AT(A, k, j);
      AT(A, k, j) = AT(A, p, j);
      AT(A, p, j) = a;
    }
    const cd rp = zero ? mkc(0.0, 0.0) : crecip(pv);
    if (k + 1 + lane < n) AT(A, k + 1 + lane, k) = cmul(AT(A, k + 1 + lane, k), rp);
    if (j > k && j < n) {
      const cd ukj = AT(A, k, j);
      // four rows per pass, every load issued before the stores (the rows' column-k entries are
      // not written here)
      for (int i = k + 1 + h; i < n; i += 8) {
        cd a[4], l[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int r = (i + 2 * u < n) ? i + 2 * u : i;
          a[u] = AT(A, r, j);
          l[u] = AT(A, r, k);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (i + 2 * u < n) AT(A, i + 2 * u, j) = csub(a[u], cmul(l[u], ukj));
      }
    }
  }
  return info;
}

// B := A^-1 B (nrhs <= 32 columns) from getrf's factors, lane (c, h) on column c: the
// interchanges, then both triangular sweeps with the row updates split by parity.
__device__ void w_getrs(int n, int nrhs, const cd* LU, const int* piv, cd* B, SmallLds& S) {
  const int lane = threadIdx.x & 63, c = lane & 31, h = lane >> 5;
  const bool live = c < nrhs;
  if (lane < n) {
    const cd u = AT(LU, lane, lane);
    S.rd[lane] = (u.x == 0.0 && u.y == 0.0) ? mkc(0.0, 0.0) : crecip(u);
  }
  for (int k = 0; k < n; ++k) {
    const int pk = piv[k];
    if (pk != k && live && h == 0) {
      const cd a = AT(B, k, c);
      AT(B, k, c) = AT(B, pk, c);
      AT(B, pk, c) = a;
    }
  }
  // (four rows per pass, loads before stores: B's column c and LU do not overlap)
  for (int k = 0; k + 1 < n; ++k) {
    if (live) {
      const cd xk = AT(B, k, c);
      for (int i = k + 1 + h; i < n; i += 8) {
        cd bv[4], lv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int r = (i + 2 * u < n) ? i + 2 * u : i;
          bv[u] = AT(B, r, c);
          lv[u] = AT(LU, r, k);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (i + 2 * u < n) AT(B, i + 2 * u, c) = csub(bv[u], cmul(lv[u], xk));
      }
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    if (live) {
      const cd xk = cmul(AT(B, k, c), S.rd[k]);
      for (int i = h; i < k; i += 8) {
        cd bv[4], lv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int r = (i + 2 * u < k) ? i + 2 * u : i;
          bv[u] = AT(B, r, c);
          lv[u] = AT(LU, r, k);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (i + 2 * u < k) AT(B, i + 2 * u, c) = csub(bv[u], cmul(lv[u], xk));
      }
      if (h == 0) AT(B, k, c) = xk;
    }
  }
}

// x (lane i holds x_i) := M^-1 x for the factors LU of M^T (M^T = Pi L U, so
// M = U^T L^T Pi^T): forward with U^T, backward with L^T, then the interchanges in reverse.
__device__ cd w_getrs_t(int n, const cd* LU, const int* piv, cd x) {
  const int lane = threadIdx.x & 63;
  cd rd = mkc(0.0, 0.0);
  if (lane < n) {
    const cd u = AT(LU, lane, lane);
    rd = (u.x == 0.0 && u.y == 0.0) ? mkc(0.0, 0.0) : crecip(u);
  }
  for (int k = 0; k < n; ++k) {
    const cd yk = cmul(readlane_c(x, k), readlane_c(rd, k));
    if (lane == k) x = yk;
    if (lane > k && lane < n) x = csub(x, cmul(AT(LU, k, lane), yk));
  }
  for (int k = n - 1; k > 0; --k) {
    const cd zk = readlane_c(x, k);
    if (lane < k) x = csub(x, cmul(AT(LU, k, lane), zk));
  }
  for (int k = n - 1; k >= 0; --k) {
    const int pk = piv[k];
    if (pk != k) {
      const cd a = readlane_c(x, k), b = readlane_c(x, pk);
      if (lane == k) x = b;
      if (lane == pk) x = a;
    }
  }
  return x;
}

// Householder reduction A := Z^H A Z to upper Hessenberg form with Z (identity on entry)
// accumulating the reflectors H_k = I - tau v v^H from the right (wb_eig.hip's k_gehrd and
// k_unghr in one pass). The right application gives lane r row r of A (r < n) or row r - n of Z;
// the left application gives lane (c, h) column c's rows of parity h.
__device__ void w_gehrd(int n, cd* A, cd* Z, SmallLds& S) {
  const int lane = threadIdx.x & 63;
  cd* v = S.hv;
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;
    const cd xi = (lane < m) ? AT(A, k + 1 + lane, k) : mkc(0.0, 0.0);
    double ss = (lane >= 1 && lane < m) ? cnorm2(xi) : 0.0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off);
    const cd alpha = readlane_c(xi, 0);
    cd tk, scale;
    double beta;
    if (ss == 0.0 && alpha.y == 0.0) {
      tk = mkc(0.0, 0.0);
      beta = alpha.x;
      scale = mkc(0.0, 0.0);
    } else {
      beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + ss), alpha.x);
      tk = mkc((beta - alpha.x) / beta, -alpha.y / beta);
      scale = crecip(mkc(alpha.x - beta, alpha.y));
    }
    if (lane < m) v[lane] = (lane == 0) ? mkc(1.0, 0.0) : cmul(xi, scale);
    if (tk.x != 0.0 || tk.y != 0.0) {
      // right: row -= tau (row . v) v^H on the rows of A and of Z
      if (lane < 2 * n) {
        cd* R = (lane < n) ? &AT(A, lane, 0) : &AT(Z, lane - n, 0);
        cd z = mkc(0.0, 0.0);
        for (int jj = 0; jj < m; ++jj) z = cadd(z, cmul(R[k + 1 + jj], v[jj]));
        z = cmul(tk, z);
        for (int jj = 0; jj < m; jj += 4) {
          cd rv[4], vv[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int q = (jj + u < m) ? jj + u : jj;
            rv[u] = R[k + 1 + q];
            vv[u] = v[q];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (jj + u < m) R[k + 1 + jj + u] = csub(rv[u], cmul(z, cconj(vv[u])));
        }
      }
      // left: rows k+1.. of columns k+1.. -= conj(tau) v (v^H column)
      const int col = k + 1 + (lane & 31), hh = lane >> 5;
      cd sdot = mkc(0.0, 0.0);
      if (col < n)
        for (int i = hh; i < m; i += 2) sdot = cadd(sdot, cmulc(v[i], AT(A, k + 1 + i, col)));
      sdot = cadd(sdot, mkc(__shfl_xor(sdot.x, 32), __shfl_xor(sdot.y, 32)));
      const cd su = cmul(cconj(tk), sdot);
      if (col < n)
        for (int i = hh; i < m; i += 8) {
          cd av[4], vv[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int q = (i + 2 * u < m) ? i + 2 * u : i;
            av[u] = AT(A, k + 1 + q, col);
            vv[u] = v[q];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (i + 2 * u < m) AT(A, k + 1 + i + 2 * u, col) = csub(av[u], cmul(vv[u], su));
        }
    }
    if (lane < m) AT(A, k + 1 + lane, k) = (lane == 0) ? mkc(beta, 0.0) : mkc(0.0, 0.0);
  }
}

// Eigenvalues and eigenvectors of a Hermitian operator from its Householder reduction: the
// Hessenberg form of a Hermitian matrix is tridiagonal, T = tridiag(e, d, e) with d = Re diag and
// e the subdiagonal (the reflectors' real beta, the last entry made real by a phase). The implicit symmetric QR step with
// Wilkinson's shift (Golub & Van Loan, Algorithm 8.3.2) chases the bulge on the real pair (d, e),
// lane i holding d_i and e_i = T(i+1, i); each plane rotation is applied to Z's columns k, k+1
// (real c, s; lane r on row r, column k+1 carried in registers to the next step). Deflation by
// LAPACK's test e_k^2 <= ulp^2 |d_k d_k+1| + safmin. On return Z holds the orthonormal
// eigenvectors of the original operator and w its (real) eigenvalues; returns 0, or 1 when an
// eigenvalue did not converge in 30 n sweeps.
__device__ inline double readlane_d(double x, int src) {
  const unsigned long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffull), src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

__device__ int w_tridiag_qr(int n, const cd* H, cd* Z, cd* w) {
  const int lane = threadIdx.x & 63;
  double dl = (lane < n) ? AT(H, lane, lane).x : 0.0;
  double el = (lane + 1 < n) ? AT(H, lane + 1, lane).x : 0.0;
  // the reflectors leave every subdiagonal real but the last, H(n-1, n-2) = h (no reflector acts
  // there): T = D^H H D with D = diag(1, .., 1, h / |h|) makes it |h|, and Z takes D
  if (n >= 2) {
    const cd h = AT(H, n - 1, n - 2);
    const double ah = cabs(h);
    if (lane == n - 2) el = ah;
    if (ah > 0.0 && lane < n) AT(Z, lane, n - 1) = cmul(AT(Z, lane, n - 1), cscale(h, 1.0 / ah));
  }
  const double eps2 = ULP * ULP;
  int hi = n - 1, sweeps = 0, bad = 0;
  while (hi > 0) {
    // lo: one past the largest k < hi with a negligible e_k (0 when none)
    const double dn = __shfl_down(dl, 1);
    const bool negl = lane < hi && el * el <= eps2 * fabs(dl * dn) + SAFMIN;
    const unsigned long long mask = __ballot(negl);
    const int lo = mask ? 64 - __clzll(mask) : 0;
    if (lo > 0 && lane == lo - 1) el = 0.0;
    if (lo == hi) { --hi; continue; }
    if (++sweeps > 30 * n) { bad = 1; break; }
    // Wilkinson shift from the trailing 2 x 2
    const double a = readlane_d(dl, hi - 1), bb = readlane_d(el, hi - 1), c = readlane_d(dl, hi);
    const double delta = 0.5 * (a - c);
    const double rt = sqrt(delta * delta + bb * bb);
    const double den = delta + (delta >= 0.0 ? rt : -rt);
    const double mu = (den != 0.0) ? c - bb * bb / den : c;
    double dk = readlane_d(dl, lo), ek = readlane_d(el, lo);
    double x = dk - mu, z = ek;
    // Z's rows by lane; lanes past n read row 0 and store into the padding column, which holds
    // no entry: the stores stay unconditional, so the next column's load, issued before this
    // column's store, is waited for without draining the store
    const int zr = (lane < n) ? lane : 0;
    cd z0 = AT(Z, zr, lo);
    cd z1 = AT(Z, zr, lo + 1);
    // d_k+1 and e_k+1 are read before the step that changes them: from a copy taken here, so the
    // reads do not wait on the chain of updates
    const double dsnap = dl, esnap = el;
    for (int k = lo; k < hi; ++k) {
      const double dk1 = readlane_d(dsnap, k + 1);
      const double enx = readlane_d(esnap, k + 1);
      const double ek1 = (k + 1 < hi) ? enx : 0.0;
      const cd z2 = AT(Z, zr, (k + 2 < n) ? k + 2 : n - 1);
      // the rotation, branch-free (x = z = 0 gives the identity)
      const double rr2 = x * x + z * z;
      const bool nz = rr2 > 0.0;
      const double ir = rsqrt_pos(nz ? rr2 : 1.0);
      const double cs = nz ? x * ir : 1.0, sn = nz ? z * ir : 0.0, r = nz ? rr2 * ir : x;
      if (k > lo && lane == k - 1) el = r;
      const double c2 = cs * cs, s2 = sn * sn, cssn = cs * sn;
      const double ndk = c2 * dk + 2.0 * cssn * ek + s2 * dk1;
      const double ndk1 = s2 * dk - 2.0 * cssn * ek + c2 * dk1;
      const double nek = cssn * (dk1 - dk) + (c2 - s2) * ek;
      if (lane == k) { dl = ndk; el = nek; }
      // Z columns k, k+1: Z <- Z G^T
      cd* dst = (lane < n) ? &AT(Z, lane, k) : &AT(Z, 0, SLD - 1);
      *dst = mkc(cs * z0.x + sn * z1.x, cs * z0.y + sn * z1.y);
      z0 = mkc(cs * z1.x - sn * z0.x, cs * z1.y - sn * z0.y);
      z1 = z2;
      x = nek;
      z = sn * ek1;
      dk = ndk1;
      ek = cs * ek1;
      if (k + 1 < hi && lane == k + 1) el = ek;
      if (k + 1 == hi && lane == hi) dl = ndk1;
    }
    {
      cd* dst = (lane < n) ? &AT(Z, lane, hi) : &AT(Z, 0, SLD - 1);
      *dst = z0;
    }
  }
  if (lane < n) w[lane] = mkc(dl, 0.0);
  return bad;
}

// Frobenius norm of the Hessenberg part of A.
__device__ double w_hess_norm(int n, const cd* A) {
  double acc = 0.0;
  for (int e = threadIdx.x; e < n * n; e += ST) {
    const int i = e / n, j = e % n;
    if (i <= j + 1) acc += cnorm2(AT(A, i, j));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  return sqrt(acc);
}

// X (upper triangular) with T X = X diag(T_kk), lane per eigenvector (k_trevc's back
// substitution; divisors below max(ulp |T_kk|, smlnum) raised to that bound).
__device__ void w_trevc(int n, const cd* T, cd* X) {
  const int k = threadIdx.x & 63;
  const bool live = k < n;
  const double smlnum = SAFMIN * ((double)n / ULP);
  const cd lam = live ? AT(T, k, k) : mkc(0.0, 0.0);
  const double smin = fmax(ULP * cabs1(lam), smlnum);
  if (live)
    for (int j = 0; j < n; ++j)
      AT(X, j, k) = (j < k) ? cneg(AT(T, j, k)) : mkc(j == k ? 1.0 : 0.0, 0.0);
  for (int i = n - 2; i >= 0; --i) {
    if (live && i < k) {
      cd d = csub(AT(T, i, i), lam);
      if (cabs1(d) < smin) d = mkc(smin, 0.0);
      const cd xi = cdiv(AT(X, i, k), d);
      AT(X, i, k) = xi;
      for (int j = 0; j < i; ++j) AT(X, j, k) = csub(AT(X, j, k), cmul(AT(T, j, i), xi));
    }
  }
}

// Unit 2-norm columns with the largest-modulus component real and positive.
__device__ void w_colnorm(int n, cd* W) {
  const int c = threadIdx.x & 63;
  if (c >= n) return;
  double s = 0.0, best = -1.0;
  cd pv = mkc(1.0, 0.0);
  for (int r = 0; r < n; ++r) {
    const cd z = AT(W, r, c);
    const double a2 = cnorm2(z);
    s += a2;
    if (a2 > best) { best = a2; pv = z; }
  }
  const double inv = (s > 0.0) ? 1.0 / sqrt(s) : 1.0;
  const double pa = sqrt(best);
  const cd ph = (pa > 0.0) ? mkc(pv.x / pa, -pv.y / pa) : mkc(1.0, 0.0);
  const cd f = cscale(ph, inv);
  for (int r = 0; r < n; ++r) AT(W, r, c) = cmul(AT(W, r, c), f);
}

__device__ void copy_out(int n, const cd* A, cd* g) {
  for (int e = threadIdx.x; e < n * n; e += ST) g[e] = AT(A, e / n, e % n);
}
__device__ void copy_in(int n, const cd* g, cd* A) {
  for (int e = threadIdx.x; e < n * n; e += ST) AT(A, e / n, e % n) = g[e];
}

__global__ __launch_bounds__(ST) void k_small_1d(int hu, int L, int mode, const double* pts,
                                                 const double* thick, BatchPtr epx, BatchPtr epy,
                                                 BatchPtr epz, int* flags, cd* facp, cd* facpp,
                                                 int* pivp, int* pivpp, cd* xvs, cd* gsave,
                                                 double* eff, cd* amp, double* pe) {
  extern __shared__ unsigned char smem[];
  SmallLds S = small_carve(smem);
  const int p = blockIdx.x, lane = threadIdx.x;
  const int n = 2 * hu + 1;
  const long long nn = (long long)n * n;
  const double* pt = pts + (long long)p * PT_STRIDE;
  const int tm = (mode == 2);
  cd** M = S.M;
  if (lane < n) point_order_setup(pt, hu, 0, mode, lane, S.ov, S.dg);
  const double k0 = 2.0 * WB_PI / pt[PT_WL];
  int pflags = 0;   // the point's flags
  for (int l = 0; l < L; ++l) {
    const long long b = (long long)p * L + l;
    int lflags = 0;   // the pair's flags
    // operator: TE Kx^2 - epy -> M0; TM B = Kx epz_i Kx - I -> M1 and epx -> M2 (A = epx B)
    if (!tm) {
      const cd* EY = epy.ptr(b);
      for (int e = lane; e < n * n; e += ST) {
        const int i = e / n, j = e % n;
        cd v = cneg(EY[e]);
        if (i == j) { const cd kx = S.ov[OV_KX * n + i]; v = cadd(v, cmul(kx, kx)); }
        AT(M[0], i, j) = v;
      }
    } else {
      const cd* EZ = epz.ptr(b);
      const cd* EX = epx.ptr(b);
      for (int e = lane; e < n * n; e += ST) {
        const int i = e / n, j = e % n;
        cd v = cmul(cmul(S.ov[OV_KX * n + i], EZ[e]), S.ov[OV_KX * n + j]);
        if (i == j) v = csub(v, mkc(1.0, 0.0));
        AT(M[1], i, j) = v;
        AT(M[2], i, j) = EX[e];
      }
    }
    // The eigensolver. A lossless layer makes the problem Hermitian to within rounding: TE's
    // operator itself (eig 1: the symmetric solver, W unitary); TM's pencil, epx B with epx
    // Hermitian positive definite and B Hermitian (eig 2: with epx = C C^H the similar operator
    // C^H B C is Hermitian; its eigenvectors Y give W = C Y, V = epx^-1 W q = C^-H Y q and
    // W^-1 = Y^H C^-1). Otherwise (eig 0) the Schur iteration of the full operator.
    int eig = 0;
    if (!tm) {
      if (w_is_herm(n, M[0])) { w_symmetrize(n, M[0]); eig = 1; }
    } else {
      if (w_is_herm(n, M[1]) && w_is_herm(n, M[2])) {
        for (int e = lane; e < n * n; e += ST) AT(M[3], e / n, e % n) = AT(M[2], e / n, e % n);
        w_symmetrize(n, M[3]);
        if (w_chol(n, M[3]) == 0) eig = 2;
      }
      if (eig == 2) {
        w_symmetrize(n, M[1]);
        w_trinv_lower(n, M[3], M[6]);                    // C^-1
        w_mm_op(n, 0, M[1], 0, M[3], nullptr, M[7]);     // B C
        w_mm_op(n, 1, M[3], 0, M[7], nullptr, M[0]);     // C^H B C
        w_symmetrize(n, M[0]);
      } else {
        w_mm(n, M[2], M[1], nullptr, M[0]);              // epx B
      }
    }
    // Hessenberg with Z in M1; general: Schur form, eigenvectors X in M3, W = Z X in M4;
    // Hermitian: tridiagonal QR, the eigenvectors in M1
    w_identity(n, M[1]);
    w_gehrd(n, M[0], M[1], S);
    if (eig) {
      if (w_tridiag_qr(n, M[0], M[1], S.w)) lflags |= WB_FLAG_NOCONVERGE;
      if (eig == 1) {
        for (int e = lane; e < n * n; e += ST) AT(M[4], e / n, e % n) = AT(M[1], e / n, e % n);
        w_colnorm(n, M[4]);
      } else {
        w_mm_op(n, 0, M[3], 0, M[1], nullptr, M[4]);     // W = C Y
      }
    } else {
      const double dfl = ULP * w_hess_norm(n, M[0]);
      if (wave_qr<true>(M[0], M[1], n, SLD, n, S.w, dfl)) lflags |= WB_FLAG_NOCONVERGE;
      w_trevc(n, M[0], M[3]);
      w_mm(n, M[1], M[3], nullptr, M[4]);
      w_colnorm(n, M[4]);
    }
    if (lane < n) S.q[lane] = mode_q(S.w[lane]);
    // V = W q (TE) or epx^-1 W q (TM) -> M5 (eig 2: C^-H Y q, and W^-1 = Y^H C^-1 -> M0 while
    // C^-1 is still in M6)
    if (eig == 2) {
      w_mm_op(n, 1, M[6], 0, M[1], nullptr, M[5]);
      for (int e = lane; e < n * n; e += ST) AT(M[5], e / n, e % n) = cmul(AT(M[5], e / n, e % n), S.q[e % n]);
      w_mm_op(n, 1, M[1], 0, M[6], nullptr, M[0]);
    } else {
      for (int e = lane; e < n * n; e += ST) {
        const int i = e / n, j = e % n;
        AT(M[5], i, j) = cmul(AT(M[4], i, j), S.q[j]);
      }
      if (tm) {
        for (int e = lane; e < n * n; e += ST) AT(M[6], e / n, e % n) = AT(M[2], e / n, e % n);
        w_getrf(n, M[6], S.piv);
        w_getrs(n, n, M[6], S.piv, M[5], S);
      }
    }
    // A'^T -> M6, B'^T -> M7 (gap: wg = 1, vg = -i y)
    for (int e = lane; e < n * n; e += ST) {
      const int i = e / n, j = e % n;
      const double y = S.ov[OV_Y * n + i].x;
      const cd Pw = AT(M[4], i, j);
      const cd Qv = cmul(AT(M[5], i, j), mkc(0.0, 1.0 / y));   // V / (-i y)
      AT(M[6], j, i) = cadd(Pw, Qv);
      AT(M[7], j, i) = csub(Pw, Qv);
    }
    // W^-1 -> M0 (eig 1: W^H, W unitary; eig 2: formed above); V^-1 = q^-1 W^-1 (TE) or
    // q^-1 W^-1 epx (TM) -> M1
    int iw = 0;
    if (eig == 1) {
      for (int e = lane; e < n * n; e += ST) AT(M[0], e / n, e % n) = cconj(AT(M[4], e % n, e / n));
    } else if (eig == 0) {
      iw = w_getrf(n, M[4], S.piv);
      w_identity(n, M[0]);
      w_getrs(n, n, M[4], S.piv, M[0], S);
    }
    if (tm) w_mm(n, M[0], M[2], nullptr, M[1]);
    int qz = 0;
    for (int e = lane; e < n * n; e += ST) {
      const int i = e / n, j = e % n;
      const cd qi = S.q[i];
      if (qi.x == 0.0 && qi.y == 0.0) qz = 1;
      AT(M[1], i, j) = cdiv(tm ? AT(M[1], i, j) : AT(M[0], i, j), qi);
    }
    if (iw || __ballot(qz)) lflags |= WB_FLAG_SINGULAR;
    // A^T -> M2, B^T -> M3: A^T[i][j] = wg_i W^-1[j][i] + vg_i V^-1[j][i]
    for (int e = lane; e < n * n; e += ST) {
      const int i = e / n, j = e % n;
      const double y = S.ov[OV_Y * n + i].x;
      const cd P0 = AT(M[0], j, i), Q0 = cmul(mkc(0.0, -y), AT(M[1], j, i));
      AT(M[2], i, j) = cadd(P0, Q0);
      AT(M[3], i, j) = csub(P0, Q0);
    }
    if (lane < n) {
      const cd x = cexp_(cscale(S.q[lane], -k0 * thick[b]));
      S.xv[lane] = x;
      xvs[b * n + lane] = x;
    }
    // G^T -> M4: Gamma_0 = diag(r22), later the previous layer's
    if (l == 0) {
      for (int e = lane; e < n * n; e += ST) {
        const int i = e / n, j = e % n;
        AT(M[4], i, j) = (i == j) ? S.dg[DG_R22 * n + i] : mkc(0.0, 0.0);
      }
    } else {
      copy_in(n, gsave + (long long)p * nn, M[4]);
    }
    // P^T = A^T + G^T B^T -> M5, Q^T = B^T + G^T A^T -> M0; R^T = X P^-T Q^T X
    w_mm(n, M[4], M[3], M[2], M[5]);
    w_mm(n, M[4], M[2], M[3], M[0]);
    const int ip = w_getrf(n, M[5], S.piv);
    copy_out(n, M[5], facp + b * nn);
    if (lane < n) pivp[b * n + lane] = S.piv[lane];
    w_getrs(n, n, M[5], S.piv, M[0], S);
    for (int e = lane; e < n * n; e += ST) {
      const int i = e / n, j = e % n;
      AT(M[0], i, j) = cmul(cmul(S.xv[i], AT(M[0], i, j)), S.xv[j]);
    }
    // P'^T = A'^T + R^T B'^T -> M1, G'^T = B'^T + R^T A'^T -> M4; G^T = P'^-T G'^T
    w_mm(n, M[0], M[7], M[6], M[1]);
    w_mm(n, M[0], M[6], M[7], M[4]);
    const int ipp = w_getrf(n, M[1], S.piv);
    copy_out(n, M[1], facpp + b * nn);
    if (lane < n) pivpp[b * n + lane] = S.piv[lane];
    w_getrs(n, n, M[1], S.piv, M[4], S);
    if (l + 1 < L) copy_out(n, M[4], gsave + (long long)p * nn);
    if (ip || ipp) pflags |= WB_FLAG_SINGULAR;
    if (lane == 0 && flags) flags[b] = lflags;
  }
  // final interface: M^T = I - G^T diag(t11) -> M1
  for (int e = lane; e < n * n; e += ST) {
    const int i = e / n, j = e % n;
    const cd v = cmul(AT(M[4], i, j), S.dg[DG_T11 * n + j]);
    AT(M[1], i, j) = csub(mkc(i == j ? 1.0 : 0.0, 0.0), v);
  }
  if (w_getrf(n, M[1], S.piv)) pflags |= WB_FLAG_SINGULAR;
  const int hot = n / 2;
  const double sig = (S.ov[OV_KX * n + hot].x < 0.0) ? -1.0 : 1.0;
  cd x = (lane == hot) ? cscale(S.dg[DG_T12 * n + hot], sig) : mkc(0.0, 0.0);
  x = w_getrs_t(n, M[1], S.piv, x);              // u
  // G u (G = (G^T)^T) and the reflected column
  cd gu = mkc(0.0, 0.0);
  for (int c = 0; c < n; ++c) {
    const cd uc = readlane_c(x, c);
    if (lane < n) gu = cadd(gu, cmul(AT(M[4], c, lane), uc));
  }
  if (lane < n) {
    cd v = cmul(S.dg[DG_T21 * n + lane], gu);
    if (lane == hot) v = cadd(v, cscale(S.dg[DG_T22 * n + hot], sig));
    S.va[lane] = v;
  }
  // back-sweep: d_l = 4 P_l^-1 X_l P'_l^-1 d_{l+1}, d_L = u
  cd d = x;
  for (int l = L - 1; l >= 0; --l) {
    const long long b = (long long)p * L + l;
    copy_in(n, facpp + b * nn, M[2]);
    copy_in(n, facp + b * nn, M[3]);
    if (lane < n) { S.piv[lane] = pivpp[b * n + lane]; S.piv2[lane] = pivp[b * n + lane]; }
    d = w_getrs_t(n, M[2], S.piv, d);
    if (lane < n) d = cscale(cmul(xvs[b * n + lane], d), 4.0);
    d = w_getrs_t(n, M[3], S.piv2, d);
  }
  if (lane < n) S.vb[lane] = cmul(S.dg[DG_R12 * n + lane], d);
  if (lane == 0 && flags && pflags)
    for (int l = 0; l < L; ++l) flags[(long long)p * L + l] |= pflags;
  point_efficiency<ST>(p, hu, 0, mode, pts, S.ov, S.va, S.vb, eff, amp, pe, S.red);
}

}  // namespace

int wb_small_1d_eligible(const wb_solve_args& a) {
  if (a.mode == 0 || a.hv != 0 || 2 * a.hu + 1 > SN || a.L < 1) return 0;
  if (a.steps || a.modes_in || a.modes_out || a.ck || a.smat || a.field_amps || a.field_ws) return 0;
  return 1;
}

hipError_t wb_small_1d_launch(const wb_solve_args& a, const wb_solve_ws& W, BatchPtr epx, BatchPtr epy,
                              BatchPtr epz, hipStream_t st) {
  const size_t sh = small_lds_bytes();
  hipLaunchKernelGGL(k_small_1d, dim3(a.P), dim3(ST), sh, st, a.hu, a.L, a.mode, a.pts, a.thick, epx,
                     epy, epz, a.flags, W.M[0], W.M[1], W.ipiv[0], W.ipiv[1], W.q, W.G0, a.eff, a.amp,
                     a.pe);
  return hipGetLastError();
}
