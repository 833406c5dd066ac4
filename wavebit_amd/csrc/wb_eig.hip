// Batched general complex FP64 eigensolver: A = W diag(w) W^-1 for square, contiguous,
// row-major A (n x n, ld = n). One workgroup (256 threads, 4 waves) owns one matrix through
// every stage, so a batch of B matrices fills B compute units with no inter-workgroup sync.
//
//   k_gehrd    Householder reduction to upper Hessenberg form, A := Q^H A Q. Each column's
//              reflector is applied from the right (a wave per row: coalesced dot + update)
//              and then from the left (a thread per column: coalesced column sweep).
//   k_unghr    forms Q explicitly from the stored reflectors (backward accumulation) and
//              clears the reflector storage below the subdiagonal.
//   k_hqr_ms   (wb_hqr.hip) windowed small-bulge multishift QR iteration with Schur vectors
//              accumulated into Z; small active blocks finish with single-shift QR
//              (Wilkinson shifts, Ahues-Tisseur deflation test; Golub & Van Loan 7.5).
//   k_trevc    eigenvectors of the triangular Schur factor by back substitution, one thread
//              per eigenvector, synchronised by row so column i of T is a broadcast read.
//   (GEMM)     W = Z X on the matrix cores, then k_colnorm scales each column to unit
//              2-norm with its largest component real, as LAPACK's zgeev returns them.
#include "wb_internal.h"
#include "wb_timing.h"
#include <float.h>


namespace {

constexpr int NT = 256;
constexpr int NTH = 1024;   // Hessenberg stages: 16 waves per matrix to hide global latency
// From this order on the Hessenberg reduction and Q run blocked (panel kernel + GEMMs); the
// scratch they need (n x HNB x 4 + panels x HNB^2) must fit the n x n buffer X.
constexpr int WB_BLOCKED_HESSENBERG_MIN = 192;
constexpr double ULP = DBL_EPSILON;            // 2^-52, LAPACK dlamch('P')
constexpr double SAFMIN = DBL_MIN;

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block-wide sum of one double per thread (all threads receive the total).
template <int NTB>
__device__ inline double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < NTB / 64; ++w) s += red[w];
  return s;
}

// ----------------------------------------------------------------------------- Hessenberg

__global__ __launch_bounds__(NTH) void k_gehrd(int n, cd* Ab, cd* taub) {
  const long long b = blockIdx.x;
  cd* A = Ab + b * (long long)n * n;
  cd* tau = taub + b * (long long)n;
  extern __shared__ cd smem[];
  cd* v = smem;            // reflector, indices k+1..n-1 stored at v[0..m)
  __shared__ double red[NTH / 64];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;   // reflector length
    // --- generate H = I - tau v v^H with H^H x = beta e1, x = A[k+1:n, k]
    double ss = 0.0;
    for (int i = k + 2 + t; i < n; i += NTH) ss += cnorm2(A[(long long)i * n + k]);
    const double xnorm2 = block_sum<NTH>(ss, red);
    const cd alpha = A[(long long)(k + 1) * n + k];
    cd tk;
    double beta;
    cd scale;
    if (xnorm2 == 0.0 && alpha.y == 0.0) {
      tk = mkc(0.0, 0.0);
      beta = alpha.x;
      scale = mkc(0.0, 0.0);
    } else {
      beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xnorm2), alpha.x);
      tk = mkc((beta - alpha.x) / beta, -alpha.y / beta);
      scale = crecip(mkc(alpha.x - beta, alpha.y));
    }
    for (int i = t; i < m; i += NTH) {
      if (i == 0) v[0] = mkc(1.0, 0.0);
      else v[i] = cmul(A[(long long)(k + 1 + i) * n + k], scale);
    }
    __syncthreads();
    if (t == 0) tau[k] = tk;
    if (tk.x != 0.0 || tk.y != 0.0) {
      // --- right: A[:, k+1:n] -= tau (A[:, k+1:n] v) v^H, a wave per row
      for (int r = wave; r < n; r += NTH / 64) {
        cd* row = A + (long long)r * n + k + 1;
        double zr = 0.0, zi = 0.0;
        for (int j = lane; j < m; j += 64) {
          cd p = cmul(row[j], v[j]);
          zr += p.x; zi += p.y;
        }
        zr = wave_sum(zr); zi = wave_sum(zi);
        const cd z = cmul(tk, mkc(zr, zi));
        for (int j = lane; j < m; j += 64) row[j] = csub(row[j], cmul(z, cconj(v[j])));
      }
      __syncthreads();
      // --- left: A[k+1:n, k+1:n] -= conj(tau) v (v^H A), a thread per column
      const cd tc = cconj(tk);
      for (int j = k + 1 + t; j < n; j += NTH) {
        double sr = 0.0, si = 0.0;
        for (int i = 0; i < m; ++i) {
          cd p = cmulc(v[i], A[(long long)(k + 1 + i) * n + j]);
          sr += p.x; si += p.y;
        }
        const cd s = cmul(tc, mkc(sr, si));
        for (int i = 0; i < m; ++i) {
          cd* p = A + (long long)(k + 1 + i) * n + j;
          *p = csub(*p, cmul(v[i], s));
        }
      }
    }
    __syncthreads();
    // --- store beta and the reflector below the subdiagonal
    for (int i = t; i < m; i += NTH) {
      cd* p = A + (long long)(k + 1 + i) * n + k;
      *p = (i == 0) ? mkc(beta, 0.0) : v[i];
    }
    __syncthreads();
  }
  if (t == 0 && n >= 2) tau[n - 2] = mkc(0.0, 0.0);
}

// Q = H_0 H_1 ... H_{n-3}; clears the reflectors below the subdiagonal of A afterwards.
__global__ __launch_bounds__(NTH) void k_unghr(int n, cd* Ab, const cd* taub, cd* Qb) {
  const long long b = blockIdx.x;
  cd* A = Ab + b * (long long)n * n;
  const cd* tau = taub + b * (long long)n;
  cd* Q = Qb + b * (long long)n * n;
  extern __shared__ cd v[];
  const int t = threadIdx.x;
  for (long long e = t; e < (long long)n * n; e += NTH) {
    const int r = (int)(e / n), c = (int)(e % n);
    Q[e] = mkc(r == c ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  for (int k = n - 3; k >= 0; --k) {
    const int m = n - k - 1;
    const cd tk = tau[k];
    for (int i = t; i < m; i += NTH) v[i] = (i == 0) ? mkc(1.0, 0.0) : A[(long long)(k + 1 + i) * n + k];
    __syncthreads();
    if (tk.x != 0.0 || tk.y != 0.0) {
      for (int j = k + 1 + t; j < n; j += NTH) {
        double sr = 0.0, si = 0.0;
        for (int i = 0; i < m; ++i) {
          cd p = cmulc(v[i], Q[(long long)(k + 1 + i) * n + j]);
          sr += p.x; si += p.y;
        }
        const cd s = cmul(tk, mkc(sr, si));
        for (int i = 0; i < m; ++i) {
          cd* p = Q + (long long)(k + 1 + i) * n + j;
          *p = csub(*p, cmul(v[i], s));
        }
      }
    }
    __syncthreads();
  }
  for (long long e = t; e < (long long)n * n; e += NTH) {
    const int r = (int)(e / n), c = (int)(e % n);
    if (r > c + 1) A[e] = mkc(0.0, 0.0);
  }
}

// ----------------------------------------------------------------------------- blocked Hessenberg
// Panel of HNB columns (LAPACK zlahr2's scheme, written for one workgroup per matrix): for
// column c = k0 + i, apply the panel's earlier reflectors (from the right through
// Y = A V T, from the left through V T^H V^H), generate reflector i from A[c+1:, c], and form
// Y[k0+1:, i] = tau (A[k0+1:, c+1:] v - Y (V^H v)) -- the one matrix-vector product per column
// that streams the trailing block (its columns are untouched within the panel). The trailing
// updates A -= Y V^H and A -= V T^H V^H A are batched GEMMs on the matrix cores.
constexpr int HNB = 32;
constexpr int HNT = 512;     // panel threads: 8 waves, each with a full row of GEMV loads in flight

// sum_{j < m} x[j] w[j] over a contiguous row segment x (m <= HNB) in unrolled chunks of 8 with
// the loads unconditional (x has HNB entries) and the terms j >= m selected away, so a chunk's
// loads are in flight together.
__device__ inline cd row_dot(const cd* x, const cd* w, int m) {
  cd acc0 = mkc(0.0, 0.0), acc1 = mkc(0.0, 0.0);
  for (int j0 = 0; j0 < m; j0 += 8) {
#pragma unroll
    for (int u = 0; u < 8; u += 2) {
      const int j = j0 + u;
      const cd p0 = cmul(x[j], w[j]), p1 = cmul(x[j + 1], w[j + 1]);
      if (j < m) acc0 = cadd(acc0, p0);
      if (j + 1 < m) acc1 = cadd(acc1, p1);
    }
  }
  return cadd(acc0, acc1);
}
// sum over rows r = r0, r0 + st, ... < rend of conj(v[r n]) a[r n] (column entries with row
// stride n), four rows in flight.
__device__ inline cd col_dotc(const cd* v, const cd* a, int n, int r0, int rend, int st) {
  cd acc[4] = {mkc(0.0, 0.0), mkc(0.0, 0.0), mkc(0.0, 0.0), mkc(0.0, 0.0)};
  int r = r0;
  for (; r + 3 * st < rend; r += 4 * st) {
#pragma unroll
    for (int u = 0; u < 4; ++u)
      acc[u] = cadd(acc[u], cmulc(v[(long long)(r + st * u) * n], a[(long long)(r + st * u) * n]));
  }
  for (; r < rend; r += st) acc[0] = cadd(acc[0], cmulc(v[(long long)r * n], a[(long long)r * n]));
  return cadd(cadd(acc[0], acc[1]), cadd(acc[2], acc[3]));
}
// the same with the second factor a vector in LDS indexed by row
__device__ inline cd col_dotc_lds(const cd* v, const cd* s, int n, int r0, int rend, int st) {
  cd acc[4] = {mkc(0.0, 0.0), mkc(0.0, 0.0), mkc(0.0, 0.0), mkc(0.0, 0.0)};
  int r = r0;
  for (; r + 3 * st < rend; r += 4 * st) {
#pragma unroll
    for (int u = 0; u < 4; ++u) acc[u] = cadd(acc[u], cmulc(v[(long long)(r + st * u) * n], s[r + st * u]));
  }
  for (; r < rend; r += st) acc[0] = cadd(acc[0], cmulc(v[(long long)r * n], s[r]));
  return cadd(cadd(acc[0], acc[1]), cadd(acc[2], acc[3]));
}

// Y and T live in the per-matrix scratch X, whose batch stride is x_stride.
__global__ __launch_bounds__(HNT) void k_hrd_panel(int n, int k0, int ib, cd* Ab, cd* Vb, cd* Yb,
                                                   cd* Tb, long long x_stride) {
  const long long b = blockIdx.x;
  cd* A = Ab + b * (long long)n * n;
  cd* V = Vb + b * (long long)n * n;            // column k0 + j holds reflector j (dense)
  cd* Y = Yb + b * x_stride;                     // n x HNB
  cd* Tg = Tb + b * x_stride;                    // this panel's T, HNB x HNB
  __shared__ cd sT[HNB][HNB + 1];
  __shared__ cd red[HNT / 32][HNB + 1];
  __shared__ cd sa[HNB], sw[HNB];
  __shared__ double sred[HNT / 64];
  __shared__ cd s_tau;
  extern __shared__ cd sv[];                     // reflector i by absolute row, n entries
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int jj = t & 31, ch = t >> 5;            // (reflector index, row chunk) for reductions
  for (int e = t; e < HNB * (HNB + 1); e += HNT) sT[e / (HNB + 1)][e % (HNB + 1)] = mkc(0.0, 0.0);
  __syncthreads();
  for (int i = 0; i < ib; ++i) {
    const int c = k0 + i;
    if (i > 0) {
      // right: A[r][c] -= sum_j Y[r][j] conj(V[c][k0+j]), rows k0+1..n-1
      if (t < i) sa[t] = cconj(V[(long long)c * n + k0 + t]);
      __syncthreads();
      for (int r = k0 + 1 + t; r < n; r += HNT)
        A[(long long)r * n + c] = csub(A[(long long)r * n + c], row_dot(Y + (long long)r * HNB, sa, i));
      __syncthreads();
      // left: w = V^H A[:, c] (rows k0+1..), w = T^H w, A[:, c] -= V w
      {
        cd acc = mkc(0.0, 0.0);
        if (jj < i) acc = col_dotc(V + k0 + jj, A + c, n, k0 + 1 + ch, n, HNT / 32);
        red[ch][jj] = acc;
      }
      __syncthreads();
      if (t < i) {
        cd s = mkc(0.0, 0.0);
        for (int q = 0; q < HNT / 32; ++q) s = cadd(s, red[q][t]);
        sw[t] = s;
      }
      __syncthreads();
      if (t < i) {
        cd s = mkc(0.0, 0.0);
        for (int l = 0; l <= t; ++l) s = cadd(s, cmulc(sT[l][t], sw[l]));
        sa[t] = s;
      }
      __syncthreads();
      for (int r = k0 + 1 + t; r < n; r += HNT)
        A[(long long)r * n + c] = csub(A[(long long)r * n + c], row_dot(V + (long long)r * n + k0, sa, i));
      __syncthreads();
    }
    // reflector from x = A[c+1:, c]
    double ss = 0.0;
    for (int r = c + 2 + t; r < n; r += HNT) ss += cnorm2(A[(long long)r * n + c]);
    const double xnorm2 = block_sum<HNT>(ss, sred);
    const cd alpha = A[(long long)(c + 1) * n + c];
    cd tk, scale;
    double beta;
    if (xnorm2 == 0.0 && alpha.y == 0.0) {
      tk = mkc(0.0, 0.0);
      beta = alpha.x;
      scale = mkc(0.0, 0.0);
    } else {
      beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xnorm2), alpha.x);
      tk = mkc((beta - alpha.x) / beta, -alpha.y / beta);
      scale = crecip(mkc(alpha.x - beta, alpha.y));
    }
    for (int r = c + 1 + t; r < n; r += HNT)
      sv[r] = (r == c + 1) ? mkc(1.0, 0.0) : cmul(A[(long long)r * n + c], scale);
    __syncthreads();
    for (int r = t; r < n; r += HNT) {
      V[(long long)r * n + k0 + i] = (r <= c) ? mkc(0.0, 0.0) : sv[r];
      if (r == c + 1) A[(long long)r * n + c] = mkc(beta, 0.0);
      else if (r > c + 1) A[(long long)r * n + c] = mkc(0.0, 0.0);
    }
    // Y[r][i] (unscaled) = A[r, c+1:] v, rows k0+1..n-1: a wave per row, all of a lane's loads
    // of the row (up to GC) issued at once (the GEMV is the HBM-bound part of the reduction:
    // the trailing matrix is streamed once per column).
    {
      constexpr int GC = 12;
      const int cb = c + 1 + lane;
      for (int r = k0 + 1 + wave; r < n; r += HNT / 64) {
        const cd* row = A + (long long)r * n;
        cd a[GC];
#pragma unroll
        for (int u = 0; u < GC; ++u) a[u] = (cb + 64 * u < n) ? row[cb + 64 * u] : mkc(0.0, 0.0);
        double zr0 = 0.0, zi0 = 0.0, zr1 = 0.0, zi1 = 0.0;
#pragma unroll
        for (int u = 0; u < GC; u += 2) {
          const cd v0 = (cb + 64 * u < n) ? sv[cb + 64 * u] : mkc(0.0, 0.0);
          const cd v1 = (cb + 64 * (u + 1) < n) ? sv[cb + 64 * (u + 1)] : mkc(0.0, 0.0);
          const cd p0 = cmul(a[u], v0), p1 = cmul(a[u + 1], v1);
          zr0 += p0.x; zi0 += p0.y;
          zr1 += p1.x; zi1 += p1.y;
        }
        for (int col = cb + 64 * GC; col < n; col += 64) {
          const cd p = cmul(row[col], sv[col]);
          zr0 += p.x;
          zi0 += p.y;
        }
        const double sr = wave_sum(zr0 + zr1);
        const double si = wave_sum(zi0 + zi1);
        if (lane == 0) Y[(long long)r * HNB + i] = mkc(sr, si);
      }
    }
    // tv = V[:, 0:i]^H v over rows c+1..n-1
    {
      cd acc = mkc(0.0, 0.0);
      if (jj < i) acc = col_dotc_lds(V + k0 + jj, sv, n, c + 1 + ch, n, HNT / 32);
      red[ch][jj] = acc;
    }
    if (t == 0) s_tau = tk;
    __syncthreads();
    if (t < i) {
      cd s = mkc(0.0, 0.0);
      for (int q = 0; q < HNT / 32; ++q) s = cadd(s, red[q][t]);
      sw[t] = s;
    }
    __syncthreads();
    // Y[:, i] = tau (y - Y[:, 0:i] tv) ; T[0:i, i] = -tau T[0:i, 0:i] tv ; T[i][i] = tau
    for (int r = k0 + 1 + t; r < n; r += HNT) {
      const cd acc = csub(Y[(long long)r * HNB + i], row_dot(Y + (long long)r * HNB, sw, i));
      Y[(long long)r * HNB + i] = cmul(tk, acc);
    }
    if (t < i) {
      cd s = mkc(0.0, 0.0);
      for (int l = t; l < i; ++l) s = cadd(s, cmul(sT[t][l], sw[l]));
      sT[t][i] = cneg(cmul(tk, s));
    }
    if (t == 0) sT[i][i] = tk;
    __syncthreads();
  }
  for (int e = t; e < HNB * HNB; e += HNT) Tg[e] = sT[e / HNB][e % HNB];
}

// Zero a (rows x cols) block per batch element.
__global__ void k_zero_block(int rows, int cols, cd* Xb, long long stride, int ld) {
  const long long b = blockIdx.y;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)rows * cols) return;
  Xb[b * stride + (e / cols) * ld + e % cols] = mkc(0.0, 0.0);
}

__global__ void k_identity(int n, cd* Qb) {
  const long long b = blockIdx.y;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n * n) return;
  Qb[b * (long long)n * n + e] = mkc((e / n) == (e % n) ? 1.0 : 0.0, 0.0);
}

// ----------------------------------------------------------------------------- eigenvectors

// X (n x n, upper triangular) with T X = X diag(T_kk): thread per eigenvector k, rows swept
// bottom-up in lockstep so column i of T is read once per step by every lane.
__global__ __launch_bounds__(NT) void k_trevc(int n, const cd* Tb, cd* Xb) {
  const long long b = blockIdx.x;
  const cd* T = Tb + b * (long long)n * n;
  cd* X = Xb + b * (long long)n * n;
  const int t = threadIdx.x;
  const double smlnum = SAFMIN * ((double)n / ULP);
  for (int k0 = 0; k0 < n; k0 += NT) {
    const int k = k0 + t;
    const bool live = k < n;
    cd lam = live ? T[(long long)k * n + k] : mkc(0.0, 0.0);
    const double smin = fmax(ULP * cabs1(lam), smlnum);
    // right-hand side b_j = -T[j][k], j < k
    if (live) {
      for (int j = 0; j < n; ++j) {
        cd val;
        if (j < k) val = cneg(T[(long long)j * n + k]);
        else if (j == k) val = mkc(1.0, 0.0);
        else val = mkc(0.0, 0.0);
        X[(long long)j * n + k] = val;
      }
    }
    const int top = (k0 + NT < n ? k0 + NT : n) - 1;   // largest k in this chunk
    for (int i = top - 1; i >= 0; --i) {
      if (live && i < k) {
        cd d = csub(T[(long long)i * n + i], lam);
        if (cabs1(d) < smin) d = mkc(smin, 0.0);
        const cd xi = cdiv(X[(long long)i * n + k], d);
        X[(long long)i * n + k] = xi;
        for (int j = 0; j < i; ++j) {
          cd* p = X + (long long)j * n + k;
          *p = csub(*p, cmul(T[(long long)j * n + i], xi));
        }
      }
    }
  }
}

// Blocked eigenvectors of the triangular Schur factor. Column k of X solves
// (T - T_kk I) x = 0 with x_k = 1 and x_j = 0 for j > k. Row blocks of TB rows are solved from
// the bottom: the coupling of block I to the rows already solved, S = T[I, i1:] X[i1:, i1:], is
// one batched GEMM on the matrix cores; k_trevc_block then finishes each column's TB x TB
// shifted triangular solve in registers (thread per column, T_II in LDS). Divisors smaller
// than max(ulp |T_kk|, smlnum) are replaced by that bound, as in LAPACK's ztrevc.
constexpr int TB = 32;

__global__ __launch_bounds__(256) void k_trevc_block(int n, int i0, const cd* Tb, cd* Xb,
                                                     const cd* Sb, long long s_stride) {
  const long long b = blockIdx.y;
  const cd* T = Tb + b * (long long)n * n;
  cd* X = Xb + b * (long long)n * n;
  const cd* S = Sb + b * s_stride;                  // TB x (n - i1), ld n
  const int i1 = (i0 + TB < n) ? i0 + TB : n;
  const int nbk = i1 - i0;
  __shared__ cd Tii[TB][TB + 1];
  for (int e = threadIdx.x; e < TB * TB; e += 256) {
    const int r = e / TB, c = e % TB;
    Tii[r][c] = (r < nbk && c < nbk) ? T[(long long)(i0 + r) * n + i0 + c] : mkc(0.0, 0.0);
  }
  __syncthreads();
  const int k = i0 + blockIdx.x * 256 + threadIdx.x;   // columns k >= i0
  if (k >= n) return;
  const cd lam = T[(long long)k * n + k];
  const double smlnum = SAFMIN * ((double)n / ULP);
  const double smin = fmax(ULP * cabs1(lam), smlnum);
  cd x[TB];
#pragma unroll
  for (int jj = TB - 1; jj >= 0; --jj) {
    const int j = i0 + jj;
    cd val = mkc(0.0, 0.0);
    if (jj < nbk && j <= k) {
      if (j == k) {
        val = mkc(1.0, 0.0);
      } else {
        cd acc = (k >= i1) ? S[(long long)jj * n + (k - i1)] : mkc(0.0, 0.0);
#pragma unroll
        for (int ll = jj + 1; ll < TB; ++ll)
          if (ll < nbk && i0 + ll <= k) acc = cadd(acc, cmul(Tii[jj][ll], x[ll]));
        cd d = csub(Tii[jj][jj], lam);
        if (cabs1(d) < smin) d = mkc(smin, 0.0);
        val = cneg(cdiv(acc, d));
      }
    }
    x[jj] = val;
  }
#pragma unroll
  for (int jj = 0; jj < TB; ++jj)
    if (jj < nbk && i0 + jj <= k) X[(long long)(i0 + jj) * n + k] = x[jj];
}

// Unit 2-norm columns with the largest-modulus component made real and positive.
__global__ __launch_bounds__(NT) void k_colnorm(int n, cd* Wb) {
  const long long b = blockIdx.x;
  cd* W = Wb + b * (long long)n * n;
  for (int c = threadIdx.x; c < n; c += NT) {
    double s = 0.0, best = -1.0;
    cd piv = mkc(1.0, 0.0);
    for (int r = 0; r < n; ++r) {
      const cd z = W[(long long)r * n + c];
      const double a2 = cnorm2(z);
      s += a2;
      if (a2 > best) { best = a2; piv = z; }
    }
    const double inv = (s > 0.0) ? 1.0 / sqrt(s) : 1.0;
    const double pa = sqrt(best);
    const cd ph = (pa > 0.0) ? mkc(piv.x / pa, -piv.y / pa) : mkc(1.0, 0.0);  // conj(piv)/|piv|
    const cd f = cscale(ph, inv);
    for (int r = 0; r < n; ++r) {
      cd* p = W + (long long)r * n + c;
      *p = cmul(*p, f);
    }
  }
}

}  // namespace

// Blocked reduction of every A to Hessenberg form: panels of HNB columns, each a panel kernel
// followed by the GEMM updates of the rest of the matrix. V (n x n per matrix) receives the
// reflectors as dense columns; X holds Y (n x HNB), the per-panel T factors and GEMM scratch.
static hipError_t wb_gehrd_blocked(int n, cd* A, cd* V, cd* X, int batch, hipStream_t stream) {
  const double n3 = (double)n * n * n;
  WbStageTimer stage(WB_T_GEHRD, stream, 40.0 / 3.0 * n3 * batch);
  const long long nn = (long long)n * n;
  const int npan = (n - 2 + HNB - 1) / HNB;
  cd* Y = X;                                             // n x HNB
  cd* T = Y + (long long)n * HNB;                        // npan x HNB x HNB
  cd* Wm = T + (long long)npan * HNB * HNB;              // HNB x n
  cd* Wm2 = Wm + (long long)HNB * n;                     // HNB x n
  const cd one = mkc(1.0, 0.0), zero = mkc(0.0, 0.0), mone = mkc(-1.0, 0.0);
  auto S = [&](cd* p, long long stride) { return bp_strided(p, stride); };
  hipError_t e;
  hipLaunchKernelGGL(k_zero_block, dim3((unsigned)((nn + 255) / 256), batch), dim3(256), 0, stream,
                     n, n, V, nn, n);
  WB_CHECK_LAUNCH();
  int p = 0;
  for (int k0 = 0; k0 < n - 2; k0 += HNB, ++p) {
    const int ib = (n - 2 - k0 < HNB) ? n - 2 - k0 : HNB;
    cd* Tp = T + (long long)p * HNB * HNB;
    hipLaunchKernelGGL(k_hrd_panel, dim3(batch), dim3(HNT), sizeof(cd) * (size_t)n, stream, n, k0, ib,
                       A, V, Y, Tp, nn);
    WB_CHECK_LAUNCH();
    const int top = k0 + 1;                 // rows 0..k0
    const int m1 = n - k0 - 1;              // rows k0+1..n-1
    const int rest = n - k0 - ib;           // trailing columns k0+ib..n-1
    // Y[0:top] = (A[0:top, k0+1:] V[k0+1:, panel]) T
    if ((e = wb_zgemm_ops_launch(0, 0, top, ib, m1, one, S(A + k0 + 1, nn), n,
                                 S(V + (long long)(k0 + 1) * n + k0, nn), n, zero, S(Wm, nn), ib,
                                 batch, stream))) return e;
    if ((e = wb_zgemm_ops_launch(0, 0, top, ib, ib, one, S(Wm, nn), ib, S(Tp, nn), HNB, zero,
                                 S(Y, nn), HNB, batch, stream))) return e;
    // panel columns k0+1..k0+ib-1, rows 0..k0: A -= Y[0:top] V[k0+1:k0+ib, panel]^H
    if (ib > 1 && (e = wb_zgemm_ops_launch(0, 1, top, ib - 1, ib, mone, S(Y, nn), HNB,
                                           S(V + (long long)(k0 + 1) * n + k0, nn), n, one,
                                           S(A + k0 + 1, nn), n, batch, stream))) return e;
    if (rest <= 0) continue;
    // right: A[:, k0+ib:] -= Y V[k0+ib:, panel]^H
    if ((e = wb_zgemm_ops_launch(0, 1, n, rest, ib, mone, S(Y, nn), HNB,
                                 S(V + (long long)(k0 + ib) * n + k0, nn), n, one,
                                 S(A + k0 + ib, nn), n, batch, stream))) return e;
    // left: A[k0+1:, k0+ib:] -= V T^H (V^H A[k0+1:, k0+ib:])
    if ((e = wb_zgemm_ops_launch(1, 0, ib, rest, m1, one, S(V + (long long)(k0 + 1) * n + k0, nn), n,
                                 S(A + (long long)(k0 + 1) * n + k0 + ib, nn), n, zero, S(Wm, nn), n,
                                 batch, stream))) return e;
    if ((e = wb_zgemm_ops_launch(1, 0, ib, rest, ib, one, S(Tp, nn), HNB, S(Wm, nn), n, zero,
                                 S(Wm2, nn), n, batch, stream))) return e;
    if ((e = wb_zgemm_ops_launch(0, 0, m1, rest, ib, mone, S(V + (long long)(k0 + 1) * n + k0, nn), n,
                                 S(Wm2, nn), n, one, S(A + (long long)(k0 + 1) * n + k0 + ib, nn), n,
                                 batch, stream))) return e;
  }
  return hipSuccess;
}

// Q^H for Q = (I - V_0 T_0 V_0^H) (I - V_1 T_1 V_1^H) ... from the blocked reduction's V and T:
// N = I, then for the panels from the last to the first N := N (I - V_p T_p^H V_p^H) on the
// trailing block N[k0+1:, k0+1:] (the only part of N that is not yet the identity).
static hipError_t wb_unghr_blocked(int n, cd* V, cd* X, cd* Qh, int batch, hipStream_t stream) {
  const double n3 = (double)n * n * n;
  WbStageTimer stage(WB_T_UNGHR, stream, 16.0 / 3.0 * n3 * batch);
  const long long nn = (long long)n * n;
  const int npan = (n - 2 + HNB - 1) / HNB;
  cd* T = X + (long long)n * HNB;
  cd* Wm = T + (long long)npan * HNB * HNB;             // n x HNB
  cd* Wm2 = Wm + (long long)HNB * n;                     // n x HNB
  const cd one = mkc(1.0, 0.0), zero = mkc(0.0, 0.0), mone = mkc(-1.0, 0.0);
  auto S = [&](cd* p, long long stride) { return bp_strided(p, stride); };
  hipError_t e;
  hipLaunchKernelGGL(k_identity, dim3((unsigned)((nn + 255) / 256), batch), dim3(256), 0, stream, n, Qh);
  WB_CHECK_LAUNCH();
  for (int p = npan - 1; p >= 0; --p) {
    const int k0 = p * HNB;
    const int ib = (n - 2 - k0 < HNB) ? n - 2 - k0 : HNB;
    const int m1 = n - k0 - 1;
    cd* Tp = T + (long long)p * HNB * HNB;
    cd* Vp = V + (long long)(k0 + 1) * n + k0;
    cd* Ns = Qh + (long long)(k0 + 1) * n + k0 + 1;
    if ((e = wb_zgemm_ops_launch(0, 0, m1, ib, m1, one, S(Ns, nn), n, S(Vp, nn), n, zero, S(Wm, nn), HNB,
                                 batch, stream))) return e;
    if ((e = wb_zgemm_ops_launch(0, 1, m1, ib, ib, one, S(Wm, nn), HNB, S(Tp, nn), HNB, zero, S(Wm2, nn),
                                 HNB, batch, stream))) return e;
    if ((e = wb_zgemm_ops_launch(0, 1, m1, m1, ib, mone, S(Wm2, nn), HNB, S(Vp, nn), n, one, S(Ns, nn), n,
                                 batch, stream))) return e;
  }
  return hipSuccess;
}

// Blocked triangular eigenvectors X of the Schur factor T (both n x n per matrix); S (n x n per
// matrix, TB rows used) is scratch for the block couplings.
static hipError_t wb_trevc_blocked(int n, const cd* T, cd* X, cd* S, int batch, hipStream_t stream) {
  const double n3 = (double)n * n * n;
  WbStageTimer stage(WB_T_TREVC, stream, 4.0 / 3.0 * n3 * batch);
  const long long nn = (long long)n * n;
  hipLaunchKernelGGL(k_identity, dim3((unsigned)((nn + 255) / 256), batch), dim3(256), 0, stream, n, X);
  WB_CHECK_LAUNCH();
  for (int i0 = ((n - 1) / TB) * TB; i0 >= 0; i0 -= TB) {
    const int i1 = (i0 + TB < n) ? i0 + TB : n;
    if (i1 < n) {
      hipError_t e = wb_zgemm_ops_launch(0, 0, i1 - i0, n - i1, n - i1, mkc(1.0, 0.0),
                                         bp_strided(T + (long long)i0 * n + i1, nn), n,
                                         bp_strided(X + (long long)i1 * n + i1, nn), n, mkc(0.0, 0.0),
                                         bp_strided(S, nn), n, batch, stream);
      if (e != hipSuccess) return e;
    }
    const int cols = n - i0;
    hipLaunchKernelGGL(k_trevc_block, dim3((cols + 255) / 256, batch), dim3(256), 0, stream, n, i0, T,
                       X, S, nn);
    WB_CHECK_LAUNCH();
  }
  return hipSuccess;
}

// A (batch of n x n) is destroyed. Outputs: w (batch x n) eigenvalues, W (batch x n x n)
// eigenvectors in columns. Workspace: Z (batch x n x n), X (batch x n x n), tau (batch x n).
hipError_t wb_zgeev_launch(int n, cd* A, cd* w, cd* W, cd* Z, cd* X, cd* tau, void* rot_ws, int* info,
                           int batch, hipStream_t stream) {
  if (n <= 0 || batch <= 0) return hipSuccess;
  const size_t sh = sizeof(cd) * (size_t)n;
  // Algorithmic flops per stage, from EIG(n) = 100 n^3 for the whole complex eigensolve:
  // Hessenberg 40/3 n^3, forming Q 16/3 n^3, triangular eigenvectors 4/3 n^3, back-transform
  // GEMM 8 n^3 (counted by the GEMM itself), QR iteration with Schur vectors the remaining 72 n^3.
  const double n3 = (double)n * n * n;
  hipError_t z = wb_zero_int_launch(info, batch, stream);
  if (z != hipSuccess) return z;
  const bool blocked = n >= WB_BLOCKED_HESSENBERG_MIN;
  if (blocked) {
    // Blocked: V (dense reflectors) in W, Y / T / GEMM scratch in X, Q^H formed in Z.
    hipError_t e = wb_gehrd_blocked(n, A, W, X, batch, stream);
    if (e != hipSuccess) return e;
    e = wb_unghr_blocked(n, W, X, Z, batch, stream);
    if (e != hipSuccess) return e;
  } else {
    WB_TIMED_F(WB_T_GEHRD, stream, 40.0 / 3.0 * n3 * batch, hipLaunchKernelGGL(k_gehrd, dim3(batch), dim3(NTH), sh, stream, n, A, tau));
    WB_CHECK_LAUNCH();
    WB_TIMED_F(WB_T_UNGHR, stream, 16.0 / 3.0 * n3 * batch,
             hipLaunchKernelGGL(k_unghr, dim3(batch), dim3(NTH), sizeof(cd) * (size_t)n, stream, n, A, tau, Z));
    WB_CHECK_LAUNCH();
  }
  {
    hipError_t e2 = wb_hqr_ms_launch(n, A, Z, blocked ? 1 : 0, w, info, rot_ws, batch, stream);
    if (e2 != hipSuccess) return e2;
  }
  if (blocked) {
    hipError_t e1 = wb_trevc_blocked(n, A, X, W, batch, stream);
    if (e1 != hipSuccess) return e1;
  } else {
    WB_TIMED_F(WB_T_TREVC, stream, 4.0 / 3.0 * n3 * batch, hipLaunchKernelGGL(k_trevc, dim3(batch), dim3(NT), 0, stream, n, A, X));
    WB_CHECK_LAUNCH();
  }
  // W = Z X (the QR keeps the Schur vectors as Z^H on the blocked path)
  hipError_t e = wb_zgemm_ops_launch(blocked ? 1 : 0, 0, n, n, n, mkc(1.0, 0.0),
                                     bp_strided(Z, (long long)n * n), n, bp_strided(X, (long long)n * n),
                                     n, mkc(0.0, 0.0), bp_strided(W, (long long)n * n), n, batch, stream);
  if (e != hipSuccess) return e;
  WB_TIMED(WB_T_COLNORM, stream, hipLaunchKernelGGL(k_colnorm, dim3(batch), dim3(NT), 0, stream, n, W));
  return hipGetLastError();
}

// Schur form only: T in A (upper triangular), the Schur vectors in Z (conjugate-transposed, Z^H,
// when *zh is set on return), eigenvalues in w. W and X are n x n scratch per matrix.
hipError_t wb_schur_launch(int n, cd* A, cd* w, cd* W, cd* Z, cd* X, cd* tau, void* rot_ws, int* info,
                           int batch, int* zh, hipStream_t stream) {
  *zh = 0;
  if (n <= 0 || batch <= 0) return hipSuccess;
  hipError_t z = wb_zero_int_launch(info, batch, stream);
  if (z != hipSuccess) return z;
  const bool blocked = n >= WB_BLOCKED_HESSENBERG_MIN;
  if (blocked) {
    hipError_t e = wb_gehrd_blocked(n, A, W, X, batch, stream);
    if (e != hipSuccess) return e;
    e = wb_unghr_blocked(n, W, X, Z, batch, stream);
    if (e != hipSuccess) return e;
  } else {
    hipLaunchKernelGGL(k_gehrd, dim3(batch), dim3(NTH), sizeof(cd) * (size_t)n, stream, n, A, tau);
    WB_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_unghr, dim3(batch), dim3(NTH), sizeof(cd) * (size_t)n, stream, n, A, tau, Z);
    WB_CHECK_LAUNCH();
  }
  *zh = blocked ? 1 : 0;
  return wb_hqr_ms_launch(n, A, Z, blocked ? 1 : 0, w, info, rot_ws, batch, stream);
}
