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
  WB_TIMED_F(WB_T_GEHRD, stream, 40.0 / 3.0 * n3 * batch, hipLaunchKernelGGL(k_gehrd, dim3(batch), dim3(NTH), sh, stream, n, A, tau));
  WB_CHECK_LAUNCH();
  WB_TIMED_F(WB_T_UNGHR, stream, 16.0 / 3.0 * n3 * batch,
           hipLaunchKernelGGL(k_unghr, dim3(batch), dim3(NTH), sizeof(cd) * (size_t)n, stream, n, A, tau, Z));
  WB_CHECK_LAUNCH();
  {
    hipError_t e2 = wb_hqr_ms_launch(n, A, Z, w, info, rot_ws, batch, stream);
    if (e2 != hipSuccess) return e2;
  }
  WB_TIMED_F(WB_T_TREVC, stream, 4.0 / 3.0 * n3 * batch, hipLaunchKernelGGL(k_trevc, dim3(batch), dim3(NT), 0, stream, n, A, X));
  WB_CHECK_LAUNCH();
  hipError_t e = wb_zgemm_launch(n, n, n, mkc(1.0, 0.0), bp_strided(Z, (long long)n * n), n,
                                 bp_strided(X, (long long)n * n), n, mkc(0.0, 0.0),
                                 bp_strided(W, (long long)n * n), n, batch, stream);
  if (e != hipSuccess) return e;
  WB_TIMED(WB_T_COLNORM, stream, hipLaunchKernelGGL(k_colnorm, dim3(batch), dim3(NT), 0, stream, n, W));
  return hipGetLastError();
}
