// Batched complex FP64 LU with partial pivoting and the triangular solves built on it.
//
// Right-looking blocked factorization in steps of WB_NP panels of WB_NB columns:
//   k_getrf_panel   one workgroup per matrix factors a tall panel (pivot search, row swap,
//                   scale, rank-1 update inside the panel);
//   k_laswp_trsm    applies a panel's row swaps to every column outside it and/or solves
//                   L11 U12 = A12 for the columns to its right (64 columns per workgroup);
//   zgemm           the step's later panels and the panel rows of the trailing columns kept
//                   up to date inside the step, then A22 -= L21 U12 with k = WB_NP WB_NB on
//                   the matrix cores.
// Solves (getrs) reuse the same split: row permutation, then per diagonal block a small
// triangular kernel followed by a GEMM update of the rows still to be solved. The pivots are
// chosen by |re| + |im| with the first maximum winning, as LAPACK's izamax does.
#include "wb_internal.h"
#include "wb_timing.h"

#define WB_NB 32
#ifndef WB_NP
#define WB_NP 4    // panels per factorization step (the trailing update has k = WB_NP WB_NB)
#endif

namespace {

__global__ __launch_bounds__(256) void k_getrf_panel(int n, int k0, int jb, BatchPtr A, int lda,
                                                     int* ipiv, int* info) {
  const long long b = blockIdx.x;
  cd* a = A.ptr(b);
  int* piv = ipiv + b * (long long)n;
  const int t = threadIdx.x;
  __shared__ double s_val[256];
  __shared__ int s_idx[256];
  __shared__ cd s_prow[WB_NB];
  for (int j = k0; j < k0 + jb; ++j) {
    // pivot search over rows j..n-1 of column j
    double best = -1.0;
    int bi = n;
    for (int i = j + t; i < n; i += 256) {
      double v = cabs1(a[(long long)i * lda + j]);
      if (v != v) v = INFINITY;            // a NaN entry is a candidate like any other
      if (v > best) { best = v; bi = i; }
    }
    s_val[t] = best;
    s_idx[t] = bi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (t < s) {
        double v2 = s_val[t + s];
        int i2 = s_idx[t + s];
        if (v2 > s_val[t] || (v2 == s_val[t] && i2 < s_idx[t])) { s_val[t] = v2; s_idx[t] = i2; }
      }
      __syncthreads();
    }
    // p lies in [j, n): every thread with a row in range offers it, and NaN counts as the largest
    const int p = s_idx[0];
    const double pv = s_val[0];
    __syncthreads();
    if (t == 0) {
      piv[j] = p;
      if (pv == 0.0 && info[b] == 0) info[b] = j + 1;
    }
    // swap rows j and p inside the panel, and stage the pivot row
    for (int c = t; c < jb; c += 256) {
      const int col = k0 + c;
      cd x = a[(long long)j * lda + col];
      cd y = a[(long long)p * lda + col];
      if (p != j) {
        a[(long long)j * lda + col] = y;
        a[(long long)p * lda + col] = x;
        s_prow[c] = y;
      } else {
        s_prow[c] = x;
      }
    }
    __syncthreads();
    const cd d = s_prow[j - k0];
    const cd rd = (pv == 0.0) ? mkc(0.0, 0.0) : crecip(d);
    for (int i = j + 1 + t; i < n; i += 256) {
      cd* row = a + (long long)i * lda;
      cd l = (pv == 0.0) ? row[j] : cmul(row[j], rd);
      row[j] = l;
      for (int c = j + 1 - k0; c < jb; ++c) row[k0 + c] = csub(row[k0 + c], cmul(l, s_prow[c]));
    }
    __syncthreads();
  }
}

// The same panel factorization with the panel held in registers: one thread per row (n <= PNT),
// its WB_NB panel entries in registers. Rows are never moved inside the panel: every thread
// keeps the logical position of the row it holds (pos), a row swap only exchanges two
// positions, and the rows are stored to their final positions at the end. A column step is a
// wave-level argmax over the candidate positions (first maximum wins), one barrier to combine
// the waves, one to broadcast the pivot row through LDS (double buffered), and a rank-1 update
// of each thread's own row, with no global traffic inside the panel.
constexpr int PNT = 768;
struct PanelCtx {
  int n, k0, jb, lane, wv, t;
  bool own;
  int* piv;
  int* info;
  long long b;
  double* s_v;
  int* s_i;
  cd (*s_prow)[WB_NB];
};
// 1 / b with one real division: b scaled by a power of two so |b|^2 neither overflows nor
// underflows, conj(b') / |b'|^2, scaled back.
__device__ inline cd crecip_scaled(cd b) {
  const double m = fmax(fabs(b.x), fabs(b.y));
  const int e = ilogb(m);
  const double bx = ldexp(b.x, -e), by = ldexp(b.y, -e);
  const double inv = 1.0 / (bx * bx + by * by);
  return mkc(ldexp(bx * inv, -e), ldexp(-by * inv, -e));
}
// Column C of the register-resident panel (compile-time C keeps x[] in registers).
template <int C, int PNW>
__device__ __forceinline__ void panel_col(cd (&x)[WB_NB], int& pos, const PanelCtx& P) {
  if (C < P.jb) {
    const int j = P.k0 + C;
    const bool cand = P.own && pos >= j;
    double v = cand ? cabs1(x[C]) : -1.0;
    if (v != v) v = INFINITY;              // NaN is the largest candidate, so the pivot stays in range
    int vi = cand ? pos : P.n;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double v2 = __shfl_xor(v, off);
      const int i2 = __shfl_xor(vi, off);
      if (v2 > v || (v2 == v && i2 < vi)) { v = v2; vi = i2; }
    }
    if (P.lane == 0) { P.s_v[P.wv] = v; P.s_i[P.wv] = vi; }
    __syncthreads();
    double pv = P.s_v[0];
    int p = P.s_i[0];
#pragma unroll
    for (int w = 1; w < PNW; ++w) {
      const double v2 = P.s_v[w];
      const int i2 = P.s_i[w];
      if (v2 > pv || (v2 == pv && i2 < p)) { pv = v2; p = i2; }
    }
    constexpr int buf = C & 1;
    const bool is_pivot = P.own && pos == p;
    if (is_pivot) {
#pragma unroll
      for (int cc = 0; cc < WB_NB; ++cc) P.s_prow[buf][cc] = x[cc];
    }
    // the swap: the pivot row takes position j, the row at j takes the pivot's position
    if (P.own && pos == j) pos = p;
    if (is_pivot) pos = j;
    if (P.t == 0) {
      P.piv[j] = p;
      if (pv == 0.0 && P.info[P.b] == 0) P.info[P.b] = j + 1;
    }
    __syncthreads();
    if (P.own && pos > j) {
      const cd l = (pv == 0.0) ? x[C] : cmul(x[C], crecip_scaled(P.s_prow[buf][C]));
      x[C] = l;
#pragma unroll
      for (int cc = C + 1; cc < WB_NB; ++cc) x[cc] = csub(x[cc], cmul(l, P.s_prow[buf][cc]));
    }
  }
  if constexpr (C + 1 < WB_NB) panel_col<C + 1, PNW>(x, pos, P);
}

// NTP threads (one per row): 768, or one wave for n <= 64, where the barriers cost nothing.
template <int NTP>
__global__ __launch_bounds__(NTP) void k_getrf_panel_reg(int n, int k0, int jb, BatchPtr A, int lda,
                                                         int* ipiv, int* info) {
  constexpr int PNW = NTP / 64;
  __shared__ double s_v[PNW];
  __shared__ int s_i[PNW];
  __shared__ cd s_prow[2][WB_NB];
  PanelCtx P;
  P.b = blockIdx.x;
  cd* a = A.ptr(P.b);
  P.n = n; P.k0 = k0; P.jb = jb;
  P.t = threadIdx.x; P.lane = P.t & 63; P.wv = P.t >> 6;
  const int r = P.t;
  P.own = r >= k0 && r < n;
  P.piv = ipiv + P.b * (long long)n;
  P.info = info;
  P.s_v = s_v; P.s_i = s_i; P.s_prow = s_prow;
  cd x[WB_NB];
#pragma unroll
  for (int c = 0; c < WB_NB; ++c) x[c] = (P.own && c < jb) ? a[(long long)r * lda + k0 + c] : mkc(0.0, 0.0);
  int pos = r;
  panel_col<0, PNW>(x, pos, P);
  if (P.own) {
#pragma unroll
    for (int c = 0; c < WB_NB; ++c)
      if (c < jb) a[(long long)pos * lda + k0 + c] = x[c];
  }
}

// Swaps for panel rows k0..k0+jb-1 applied to every column outside the panel; then the unit
// lower triangular solve for the columns right of the panel. 64 columns per workgroup.
// mode 0: both; 1: the swaps only; 2: the triangular solve only (the grid then covers just the
// columns right of the panel).
__global__ __launch_bounds__(64) void k_laswp_trsm(int n, int k0, int jb, BatchPtr A, int lda,
                                                   const int* ipiv, int mode) {
  const long long b = blockIdx.y;
  cd* a = A.ptr(b);
  const int* piv = ipiv + b * (long long)n;
  const int t = threadIdx.x;
  __shared__ cd L[WB_NB][WB_NB + 1];
  const int ci = blockIdx.x * 64 + t;
  int col = -1;
  if (mode == 2) {
    if (ci < n - k0 - jb) col = k0 + jb + ci;
  } else if (ci < n - jb) {
    col = (ci < k0) ? ci : ci + jb;
  }
  if (mode == 1) {
    if (col >= 0) {
      for (int j = k0; j < k0 + jb; ++j) {
        const int p = piv[j];
        if (p != j) {
          cd x = a[(long long)j * lda + col];
          a[(long long)j * lda + col] = a[(long long)p * lda + col];
          a[(long long)p * lda + col] = x;
        }
      }
    }
    return;
  }
  if (col >= 0 && mode == 0) {
    for (int j = k0; j < k0 + jb; ++j) {
      const int p = piv[j];
      if (p != j) {
        cd x = a[(long long)j * lda + col];
        a[(long long)j * lda + col] = a[(long long)p * lda + col];
        a[(long long)p * lda + col] = x;
      }
    }
  }
  // stage L11 (after the swaps of the panel itself, already applied by the panel kernel)
  for (int e = t; e < jb * jb; e += 64) {
    const int r = e / jb, c = e % jb;
    L[r][c] = a[(long long)(k0 + r) * lda + k0 + c];
  }
  __syncthreads();
  if (col >= k0 + jb) {
    cd x[WB_NB];
#pragma unroll
    for (int r = 0; r < WB_NB; ++r)
      if (r < jb) x[r] = a[(long long)(k0 + r) * lda + col];
#pragma unroll
    for (int r = 1; r < WB_NB; ++r) {
      if (r < jb) {
        cd s = x[r];
#pragma unroll
        for (int q = 0; q < WB_NB; ++q)
          if (q < r) s = csub(s, cmul(L[r][q], x[q]));
        x[r] = s;
      }
    }
#pragma unroll
    for (int r = 0; r < WB_NB; ++r)
      if (r < jb) a[(long long)(k0 + r) * lda + col] = x[r];
  }
}

// Row permutation of B by the LU pivots (sequential swaps, as LAPACK's zlaswp); reverse:
// the swaps in the opposite order (the inverse permutation).
__global__ __launch_bounds__(64) void k_laswp_rhs(int n, int nrhs, BatchPtr B, int ldb,
                                                  const int* ipiv, int piv_stride, int reverse) {
  const long long b = blockIdx.y;
  cd* x = B.ptr(b);
  const int* piv = ipiv + b * (long long)piv_stride;
  const int col = blockIdx.x * 64 + threadIdx.x;
  if (col >= nrhs) return;
  for (int i = 0; i < n; ++i) {
    const int j = reverse ? n - 1 - i : i;
    const int p = piv[j];
    if (p != j) {
      cd v = x[(long long)j * ldb + col];
      x[(long long)j * ldb + col] = x[(long long)p * ldb + col];
      x[(long long)p * ldb + col] = v;
    }
  }
}

// Solve the jb x jb diagonal block (rows r0..r0+jb-1) of a triangular factor against the
// same rows of B, in place. KIND 0: unit lower L; 1: non-unit upper U; 2: non-unit lower U^H;
// 3: unit upper L^H (kinds 2 and 3 read the packed LU factors transposed and conjugated).
template <int KIND>
__global__ __launch_bounds__(64) void k_trsm_diag(int r0, int jb, int nrhs, BatchPtr A, int lda,
                                                  BatchPtr B, int ldb) {
  const long long b = blockIdx.y;
  const cd* a = A.ptr(b);
  cd* x = B.ptr(b);
  const int t = threadIdx.x;
  __shared__ cd T[WB_NB][WB_NB + 1];
  for (int e = t; e < jb * jb; e += 64) {
    const int r = e / jb, c = e % jb;
    if (KIND < 2) T[r][c] = a[(long long)(r0 + r) * lda + r0 + c];
    else T[r][c] = cconj(a[(long long)(r0 + c) * lda + r0 + r]);
  }
  __syncthreads();
  const int col = blockIdx.x * 64 + t;
  if (col >= nrhs) return;
  cd v[WB_NB];
#pragma unroll
  for (int r = 0; r < WB_NB; ++r)
    if (r < jb) v[r] = x[(long long)(r0 + r) * ldb + col];
  if (KIND == 0 || KIND == 2) {
#pragma unroll
    for (int r = 0; r < WB_NB; ++r) {
      if (r < jb) {
        cd s = v[r];
#pragma unroll
        for (int q = 0; q < WB_NB; ++q)
          if (q < r) s = csub(s, cmul(T[r][q], v[q]));
        v[r] = (KIND == 2) ? cdiv(s, T[r][r]) : s;
      }
    }
  } else {
#pragma unroll
    for (int r = WB_NB - 1; r >= 0; --r) {
      if (r < jb) {
        cd s = v[r];
#pragma unroll
        for (int q = 0; q < WB_NB; ++q)
          if (q > r && q < jb) s = csub(s, cmul(T[r][q], v[q]));
        v[r] = (KIND == 1) ? cdiv(s, T[r][r]) : s;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < WB_NB; ++r)
    if (r < jb) x[(long long)(r0 + r) * ldb + col] = v[r];
}

__global__ void k_zero_int(int* x, int count) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < count) x[e] = 0;
}

__global__ void k_set_identity(int n, BatchPtr B, int ldb) {
  const long long b = blockIdx.y;
  cd* x = B.ptr(b);
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n * n) return;
  const int r = (int)(e / n), c = (int)(e % n);
  x[(long long)r * ldb + c] = mkc(r == c ? 1.0 : 0.0, 0.0);
}

inline BatchPtr offset(BatchPtr p, long long off) { p.base += off; return p; }

// ----------------------------------------------------------------------------- few-column solves
// A X = B (CONJ = false) or A^H X = B (CONJ = true) from the packed LU factors for at most
// FEW_NR right-hand sides, the whole solve in one 512-thread workgroup per matrix: B and the
// pivots staged in LDS (dynamic: n x nrhs complex + n ints), the row swaps applied there by one
// lane per column, then both triangular sweeps in 32-row blocks -- the block's diagonal
// triangle (staged in LDS) solved by wave 0 with one row per lane and the solved values
// broadcast by readlane, the rows beyond the block updated by all threads -- and the result
// written back once. Replaces 4 n / 32 small launches and a swap loop over global memory.
constexpr int FEW_NR = 4;
constexpr int FEW_B = 32;
constexpr int FEW_NT = 512;

__device__ inline cd few_readlane(cd x, int src) {
  const unsigned long long rb = __double_as_longlong(x.x), ib = __double_as_longlong(x.y);
  const int r0 = __builtin_amdgcn_readlane((int)(rb & 0xffffffffull), src);
  const int r1 = __builtin_amdgcn_readlane((int)(rb >> 32), src);
  const int i0 = __builtin_amdgcn_readlane((int)(ib & 0xffffffffull), src);
  const int i1 = __builtin_amdgcn_readlane((int)(ib >> 32), src);
  return mkc(__longlong_as_double(((unsigned long long)(unsigned)r1 << 32) | (unsigned)r0),
             __longlong_as_double(((unsigned long long)(unsigned)i1 << 32) | (unsigned)i0));
}

template <bool CONJ>
__global__ __launch_bounds__(FEW_NT) void k_getrs_few(int n, int nrhs, BatchPtr LU, int lda, const int* ipiv,
                                                     BatchPtr B, int ldb) {
  extern __shared__ unsigned char few_smem[];
  cd* __restrict__ sx = (cd*)few_smem;                               // n x nrhs
  int* __restrict__ sp = (int*)(few_smem + sizeof(cd) * (size_t)n * nrhs);   // n pivots
  __shared__ cd T[2][FEW_B][FEW_B + 1];                              // diagonal blocks, double buffered
  const long long b = blockIdx.x;
  const cd* a = LU.ptr(b);
  cd* x = B.ptr(b);
  const int* piv = ipiv + b * (long long)n;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  for (int e = t; e < n * nrhs; e += FEW_NT) {
    const int r = e / nrhs, c = e % nrhs;
    sx[e] = x[(long long)r * ldb + c];
  }
  for (int e = t; e < n; e += FEW_NT) sp[e] = piv[e];
  __syncthreads();
  auto swaps = [&](bool reverse) {
    if (t < nrhs) {
      for (int i = 0; i < n; ++i) {
        const int j = reverse ? n - 1 - i : i;
        const int q = sp[j];
        if (q != j) {
          const cd u = sx[j * nrhs + t];
          sx[j * nrhs + t] = sx[q * nrhs + t];
          sx[q * nrhs + t] = u;
        }
      }
    }
    __syncthreads();
  };
  // factor entries: the lower sweep uses L (!CONJ, unit) or U^H (CONJ, non-unit), the upper
  // sweep U (!CONJ, non-unit) or L^H (CONJ, unit); (r, c) in the swept triangle's coordinates
  auto fac = [&](int r, int c) -> cd {
    return CONJ ? cconj(a[(long long)c * lda + r]) : a[(long long)r * lda + c];
  };
  constexpr bool LOWER_UNIT = !CONJ, UPPER_UNIT = CONJ;
  auto block_len = [&](int blk) { return (n - blk < FEW_B) ? n - blk : FEW_B; };
  // the block is stored transposed (T[c][r] = entry (r, c)): in the solve, lane r reads entry
  // (r, j) of column j, consecutive lanes on consecutive LDS addresses
  auto load_block = [&](int buf, int blk) {
    const int jb = block_len(blk);
    for (int e = t; e < jb * jb; e += FEW_NT) {
      const int r = e / jb, c = e % jb;
      T[buf][c][r] = fac(blk + r, blk + c);
    }
  };
  // row i -= (factor row i, columns blk..blk+jb-1) x (solved rows blk..blk+jb-1): the block's
  // factor entries are loaded together (a fixed-length unrolled loop), then accumulated
  auto update_row = [&](int i, int blk, int jb) {
    cd m[FEW_B];
#pragma unroll
    for (int k = 0; k < FEW_B; ++k) m[k] = (k < jb) ? fac(i, blk + k) : mkc(0.0, 0.0);
    cd acc[FEW_NR];
#pragma unroll
    for (int c = 0; c < FEW_NR; ++c) acc[c] = (c < nrhs) ? sx[i * nrhs + c] : mkc(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < FEW_B; ++k) {
      if (k < jb) {
#pragma unroll
        for (int c = 0; c < FEW_NR; ++c)
          if (c < nrhs) acc[c] = csub(acc[c], cmul(m[k], sx[(blk + k) * nrhs + c]));
      }
    }
#pragma unroll
    for (int c = 0; c < FEW_NR; ++c)
      if (c < nrhs) sx[i * nrhs + c] = acc[c];
  };
  // wave 0: the block's triangle, one row per lane (lower: rows in increasing order)
  auto solve_block = [&](int buf, int blk, bool lower, bool unit) {
    const int jb = block_len(blk);
    cd xi[FEW_NR];
#pragma unroll
    for (int c = 0; c < FEW_NR; ++c) xi[c] = (lane < jb && c < nrhs) ? sx[(blk + lane) * nrhs + c] : mkc(0.0, 0.0);
    // non-unit diagonal: every lane forms its own row's reciprocal before the serial loop
    const cd rd = (!unit && lane < jb) ? crecip(T[buf][lane][lane]) : mkc(1.0, 0.0);
    for (int s = 0; s < jb; ++s) {
      const int j = lower ? s : jb - 1 - s;
      if (!unit && lane == j) {
#pragma unroll
        for (int c = 0; c < FEW_NR; ++c) xi[c] = cmul(xi[c], rd);
      }
      const bool below = lower ? (lane > j && lane < jb) : (lane < j);
      const cd tij = below ? T[buf][j][lane] : mkc(0.0, 0.0);
#pragma unroll
      for (int c = 0; c < FEW_NR; ++c) {
        if (c < nrhs) {
          const cd xj = few_readlane(xi[c], j);
          if (below) xi[c] = csub(xi[c], cmul(tij, xj));
        }
      }
    }
#pragma unroll
    for (int c = 0; c < FEW_NR; ++c)
      if (lane < jb && c < nrhs) sx[(blk + lane) * nrhs + c] = xi[c];
  };
  if (!CONJ) swaps(false);
  // lower sweep: per block, wave 0 solves the triangle, then every thread updates the rows
  // below it while the next block's triangle is fetched into the other buffer
  const int nb = (n + FEW_B - 1) / FEW_B;
  load_block(0, 0);
  __syncthreads();
  for (int q = 0; q < nb; ++q) {
    const int blk = q * FEW_B, jb = block_len(blk), buf = q & 1;
    if (wv == 0) solve_block(buf, blk, true, LOWER_UNIT);
    __syncthreads();
    if (q + 1 < nb) load_block(buf ^ 1, blk + FEW_B);
    for (int i = blk + jb + t; i < n; i += FEW_NT) update_row(i, blk, jb);
    __syncthreads();
  }
  // upper sweep, blocks from the last
  load_block(0, (nb - 1) * FEW_B);
  __syncthreads();
  for (int q = nb - 1, it = 0; q >= 0; --q, ++it) {
    const int blk = q * FEW_B, jb = block_len(blk), buf = it & 1;
    if (wv == 0) solve_block(buf, blk, false, UPPER_UNIT);
    __syncthreads();
    if (q > 0) load_block(buf ^ 1, blk - FEW_B);
    for (int i = t; i < blk; i += FEW_NT) update_row(i, blk, jb);
    __syncthreads();
  }
  if (CONJ) swaps(true);
  for (int e = t; e < n * nrhs; e += FEW_NT) {
    const int r = e / nrhs, c = e % nrhs;
    x[(long long)r * ldb + c] = sx[e];
  }
}

// dynamic LDS of k_getrs_few, or 0 when the solve does not fit it
inline size_t getrs_few_lds(int n, int nrhs) {
  if (nrhs < 1 || nrhs > FEW_NR) return 0;
  const size_t bytes = sizeof(cd) * (size_t)n * nrhs + sizeof(int) * (size_t)n;
  return bytes <= 65536 ? bytes : 0;
}

}  // namespace

hipError_t wb_getrf_launch(int n, BatchPtr A, int lda, int* ipiv, int* info, int batch,
                           hipStream_t stream) {
  if (n <= 0 || batch <= 0) return hipSuccess;
  // a fill kernel rather than a memset node, so the sequence stays safe under graph capture
  hipLaunchKernelGGL(k_zero_int, dim3((batch + 255) / 256), dim3(256), 0, stream, info, batch);
  WB_CHECK_LAUNCH();
  // Right-looking, WB_NP panels of WB_NB columns per step: inside the step each panel is
  // factored, its swaps reach every column outside it (as in LAPACK), its rows of the trailing
  // columns are brought up to date with the step's earlier panels before its U rows are
  // solved, and the step's later panels get its update; the trailing matrix then gets one
  // k = WB_NP WB_NB update, so it is read and written once per step.
  const cd one = mkc(1.0, 0.0), mone = mkc(-1.0, 0.0);
  auto panel = [&](int k0, int jb) -> hipError_t {
    if (n <= 64) {
      WB_TIMED(WB_T_GETRF_PANEL, stream,
               hipLaunchKernelGGL(k_getrf_panel_reg<64>, dim3(batch), dim3(64), 0, stream, n, k0, jb, A, lda,
                                  ipiv, info));
    } else if (n <= PNT) {
      WB_TIMED(WB_T_GETRF_PANEL, stream,
               hipLaunchKernelGGL(k_getrf_panel_reg<PNT>, dim3(batch), dim3(PNT), 0, stream, n, k0, jb, A,
                                  lda, ipiv, info));
    } else {
      WB_TIMED(WB_T_GETRF_PANEL, stream,
               hipLaunchKernelGGL(k_getrf_panel, dim3(batch), dim3(256), 0, stream, n, k0, jb, A, lda,
                                  ipiv, info));
    }
    return hipGetLastError();
  };
  auto swap_trsm = [&](int k0, int jb, int mode) -> hipError_t {
    const int cols = (mode == 2) ? n - k0 - jb : n - jb;
    if (cols <= 0) return hipSuccess;
    WB_TIMED(WB_T_LASWP_TRSM, stream,
             hipLaunchKernelGGL(k_laswp_trsm, dim3((cols + 63) / 64, batch), dim3(64), 0, stream, n, k0, jb, A,
                                lda, ipiv, mode));
    return hipGetLastError();
  };
  // C[r0.., c0..] (m x nc) -= A[r0.., a0..] (m x kk) * A[a0.., c0..] (kk x nc)
  auto update = [&](int r0, int c0, int m, int nc, int a0, int kk) -> hipError_t {
    if (m <= 0 || nc <= 0 || kk <= 0) return hipSuccess;
    return wb_zgemm_launch(m, nc, kk, mone, offset(A, (long long)r0 * lda + a0), lda,
                           offset(A, (long long)a0 * lda + c0), lda, one, offset(A, (long long)r0 * lda + c0),
                           lda, batch, stream);
  };
  hipError_t e;
  constexpr int SB = WB_NP * WB_NB;                 // columns per step
  for (int k0 = 0; k0 < n; k0 += SB) {
    const int jb = (n - k0 < SB) ? n - k0 : SB;
    const int kend = k0 + jb;
    for (int cq = k0; cq < kend; cq += WB_NB) {
      const int wq = (kend - cq < WB_NB) ? kend - cq : WB_NB;
      const int cr = cq + wq;                       // first column right of this panel
      if ((e = panel(cq, wq))) return e;            // its columns are up to date (below)
      if ((e = swap_trsm(cq, wq, 1))) return e;     // its swaps, every other column
      // its rows of the trailing columns: the earlier panels of this step, then its U rows
      if ((e = update(cq, kend, wq, n - kend, k0, cq - k0))) return e;
      if ((e = swap_trsm(cq, wq, 2))) return e;
      // the later panels of this step: every row below this panel (right-looking in the step)
      if ((e = update(cr, cr, n - cr, kend - cr, cq, wq))) return e;
    }
    if ((e = update(kend, kend, n - kend, n - kend, k0, jb))) return e;   // trailing, k = jb
  }
  return hipSuccess;
}

// Solve A X = B in place given the LU factors of A.
hipError_t wb_getrs_launch(int n, int nrhs, BatchPtr LU, int lda, const int* ipiv, BatchPtr B,
                           int ldb, int batch, hipStream_t stream) {
  if (n <= 0 || nrhs <= 0 || batch <= 0) return hipSuccess;
  if (const size_t sh = getrs_few_lds(n, nrhs)) {
    WB_TIMED(WB_T_TRSM_DIAG, stream,
             hipLaunchKernelGGL(k_getrs_few<false>, dim3(batch), dim3(FEW_NT), sh, stream, n, nrhs, LU, lda, ipiv,
                                B, ldb));
    return hipGetLastError();
  }
  dim3 g((nrhs + 63) / 64, batch);
  hipLaunchKernelGGL(k_laswp_rhs, g, dim3(64), 0, stream, n, nrhs, B, ldb, ipiv, n, 0);
  WB_CHECK_LAUNCH();
  for (int r0 = 0; r0 < n; r0 += WB_NB) {
    const int jb = (n - r0 < WB_NB) ? n - r0 : WB_NB;
    WB_TIMED(WB_T_TRSM_DIAG, stream,
             hipLaunchKernelGGL(k_trsm_diag<0>, g, dim3(64), 0, stream, r0, jb, nrhs, LU, lda, B, ldb));
    WB_CHECK_LAUNCH();
    const int rest = n - r0 - jb;
    if (rest > 0) {
      hipError_t e = wb_zgemm_launch(rest, nrhs, jb, mkc(-1.0, 0.0),
                                     offset(LU, (long long)(r0 + jb) * lda + r0), lda,
                                     offset(B, (long long)r0 * ldb), ldb, mkc(1.0, 0.0),
                                     offset(B, (long long)(r0 + jb) * ldb), ldb, batch, stream);
      if (e != hipSuccess) return e;
    }
  }
  const int last = ((n - 1) / WB_NB) * WB_NB;
  for (int r0 = last; r0 >= 0; r0 -= WB_NB) {
    const int jb = (n - r0 < WB_NB) ? n - r0 : WB_NB;
    WB_TIMED(WB_T_TRSM_DIAG, stream,
             hipLaunchKernelGGL(k_trsm_diag<1>, g, dim3(64), 0, stream, r0, jb, nrhs, LU, lda, B, ldb));
    WB_CHECK_LAUNCH();
    if (r0 > 0) {
      hipError_t e = wb_zgemm_launch(r0, nrhs, jb, mkc(-1.0, 0.0), offset(LU, (long long)r0), lda,
                                     offset(B, (long long)r0 * ldb), ldb, mkc(1.0, 0.0), B, ldb,
                                     batch, stream);
      if (e != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

// Solve A^H X = B in place given the LU factors of A (A = P^T L U, so A^H = U^H L^H P):
// forward substitution with U^H, backward with L^H, then the swaps in reverse order. The
// off-diagonal updates are conjugate-transposed GEMMs reading the same packed factors.
hipError_t wb_getrs_conj_launch(int n, int nrhs, BatchPtr LU, int lda, const int* ipiv, BatchPtr B,
                                int ldb, int batch, hipStream_t stream) {
  if (n <= 0 || nrhs <= 0 || batch <= 0) return hipSuccess;
  if (const size_t sh = getrs_few_lds(n, nrhs)) {
    WB_TIMED(WB_T_TRSM_DIAG, stream,
             hipLaunchKernelGGL(k_getrs_few<true>, dim3(batch), dim3(FEW_NT), sh, stream, n, nrhs, LU, lda, ipiv,
                                B, ldb));
    return hipGetLastError();
  }
  dim3 g((nrhs + 63) / 64, batch);
  for (int r0 = 0; r0 < n; r0 += WB_NB) {
    const int jb = (n - r0 < WB_NB) ? n - r0 : WB_NB;
    WB_TIMED(WB_T_TRSM_DIAG, stream,
             hipLaunchKernelGGL(k_trsm_diag<2>, g, dim3(64), 0, stream, r0, jb, nrhs, LU, lda, B, ldb));
    WB_CHECK_LAUNCH();
    const int rest = n - r0 - jb;
    if (rest > 0) {
      // B[r0+jb:] -= (U[r0:r0+jb, r0+jb:])^H B[r0:r0+jb]
      hipError_t e = wb_zgemm_ops_launch(1, 0, rest, nrhs, jb, mkc(-1.0, 0.0),
                                         offset(LU, (long long)r0 * lda + r0 + jb), lda,
                                         offset(B, (long long)r0 * ldb), ldb, mkc(1.0, 0.0),
                                         offset(B, (long long)(r0 + jb) * ldb), ldb, batch, stream);
      if (e != hipSuccess) return e;
    }
  }
  const int last = ((n - 1) / WB_NB) * WB_NB;
  for (int r0 = last; r0 >= 0; r0 -= WB_NB) {
    const int jb = (n - r0 < WB_NB) ? n - r0 : WB_NB;
    WB_TIMED(WB_T_TRSM_DIAG, stream,
             hipLaunchKernelGGL(k_trsm_diag<3>, g, dim3(64), 0, stream, r0, jb, nrhs, LU, lda, B, ldb));
    WB_CHECK_LAUNCH();
    if (r0 > 0) {
      // B[0:r0] -= (L[r0:r0+jb, 0:r0])^H B[r0:r0+jb]
      hipError_t e = wb_zgemm_ops_launch(1, 0, r0, nrhs, jb, mkc(-1.0, 0.0),
                                         offset(LU, (long long)r0 * lda), lda,
                                         offset(B, (long long)r0 * ldb), ldb, mkc(1.0, 0.0), B, ldb,
                                         batch, stream);
      if (e != hipSuccess) return e;
    }
  }
  hipLaunchKernelGGL(k_laswp_rhs, g, dim3(64), 0, stream, n, nrhs, B, ldb, ipiv, n, 1);
  WB_CHECK_LAUNCH();
  return hipSuccess;
}

// ----------------------------------------------------------------------------- blocked solves
// Solves with many right-hand sides run in blocks of LB = 64 rows: the 64 x 64 diagonal blocks
// of L and U are inverted once (k_tri_inv, in LDS), so every step of both sweeps is a k = 64
// GEMM on the matrix cores -- the diagonal block's solve (X_b = inv(T_bb) B_b, through a
// scratch block) and the update of the rows still to be solved. A^H solves use the same
// inverses conjugate-transposed.
namespace {

constexpr int LB = 64;

// inv(unit lower L_bb) (kind 0) or inv(upper U_bb) (kind 1) of diagonal block bb, LB x LB
// row-major per (matrix, block, kind). Four lanes per column of the inverse: row i of the
// inverse is one dot product per column over the rows already found, split four ways and
// summed across the lane quad; rows go top-down (kind 0) or bottom-up (kind 1).
__global__ __launch_bounds__(4 * LB) void k_tri_inv(int n, BatchPtr LU, int lda, cd* inv, int nblk) {
  const int bb = blockIdx.x, kind = blockIdx.z;
  const long long b = blockIdx.y;
  const cd* a = LU.ptr(b);
  const int r0 = bb * LB;
  const int jb = (n - r0 < LB) ? n - r0 : LB;
  __shared__ cd T[LB][LB + 1];
  __shared__ cd X[LB][LB + 1];
  const int t = threadIdx.x;
  for (int e = t; e < LB * LB; e += 4 * LB) {
    const int r = e / LB, c = e % LB;
    T[r][c] = (r < jb && c < jb) ? a[(long long)(r0 + r) * lda + r0 + c] : mkc(0.0, 0.0);
  }
  __syncthreads();
  const int j = t >> 2, p = t & 3;   // column of the inverse, quarter of the dot product
  const bool live = j < jb;
  if (kind == 0) {
    for (int i = 0; i < LB; ++i) {
      cd acc = mkc(0.0, 0.0);
      if (live && i < jb)
        for (int l = j + p; l < i; l += 4) acc = cadd(acc, cmul(T[i][l], X[l][j]));
      acc.x += __shfl_xor(acc.x, 1); acc.y += __shfl_xor(acc.y, 1);
      acc.x += __shfl_xor(acc.x, 2); acc.y += __shfl_xor(acc.y, 2);
      if (p == 0) {
        cd x = mkc(0.0, 0.0);
        if (live && i < jb) x = (i == j) ? mkc(1.0, 0.0) : (i > j ? cneg(acc) : mkc(0.0, 0.0));
        X[i][j] = x;
      }
    }
  } else {
    for (int i = LB - 1; i >= 0; --i) {
      cd acc = mkc(0.0, 0.0);
      if (live && i < jb)
        for (int l = i + 1 + p; l <= j; l += 4) acc = cadd(acc, cmul(T[i][l], X[l][j]));
      acc.x += __shfl_xor(acc.x, 1); acc.y += __shfl_xor(acc.y, 1);
      acc.x += __shfl_xor(acc.x, 2); acc.y += __shfl_xor(acc.y, 2);
      if (p == 0) {
        cd x = mkc(0.0, 0.0);
        if (live && i < jb && i <= j) x = cdiv(csub(mkc(i == j ? 1.0 : 0.0, 0.0), acc), T[i][i]);
        X[i][j] = x;
      }
    }
  }
  __syncthreads();
  cd* out = inv + ((b * nblk + bb) * 2 + kind) * (long long)LB * LB;
  for (int e = t; e < LB * LB; e += 4 * LB) out[e] = X[e / LB][e % LB];
}


}  // namespace

size_t wb_getrs_ws_bytes(int n, int nrhs, int batch) {
  const size_t nblk = (size_t)(n + LB - 1) / LB;
  (void)nrhs;
  return sizeof(cd) * (size_t)batch * nblk * 2 * LB * LB + 256;
}

// A X = B (conj = 0) or A^H X = B (conj = 1) in place, from wb_getrf_launch's factors of A,
// with scratch ws of wb_getrs_ws_bytes(n, nrhs, batch) bytes.
hipError_t wb_getrs_blocked(int n, int nrhs, BatchPtr LU, int lda, const int* ipiv, BatchPtr B,
                            int ldb, int batch, int conj, void* ws, hipStream_t stream) {
  if (n <= 0 || nrhs <= 0 || batch <= 0) return hipSuccess;
  const int nblk = (n + LB - 1) / LB;
  cd* inv = (cd*)ws;
  const long long inv_stride = (long long)nblk * 2 * LB * LB;
  const cd one = mkc(1.0, 0.0), zero = mkc(0.0, 0.0), mone = mkc(-1.0, 0.0);
  hipError_t e;
  dim3 g((nrhs + 63) / 64, batch);
  if (!conj) {
    hipLaunchKernelGGL(k_laswp_rhs, g, dim3(64), 0, stream, n, nrhs, B, ldb, ipiv, n, 0);
    WB_CHECK_LAUNCH();
  }
  WB_TIMED(WB_T_TRSM_DIAG, stream,
           hipLaunchKernelGGL(k_tri_inv, dim3(nblk, batch, 2), dim3(4 * LB), 0, stream, n, LU, lda, inv, nblk));
  WB_CHECK_LAUNCH();
  auto INV = [&](int bb, int kind) { return bp_strided(inv + ((long long)bb * 2 + kind) * LB * LB, inv_stride); };
  // B_b := inv(T_bb) B_b in place: every GEMM tile configuration covers all jb <= 64 rows of
  // an output tile in one workgroup, which reads its whole column slice of B_b (the K loop)
  // before its epilogue writes that slice, and no other workgroup touches those columns
  static_assert(LB <= 64, "in-place diagonal-block products need one row tile per column slice");
  auto diag_solve = [&](int bb, int kind, int opa) -> hipError_t {
    const int r0 = bb * LB, jb = (n - r0 < LB) ? n - r0 : LB;
    return wb_zgemm_ops_launch(opa, 0, jb, nrhs, jb, one, INV(bb, kind), LB, offset(B, (long long)r0 * ldb), ldb,
                               zero, offset(B, (long long)r0 * ldb), ldb, batch, stream);
  };
  if (!conj) {
    for (int bb = 0; bb < nblk; ++bb) {           // L: forward
      const int r0 = bb * LB, jb = (n - r0 < LB) ? n - r0 : LB;
      if ((e = diag_solve(bb, 0, 0))) return e;
      const int rest = n - r0 - jb;
      if (rest > 0 && (e = wb_zgemm_ops_launch(0, 0, rest, nrhs, jb, mone, offset(LU, (long long)(r0 + jb) * lda + r0), lda,
                                               offset(B, (long long)r0 * ldb), ldb, one,
                                               offset(B, (long long)(r0 + jb) * ldb), ldb, batch, stream))) return e;
    }
    for (int bb = nblk - 1; bb >= 0; --bb) {      // U: backward
      const int r0 = bb * LB;
      if ((e = diag_solve(bb, 1, 0))) return e;
      if (r0 > 0) {
        const int jb = (n - r0 < LB) ? n - r0 : LB;
        if ((e = wb_zgemm_ops_launch(0, 0, r0, nrhs, jb, mone, offset(LU, (long long)r0), lda,
                                     offset(B, (long long)r0 * ldb), ldb, one, B, ldb, batch, stream))) return e;
      }
    }
  } else {
    for (int bb = 0; bb < nblk; ++bb) {           // U^H: forward
      const int r0 = bb * LB, jb = (n - r0 < LB) ? n - r0 : LB;
      if ((e = diag_solve(bb, 1, 1))) return e;
      const int rest = n - r0 - jb;
      if (rest > 0 && (e = wb_zgemm_ops_launch(1, 0, rest, nrhs, jb, mone, offset(LU, (long long)r0 * lda + r0 + jb), lda,
                                               offset(B, (long long)r0 * ldb), ldb, one,
                                               offset(B, (long long)(r0 + jb) * ldb), ldb, batch, stream))) return e;
    }
    for (int bb = nblk - 1; bb >= 0; --bb) {      // L^H: backward
      const int r0 = bb * LB;
      if ((e = diag_solve(bb, 0, 1))) return e;
      if (r0 > 0) {
        const int jb = (n - r0 < LB) ? n - r0 : LB;
        if ((e = wb_zgemm_ops_launch(1, 0, r0, nrhs, jb, mone, offset(LU, (long long)r0 * lda), lda,
                                     offset(B, (long long)r0 * ldb), ldb, one, B, ldb, batch, stream))) return e;
      }
    }
    hipLaunchKernelGGL(k_laswp_rhs, g, dim3(64), 0, stream, n, nrhs, B, ldb, ipiv, n, 1);
    WB_CHECK_LAUNCH();
  }
  return hipSuccess;
}

// inv(A) with the blocked solve: A is overwritten by its LU factors, Ainv receives the inverse.
hipError_t wb_inverse_blocked(int n, BatchPtr A, int lda, BatchPtr Ainv, int ldi, int* ipiv, int* info,
                              int batch, void* ws, hipStream_t stream) {
  hipError_t e = wb_getrf_launch(n, A, lda, ipiv, info, batch, stream);
  if (e != hipSuccess) return e;
  e = wb_set_identity_launch(n, Ainv, ldi, batch, stream);
  if (e != hipSuccess) return e;
  return wb_getrs_blocked(n, n, A, lda, ipiv, Ainv, ldi, batch, 0, ws, stream);
}

hipError_t wb_zero_int_launch(int* x, int count, hipStream_t stream) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_int, dim3((count + 255) / 256), dim3(256), 0, stream, x, count);
  return hipGetLastError();
}

hipError_t wb_set_identity_launch(int n, BatchPtr B, int ldb, int batch, hipStream_t stream) {
  long long total = (long long)n * n;
  dim3 g((unsigned)((total + 255) / 256), batch);
  hipLaunchKernelGGL(k_set_identity, g, dim3(256), 0, stream, n, B, ldb);
  return hipGetLastError();
}

// inv(A): A is overwritten by its LU factors, Ainv receives the inverse.
hipError_t wb_inverse_launch(int n, BatchPtr A, int lda, BatchPtr Ainv, int ldi, int* ipiv,
                             int* info, int batch, hipStream_t stream) {
  hipError_t e = wb_getrf_launch(n, A, lda, ipiv, info, batch, stream);
  if (e != hipSuccess) return e;
  e = wb_set_identity_launch(n, Ainv, ldi, batch, stream);
  if (e != hipSuccess) return e;
  return wb_getrs_launch(n, n, A, lda, ipiv, Ainv, ldi, batch, stream);
}
