// Batched complex FP64 GEMM on CDNA4 matrix cores.
//
// C[b] = alpha * op(A[b]) * op(B[b]) + beta * C[b], op(A) m x k, op(B) k x n, C m x n, all
// row-major; op is the identity or the conjugate transpose (then A is stored k x m, B n x k,
// and the tile loads walk the stored rows so global reads stay coalesced).
// A complex product is four real v_mfma_f64_16x16x4f64 streams (re*re, -im*im, re*im, im*re)
// into two accumulators. A 256-thread workgroup owns a 64 x 64 tile of C, its four waves a
// 32 x 32 quadrant each (2 x 2 MFMA blocks); K is staged through LDS 16 at a time, split into
// real and imaginary planes so every MFMA operand is one 8-byte LDS read, in two LDS stages: the
// next K tile is fetched into registers while the current one feeds the matrix cores.
//
// MFMA f64 16x16x4 lane layout (measured on gfx950): A[i][kk] in lane i + 16*kk, B[kk][j] in
// lane j + 16*kk, D[i][j] in lane j + 16*(i%4), register i/4.
#include "wb_common.h"
#include "wb_timing.h"

namespace {

constexpr int BM = 64, BN = 64, BK = 16, PAD = 2;
typedef double v4d __attribute__((ext_vector_type(4)));

template <bool CA, bool CB>
__global__ __launch_bounds__(256) void k_zgemm(int m, int n, int k, cd alpha, BatchPtr A, int lda,
                                               BatchPtr B, int ldb, cd beta, BatchPtr C, int ldc,
                                               int beta_zero) {
  const long long b = blockIdx.z;
  const cd* __restrict__ a = A.ptr(b);
  const cd* __restrict__ bb = B.ptr(b);
  cd* __restrict__ c = C.ptr(b);
  const int tm = blockIdx.y * BM, tn = blockIdx.x * BN;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;

  // two LDS stages: the next K tile is fetched into registers while the current one feeds the
  // matrix cores, then written to the other stage; one barrier per K tile
  __shared__ double As_r[2][BK][BM + PAD], As_i[2][BK][BM + PAD];
  __shared__ double Bs_r[2][BK][BN + PAD], Bs_i[2][BK][BN + PAD];

  v4d acc_r[2][2], acc_i[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) { acc_r[i][j] = (v4d){0, 0, 0, 0}; acc_i[i][j] = (v4d){0, 0, 0, 0}; }

  // global -> register assignment (4 elements of A and 4 of B per thread per K tile)
  const int ar = t >> 2, ak = (t & 3) * 4;          // A as stored: row ar, k ak..ak+3
  const int ak2 = t >> 4, ar2 = (t & 15) * 4;       // A^H: stored row (k) ak2, m ar2..ar2+3
  const int bk = t >> 4, bn = (t & 15) * 4;         // B as stored: k row bk, cols bn..bn+3
  const int bn2 = t >> 2, bk2 = (t & 3) * 4;        // B^H: stored row (n) bn2, k bk2..bk2+3
  cd ra[4], rb[4];

  auto fetch = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      cd v = mkc(0.0, 0.0);
      if (!CA) {
        const int row = tm + ar, kk = k0 + ak + q;
        if (row < m && kk < k) v = a[(long long)row * lda + kk];
      } else {
        const int kk = k0 + ak2, row = tm + ar2 + q;
        if (row < m && kk < k) v = a[(long long)kk * lda + row];
      }
      ra[q] = v;
      cd w = mkc(0.0, 0.0);
      if (!CB) {
        const int kk = k0 + bk, col = tn + bn + q;
        if (kk < k && col < n) w = bb[(long long)kk * ldb + col];
      } else {
        const int col = tn + bn2, kk = k0 + bk2 + q;
        if (kk < k && col < n) w = bb[(long long)col * ldb + kk];
      }
      rb[q] = w;
    }
  };
  auto stash = [&](int st) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (!CA) { As_r[st][ak + q][ar] = ra[q].x; As_i[st][ak + q][ar] = ra[q].y; }
      else { As_r[st][ak2][ar2 + q] = ra[q].x; As_i[st][ak2][ar2 + q] = -ra[q].y; }
      if (!CB) { Bs_r[st][bk][bn + q] = rb[q].x; Bs_i[st][bk][bn + q] = rb[q].y; }
      else { Bs_r[st][bk2 + q][bn2] = rb[q].x; Bs_i[st][bk2 + q][bn2] = -rb[q].y; }
    }
  };

  const int nk = (k + BK - 1) / BK;
  if (nk > 0) {
    fetch(0);
    stash(0);
  }
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int st = kt & 1;
    if (kt + 1 < nk) fetch((kt + 1) * BK);
#pragma unroll
    for (int s = 0; s < BK / 4; ++s) {
      const int kr = s * 4 + (lane >> 4);
      double ar_[2], ai_[2], br_[2], bi_[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        ar_[i] = As_r[st][kr][wm + i * 16 + (lane & 15)];
        ai_[i] = As_i[st][kr][wm + i * 16 + (lane & 15)];
        br_[i] = Bs_r[st][kr][wn + i * 16 + (lane & 15)];
        bi_[i] = Bs_i[st][kr][wn + i * 16 + (lane & 15)];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          acc_r[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar_[i], br_[j], acc_r[i][j], 0, 0, 0);
          acc_r[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai_[i], bi_[j], acc_r[i][j], 0, 0, 0);
          acc_i[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar_[i], bi_[j], acc_i[i][j], 0, 0, 0);
          acc_i[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai_[i], br_[j], acc_i[i][j], 0, 0, 0);
        }
    }
    if (kt + 1 < nk) stash(st ^ 1);
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = tm + wm + i * 16 + (lane >> 4) + 4 * r;
        const int col = tn + wn + j * 16 + (lane & 15);
        if (row < m && col < n) {
          cd v = cmul(alpha, mkc(acc_r[i][j][r], acc_i[i][j][r]));
          cd* p = c + (long long)row * ldc + col;
          if (!beta_zero) v = cadd(v, cmul(beta, *p));
          *p = v;
        }
      }
}

}  // namespace

hipError_t wb_zgemm_ops_launch(int opa, int opb, int m, int n, int k, cd alpha, BatchPtr A, int lda,
                               BatchPtr B, int ldb, cd beta, BatchPtr C, int ldc, int batch,
                               hipStream_t stream) {
  if (m <= 0 || n <= 0 || batch <= 0) return hipSuccess;
  dim3 grid((n + BN - 1) / BN, (m + BM - 1) / BM, batch);
  int beta_zero = (beta.x == 0.0 && beta.y == 0.0) ? 1 : 0;
  const double fl = 8.0 * m * n * (double)(k > 0 ? k : 0) * batch;
  if (!opa && !opb) {
    WB_TIMED_F(WB_T_ZGEMM, stream, fl,
               hipLaunchKernelGGL((k_zgemm<false, false>), grid, dim3(256), 0, stream, m, n, k, alpha,
                                  A, lda, B, ldb, beta, C, ldc, beta_zero));
  } else if (opa && !opb) {
    WB_TIMED_F(WB_T_ZGEMM, stream, fl,
               hipLaunchKernelGGL((k_zgemm<true, false>), grid, dim3(256), 0, stream, m, n, k, alpha,
                                  A, lda, B, ldb, beta, C, ldc, beta_zero));
  } else if (!opa && opb) {
    WB_TIMED_F(WB_T_ZGEMM, stream, fl,
               hipLaunchKernelGGL((k_zgemm<false, true>), grid, dim3(256), 0, stream, m, n, k, alpha,
                                  A, lda, B, ldb, beta, C, ldc, beta_zero));
  } else {
    WB_TIMED_F(WB_T_ZGEMM, stream, fl,
               hipLaunchKernelGGL((k_zgemm<true, true>), grid, dim3(256), 0, stream, m, n, k, alpha,
                                  A, lda, B, ldb, beta, C, ldc, beta_zero));
  }
  return hipGetLastError();
}

hipError_t wb_zgemm_launch(int m, int n, int k, cd alpha, BatchPtr A, int lda, BatchPtr B, int ldb,
                           cd beta, BatchPtr C, int ldc, int batch, hipStream_t stream) {
  return wb_zgemm_ops_launch(0, 0, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, batch, stream);
}
