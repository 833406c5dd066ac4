// Batched complex FP64 GEMM on CDNA4 matrix cores.
//
// C[b] = alpha * op(A[b]) * op(B[b]) + beta * C[b], op(A) m x k, op(B) k x n, C m x n, all
// row-major; op is the identity or the conjugate transpose (then A is stored k x m, B n x k,
// and the tile loads walk the stored rows so global reads stay coalesced).
// A complex product is four real v_mfma_f64_16x16x4f64 streams (re*re, -im*im, re*im, im*re)
// into two accumulators. A workgroup of 4 or 8 waves owns a BM x BN tile of C (64 x 64 by
// default: four waves of 2 x 2 MFMA blocks); K is staged through LDS 8 at a time (16 for the
// skinny tile), split into real and imaginary planes so every MFMA operand is one 8-byte LDS
// read, in two LDS stages: the next K tile is fetched into registers while the current one
// feeds the matrix cores. The short K tile keeps the LDS stages at 33 KiB and the registers
// under 128 per lane, so four waves share each SIMD and one wave's barrier wait is covered by
// the others' MFMAs (measured: 60 TF/s on 256 x 722^3 against 51 TF/s with 16-deep tiles and
// two waves per SIMD). The tile shape is picked per GEMM shape by the autotuner below.
//
// MFMA f64 16x16x4 lane layout (measured on gfx950): A[i][kk] in lane i + 16*kk, B[kk][j] in
// lane j + 16*kk, D[i][j] in lane j + 16*(i%4), register i/4.
#include "wb_common.h"
#include "wb_timing.h"

#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

namespace {

constexpr int PAD = 2;
typedef double v4d __attribute__((ext_vector_type(4)));

// Tile configuration: BM x BN of C per workgroup of NW waves, K staged BK at a time, the waves
// laid out WGM (along M) x NW / WGM (along N), each owning a block of 16 x 16 MFMA tiles; MINB
// workgroups per CU bound the registers. Every configuration accumulates each C entry over k
// in the same order (MFMA steps of 4 in increasing k), so they give bitwise identical results
// and the choice is a pure speed question.
template <int BM, int BN, int BK, int WGM, int NW, int MINB, bool CA, bool CB>
__global__ __launch_bounds__(NW * 64, MINB) void k_zgemm(int m, int n, int k, cd alpha, BatchPtr A, int lda,
                                                         BatchPtr B, int ldb, cd beta, BatchPtr C, int ldc,
                                                         int beta_zero) {
  constexpr int NT = NW * 64;
  constexpr int WGN = NW / WGM;
  constexpr int WM = BM / WGM, WN = BN / WGN;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int EA = BM * BK / NT, EB = BK * BN / NT;      // elements per thread per K tile
  constexpr int VA = EA < 4 ? EA : 4, VB = EB < 4 ? EB : 4; // consecutive elements per load group
  constexpr int GA = EA / VA, GB = EB / VB;
  static_assert(WGM * WGN == NW && FM * 16 == WM && FN * 16 == WN && BK % 4 == 0, "tile shape");
  static_assert(EA >= 1 && EB >= 1 && GA * VA == EA && GB * VB == EB, "tile loads");
  const long long b = blockIdx.z;
  const cd* __restrict__ a = A.ptr(b);
  const cd* __restrict__ bb = B.ptr(b);
  cd* __restrict__ c = C.ptr(b);
  const int tm = blockIdx.y * BM, tn = blockIdx.x * BN;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = (wave / WGN) * WM, wn = (wave % WGN) * WN;

  // two LDS stages: the next K tile is fetched into registers while the current one feeds the
  // matrix cores, then written to the other stage; one barrier per K tile
  __shared__ double As_r[2][BK][BM + PAD], As_i[2][BK][BM + PAD];
  __shared__ double Bs_r[2][BK][BN + PAD], Bs_i[2][BK][BN + PAD];

  v4d acc_r[FM][FN], acc_i[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) { acc_r[i][j] = (v4d){0, 0, 0, 0}; acc_i[i][j] = (v4d){0, 0, 0, 0}; }

  cd ra[GA][VA], rb[GB][VB];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int g = 0; g < GA; ++g) {
      const int f = t + NT * g;
#pragma unroll
      for (int q = 0; q < VA; ++q) {
        cd v = mkc(0.0, 0.0);
        if (!CA) {   // A stored m x k: row f / (BK/VA), k (f % (BK/VA)) * VA + q
          const int row = tm + f / (BK / VA), kk = k0 + (f % (BK / VA)) * VA + q;
          if (row < m && kk < k) v = a[(long long)row * lda + kk];
        } else {     // A stored k x m
          const int kk = k0 + f / (BM / VA), row = tm + (f % (BM / VA)) * VA + q;
          if (row < m && kk < k) v = a[(long long)kk * lda + row];
        }
        ra[g][q] = v;
      }
    }
#pragma unroll
    for (int g = 0; g < GB; ++g) {
      const int f = t + NT * g;
#pragma unroll
      for (int q = 0; q < VB; ++q) {
        cd w = mkc(0.0, 0.0);
        if (!CB) {   // B stored k x n
          const int kk = k0 + f / (BN / VB), col = tn + (f % (BN / VB)) * VB + q;
          if (kk < k && col < n) w = bb[(long long)kk * ldb + col];
        } else {     // B stored n x k
          const int col = tn + f / (BK / VB), kk = k0 + (f % (BK / VB)) * VB + q;
          if (kk < k && col < n) w = bb[(long long)col * ldb + kk];
        }
        rb[g][q] = w;
      }
    }
  };
  auto stash = [&](int st) {
#pragma unroll
    for (int g = 0; g < GA; ++g) {
      const int f = t + NT * g;
#pragma unroll
      for (int q = 0; q < VA; ++q) {
        if (!CA) {
          const int r = f / (BK / VA), kk = (f % (BK / VA)) * VA + q;
          As_r[st][kk][r] = ra[g][q].x; As_i[st][kk][r] = ra[g][q].y;
        } else {
          const int kk = f / (BM / VA), r = (f % (BM / VA)) * VA + q;
          As_r[st][kk][r] = ra[g][q].x; As_i[st][kk][r] = -ra[g][q].y;
        }
      }
    }
#pragma unroll
    for (int g = 0; g < GB; ++g) {
      const int f = t + NT * g;
#pragma unroll
      for (int q = 0; q < VB; ++q) {
        if (!CB) {
          const int kk = f / (BN / VB), cc = (f % (BN / VB)) * VB + q;
          Bs_r[st][kk][cc] = rb[g][q].x; Bs_i[st][kk][cc] = rb[g][q].y;
        } else {
          const int cc = f / (BK / VB), kk = (f % (BK / VB)) * VB + q;
          Bs_r[st][kk][cc] = rb[g][q].x; Bs_i[st][kk][cc] = -rb[g][q].y;
        }
      }
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
    for (int s4 = 0; s4 < BK / 4; ++s4) {
      const int kr = s4 * 4 + (lane >> 4);
      double ar_[FM], ai_[FM], br_[FN], bi_[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        ar_[i] = As_r[st][kr][wm + i * 16 + (lane & 15)];
        ai_[i] = As_i[st][kr][wm + i * 16 + (lane & 15)];
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        br_[j] = Bs_r[st][kr][wn + j * 16 + (lane & 15)];
        bi_[j] = Bs_i[st][kr][wn + j * 16 + (lane & 15)];
      }
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
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
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
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

// The candidate configurations.
struct GemmCfg { int bm, bn, nw; const char* name; };
constexpr int NCFG = 4;
constexpr GemmCfg CFGS[NCFG] = {{64, 64, 4, "64x64"}, {128, 64, 8, "128x64"}, {64, 128, 8, "64x128"},
                                {128, 16, 4, "128x16"}};

template <int CFG, bool CA, bool CB>
void launch_cfg_ops(dim3 grid, hipStream_t s, int m, int n, int k, cd alpha, BatchPtr A, int lda, BatchPtr B,
                    int ldb, cd beta, BatchPtr C, int ldc, int bz) {
  if constexpr (CFG == 0)
    hipLaunchKernelGGL((k_zgemm<64, 64, 8, 2, 4, 3, CA, CB>), grid, dim3(256), 0, s, m, n, k, alpha, A, lda, B,
                       ldb, beta, C, ldc, bz);
  else if constexpr (CFG == 1)
    hipLaunchKernelGGL((k_zgemm<128, 64, 8, 4, 8, 2, CA, CB>), grid, dim3(512), 0, s, m, n, k, alpha, A, lda, B,
                       ldb, beta, C, ldc, bz);
  else if constexpr (CFG == 2)
    hipLaunchKernelGGL((k_zgemm<64, 128, 8, 2, 8, 2, CA, CB>), grid, dim3(512), 0, s, m, n, k, alpha, A, lda, B,
                       ldb, beta, C, ldc, bz);
  else
    hipLaunchKernelGGL((k_zgemm<128, 16, 16, 4, 4, 1, CA, CB>), grid, dim3(256), 0, s, m, n, k, alpha, A, lda, B,
                       ldb, beta, C, ldc, bz);
}

template <int CFG>
void launch_cfg(int opa, int opb, dim3 grid, hipStream_t s, int m, int n, int k, cd alpha, BatchPtr A, int lda,
                BatchPtr B, int ldb, cd beta, BatchPtr C, int ldc, int bz) {
  if (!opa && !opb) launch_cfg_ops<CFG, false, false>(grid, s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, bz);
  else if (opa && !opb) launch_cfg_ops<CFG, true, false>(grid, s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, bz);
  else if (!opa && opb) launch_cfg_ops<CFG, false, true>(grid, s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, bz);
  else launch_cfg_ops<CFG, true, true>(grid, s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, bz);
}

void launch_any(int cfg, int opa, int opb, int m, int n, int k, cd alpha, BatchPtr A, int lda, BatchPtr B, int ldb,
                cd beta, BatchPtr C, int ldc, int batch, hipStream_t s) {
  const dim3 grid((n + CFGS[cfg].bn - 1) / CFGS[cfg].bn, (m + CFGS[cfg].bm - 1) / CFGS[cfg].bm, batch);
  const int bz = (beta.x == 0.0 && beta.y == 0.0) ? 1 : 0;
  switch (cfg) {
    case 1: launch_cfg<1>(opa, opb, grid, s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, bz); break;
    case 2: launch_cfg<2>(opa, opb, grid, s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, bz); break;
    case 3: launch_cfg<3>(opa, opb, grid, s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, bz); break;
    default: launch_cfg<0>(opa, opb, grid, s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, bz); break;
  }
}

// ---------------------------------------------------------------------------- autotuner
// Shape key (op(A), op(B), m, n, k, batch) -> configuration. Untuned shapes use configuration 0.
// With recording on, launches note their shapes; wb_gemm_autotune_recorded then times every
// candidate on the caller's scratch for each recorded shape (outside any solve: it waits for
// the stream) and keeps the fastest. The table is per process (shared by all devices: one GPU
// model).
typedef std::tuple<int, int, int, int, int, int> ShapeKey;
std::mutex g_tune_mu;
std::map<ShapeKey, int> g_table;
std::map<ShapeKey, int> g_recorded;
bool g_record = false;

int lookup(const ShapeKey& key) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (g_record) g_recorded.emplace(key, 0);
  auto it = g_table.find(key);
  return it == g_table.end() ? 0 : it->second;
}

}  // namespace

hipError_t wb_zgemm_ops_launch(int opa, int opb, int m, int n, int k, cd alpha, BatchPtr A, int lda,
                               BatchPtr B, int ldb, cd beta, BatchPtr C, int ldc, int batch,
                               hipStream_t stream) {
  if (m <= 0 || n <= 0 || batch <= 0) return hipSuccess;
  const int cfg = lookup(ShapeKey(opa, opb, m, n, k, batch));
  const double fl = 8.0 * m * n * (double)(k > 0 ? k : 0) * batch;
  WB_TIMED_F(WB_T_ZGEMM, stream, fl,
             launch_any(cfg, opa, opb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, batch, stream));
  return hipGetLastError();
}

hipError_t wb_zgemm_launch(int m, int n, int k, cd alpha, BatchPtr A, int lda, BatchPtr B, int ldb,
                           cd beta, BatchPtr C, int ldc, int batch, hipStream_t stream) {
  return wb_zgemm_ops_launch(0, 0, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, batch, stream);
}

void wb_gemm_record(int on) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_record = on != 0;
}

// Scratch bytes wb_gemm_autotune_recorded needs for the recorded, not yet tuned shapes.
size_t wb_gemm_autotune_bytes() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  size_t need = 0;
  for (auto& kv : g_recorded) {
    if (g_table.count(kv.first)) continue;
    const ShapeKey& key = kv.first;
    const size_t m = std::get<2>(key), n = std::get<3>(key), k = std::get<4>(key), b = std::get<5>(key);
    const size_t e = (m * k + k * n + m * n) * b;
    if (e > need) need = e;
  }
  return need * sizeof(cd);
}

// Tune every recorded, not yet tuned shape on the caller's scratch buffer (zeroed here): each
// candidate runs once to warm up, then `reps` times between events; the fastest wins. The call
// waits for the stream. Returns the number of shapes tuned, or -1 (HIP error in *err).
int wb_gemm_autotune_recorded(int reps, void* scratch, size_t bytes, hipStream_t stream, hipError_t* err) {
  std::vector<ShapeKey> todo;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (auto& kv : g_recorded)
      if (!g_table.count(kv.first)) todo.push_back(kv.first);
  }
  *err = hipSuccess;
  if (todo.empty()) return 0;
  cd* buf = (cd*)scratch;
  if ((*err = hipMemsetAsync(buf, 0, bytes, stream)) != hipSuccess) return -1;
  hipEvent_t e0, e1;
  if ((*err = hipEventCreate(&e0)) != hipSuccess) return -1;
  if ((*err = hipEventCreate(&e1)) != hipSuccess) { (void)hipEventDestroy(e0); return -1; }
  int tuned = 0;
  for (auto& key : todo) {
    const int opa = std::get<0>(key), opb = std::get<1>(key);
    const int m = std::get<2>(key), n = std::get<3>(key), k = std::get<4>(key), b = std::get<5>(key);
    if (((size_t)m * k + (size_t)k * n + (size_t)m * n) * b * sizeof(cd) > bytes) continue;
    cd* A = buf;
    cd* B = A + (size_t)m * k * b;
    cd* C = B + (size_t)k * n * b;
    const int lda = opa ? m : k, ldb = opb ? k : n;
    const BatchPtr pa = bp_strided(A, (long long)m * k), pb = bp_strided(B, (long long)k * n),
                   pc = bp_strided(C, (long long)m * n);
    int best = 0;
    float best_ms = 0.0f;
    for (int cfg = 0; cfg < NCFG; ++cfg) {
      launch_any(cfg, opa, opb, m, n, k, mkc(1.0, 0.0), pa, lda, pb, ldb, mkc(0.0, 0.0), pc, n, b, stream);
      (void)hipEventRecord(e0, stream);
      for (int r = 0; r < reps; ++r)
        launch_any(cfg, opa, opb, m, n, k, mkc(1.0, 0.0), pa, lda, pb, ldb, mkc(0.0, 0.0), pc, n, b, stream);
      (void)hipEventRecord(e1, stream);
      if ((*err = hipEventSynchronize(e1)) != hipSuccess) break;
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (cfg == 0 || ms < best_ms) { best = cfg; best_ms = ms; }
    }
    if (*err != hipSuccess) break;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_table[key] = best;
    ++tuned;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return *err == hipSuccess ? tuned : -1;
}

// The table as text lines "opa opb m n k batch config", and back.
size_t wb_gemm_table_text(char* out, size_t cap) {
  std::string s;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    char line[128];
    for (auto& kv : g_table) {
      const ShapeKey& k = kv.first;
      snprintf(line, sizeof line, "%d %d %d %d %d %d %d\n", std::get<0>(k), std::get<1>(k), std::get<2>(k),
               std::get<3>(k), std::get<4>(k), std::get<5>(k), kv.second);
      s += line;
    }
  }
  if (out && cap > 0) {
    const size_t nc = s.size() < cap - 1 ? s.size() : cap - 1;
    s.copy(out, nc);
    out[nc] = '\0';
  }
  return s.size() + 1;
}

int wb_gemm_table_load(const char* text) {
  if (!text) return 0;
  int added = 0;
  const char* p = text;
  while (*p) {
    int v[7];
    int used = 0;
    if (sscanf(p, "%d %d %d %d %d %d %d%n", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &used) != 7) break;
    if (v[6] >= 0 && v[6] < NCFG) {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      g_table[ShapeKey(v[0], v[1], v[2], v[3], v[4], v[5])] = v[6];
      ++added;
    }
    p += used;
    while (*p == '\n' || *p == ' ') ++p;
  }
  return added;
}

void wb_gemm_table_clear() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_table.clear();
  g_recorded.clear();
}

const char* wb_gemm_config_name(int cfg) { return (cfg >= 0 && cfg < NCFG) ? CFGS[cfg].name : "?"; }
int wb_gemm_config_count() { return NCFG; }

// Force a configuration for one shape (tests: every configuration must agree bit for bit).
void wb_gemm_table_set(int opa, int opb, int m, int n, int k, int batch, int cfg) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (cfg < 0) g_table.erase(ShapeKey(opa, opb, m, n, k, batch));
  else if (cfg < NCFG) g_table[ShapeKey(opa, opb, m, n, k, batch)] = cfg;
}
