// Grazing-cluster rebuild of 2D layer modes (reference src/wavebit/rcwa/_engine/solve_2d.py:
// 118-228, _rebuild_grazing_cluster, entered at :2463-2485 whenever min|w| < 1e-9 max|w|).
//
// At a Rayleigh point an order sits on a layer's light line: the product operator Omega2 =
// Omega_L Omega_R has a cluster of eigenvalues near zero whose eigenvectors are fixed by round-off,
// so V = Omega_R W / q satisfies Omega_R W = V q but not Omega_L V = W q. What is stable is the
// invariant subspace pair: S_E (the cluster of Omega_L Omega_R, columns J of W) and S_H (the same
// cluster of the swapped product Omega_R Omega_L). With orthonormal bases Q_E, Q_H, the factors
// restricted to them, R_r = Q_H^H Omega_R Q_E and L_r = Q_E^H Omega_L Q_H, give the k x k
// first-order problem [[0, L_r], [R_r, 0]] g = mu g whose eigenpairs satisfy both relations; of
// its 2k eigenvectors (+-mu twins share a field pattern) k are picked greedily so that their E
// parts and their H parts both stay independent, each branch-fixed to mode_q(mu^2). A picked rate
// below 1e-10 sqrt(max|w|) is a numerically defective cluster: no exponential mode basis exists
// there and the layer is refused (WB_FLAG_DEFECTIVE).
//
// Device stages, per (point, layer) pair of a batch (one launch sequence for the whole batch):
//   k_omega_lr      dense Omega_L, Omega_R (two n x n per pair) from the conv matrices
//   GEMM + zgeev    H = Omega_R Omega_L and its eigenpairs
//   k_graze_basis   the cluster J of w (first eigensolve) and Jh of H, orthonormal bases Q_E, Q_H
//                   (classical Gram-Schmidt, two passes, one workgroup per pair)
//   GEMMs           R_r = Q_H^H Omega_R Q_E, L_r = Q_E^H Omega_L Q_H (K x K, K = GRAZE_KMAX)
//   k_graze_small   [[0, L_r], [R_r, 0]] in the leading 2k x 2k block of a 2K x 2K matrix whose
//                   trailing diagonal holds distinct values far outside its spectrum
//   zgeev (2K)      its eigenpairs; the 2k inside the block's spectral radius are the block's
//   k_graze_finish  defect test, greedy pick, root and branch sign
//   GEMMs           W_c = Q_E g_E, V_c = Q_H g_H
//   k_graze_patch   (after the s/p rotation) the rotated W_c, V_c and q_c replace columns J
// A pair whose cluster is larger than GRAZE_KMAX, whose two products disagree on its size, or whose
// small eigensolve fails keeps its modes as computed (the reference keeps them and warns).
#include "wb_rcwa.h"

namespace {

constexpr double GRAZE_REL = 1e-9;       // solve_2d.py:44-56 (_COUPLED_REL)
constexpr double DEFECTIVE_REL = 1e-10;  // solve_2d.py:58-69 (_DEFECTIVE_REL)
constexpr int K = WB_GRAZE_KMAX;         // largest cluster rebuilt
constexpr int M = 2 * K;                 // order of the padded first-order eigenproblem
constexpr int NTB = 256;                 // threads of the per-pair workgroups
constexpr int RMAX = WB_GRAZE_NMAX / NTB;   // rows per thread in k_graze_basis

__device__ inline cd block_sum(cd x, cd* sh) {
  const int t = threadIdx.x;
  sh[t] = x;
  __syncthreads();
  for (int s = NTB / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] = cadd(sh[t], sh[t + s]);
    __syncthreads();
  }
  const cd r = sh[0];
  __syncthreads();
  return r;
}

__device__ inline double block_max(double x, double* sh) {
  const int t = threadIdx.x;
  sh[t] = x;
  __syncthreads();
  for (int s = NTB / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] = fmax(sh[t], sh[t + s]);
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// Dense Omega_L = [[-Cc, B], [-D, E]] and Omega_R = [[-KxKy + epyx, Kx^2 - epy],
// [epx - Ky^2, KyKx - epxy]] of each pair (the off-diagonal slots only when offd): Omega_L at
// olr + b * 2 n^2, Omega_R n^2 after it (B = Kx epz_i Kx - I, Cc = Kx epz_i Ky,
// D = Ky epz_i Ky - I, E = Ky epz_i Kx; solve_2d.py:2377-2411, 2474-2476).
__global__ void k_omega_lr(int N, int L, BatchPtr epx, BatchPtr epy, BatchPtr epz, BatchPtr epxy,
                           BatchPtr epyx, int offd, const cd* ov, cd* olr) {
  const long long b = blockIdx.y;
  const int p = (int)(b / L);
  const int n = 2 * N;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)N * N) return;
  const int i = (int)(e / N), j = (int)(e % N);
  const cd* o = ov + (long long)p * OV_COUNT * N;
  const cd kxi = o[OV_KX * N + i], kyi = o[OV_KY * N + i];
  const cd kxj = o[OV_KX * N + j], kyj = o[OV_KY * N + j];
  const cd z = epz.ptr(b)[e];
  const bool dg = (i == j);
  const cd one = mkc(1.0, 0.0), zero = mkc(0.0, 0.0);
  cd* OL = olr + b * 2LL * n * n;
  cd* OR = OL + (long long)n * n;
  const cd Bm = dg ? csub(cmul(cmul(kxi, z), kxj), one) : cmul(cmul(kxi, z), kxj);
  const cd Dm = dg ? csub(cmul(cmul(kyi, z), kyj), one) : cmul(cmul(kyi, z), kyj);
  OL[(long long)i * n + j] = cneg(cmul(cmul(kxi, z), kyj));
  OL[(long long)i * n + N + j] = Bm;
  OL[(long long)(N + i) * n + j] = cneg(Dm);
  OL[(long long)(N + i) * n + N + j] = cmul(cmul(kyi, z), kxj);
  const cd ex = epx.ptr(b)[e], ey = epy.ptr(b)[e];
  cd o00 = dg ? cneg(cmul(kxi, kyi)) : zero;
  cd o11 = dg ? cmul(kyi, kxi) : zero;
  if (offd) {
    o00 = cadd(o00, epyx.ptr(b)[e]);
    o11 = csub(o11, epxy.ptr(b)[e]);
  }
  OR[(long long)i * n + j] = o00;
  OR[(long long)i * n + N + j] = dg ? csub(cmul(kxi, kxi), ey) : cneg(ey);
  OR[(long long)(N + i) * n + j] = dg ? csub(ex, cmul(kyi, kyi)) : ex;
  OR[(long long)(N + i) * n + N + j] = o11;
}

// Orthonormal basis (n x K, row-major, columns >= k zero) of columns idx[0..k) of the n x n
// matrix S, by classical Gram-Schmidt with a second pass. Thread t owns rows t + NTB m. Returns
// false (uniformly) when a column is dependent on the earlier ones.
__device__ bool orthonormal_basis(int n, int k, const int* idx, const cd* S, cd* Q, cd* sh) {
  const int t = threadIdx.x;
  for (int i = 0; i < k; ++i) {
    cd v[RMAX];
#pragma unroll
    for (int m = 0; m < RMAX; ++m) {
      const int r = t + NTB * m;
      v[m] = (r < n) ? S[(long long)r * n + idx[i]] : mkc(0.0, 0.0);
    }
    for (int pass = 0; pass < 2; ++pass) {
      for (int j = 0; j < i; ++j) {
        cd part = mkc(0.0, 0.0);
#pragma unroll
        for (int m = 0; m < RMAX; ++m) {
          const int r = t + NTB * m;
          if (r < n) part = cadd(part, cmulc(Q[(long long)r * K + j], v[m]));
        }
        const cd d = block_sum(part, sh);
#pragma unroll
        for (int m = 0; m < RMAX; ++m) {
          const int r = t + NTB * m;
          if (r < n) v[m] = csub(v[m], cmul(Q[(long long)r * K + j], d));
        }
      }
    }
    double part = 0.0;
#pragma unroll
    for (int m = 0; m < RMAX; ++m) part += cnorm2(v[m]);
    const double nrm = sqrt(block_sum(mkc(part, 0.0), sh).x);
    if (!(nrm > 1e-13)) return false;
    const double inv = 1.0 / nrm;
#pragma unroll
    for (int m = 0; m < RMAX; ++m) {
      const int r = t + NTB * m;
      if (r < n) Q[(long long)r * K + i] = cscale(v[m], inv);
    }
  }
  return true;
}

// The cluster of each pair and its two bases. One workgroup per pair.
__global__ __launch_bounds__(NTB) void k_graze_basis(int n, const cd* w, const cd* W, const cd* wh,
                                                     const cd* Vh, wb_graze_ws g) {
  __shared__ cd sh[NTB];
  __shared__ double shd[NTB];
  __shared__ int sJ[K], sJh[K];
  __shared__ int s_status, s_k;
  const long long b = blockIdx.x;
  const int t = threadIdx.x;
  const cd* wb = w + b * n;
  const cd* whb = wh + b * n;
  double mx = 0.0, mxh = 0.0;
  for (int i = t; i < n; i += NTB) {
    mx = fmax(mx, cabs(wb[i]));
    mxh = fmax(mxh, cabs(whb[i]));
  }
  const double wmax = block_max(mx, shd);
  const double whmax = block_max(mxh, shd);
  if (t == 0) {
    int cnt = 0, cnth = 0;
    for (int i = 0; i < n; ++i) {
      if (cabs(wb[i]) < GRAZE_REL * wmax) { if (cnt < K) sJ[cnt] = i; ++cnt; }
      if (cabs(whb[i]) < GRAZE_REL * whmax) { if (cnth < K) sJh[cnth] = i; ++cnth; }
    }
    int st = WB_GRAZE_ACTIVE;
    if (cnt == 0) st = WB_GRAZE_NONE;
    else if (cnt > K || cnth != cnt || g.einfo[b] != 0) st = WB_GRAZE_KEPT;
    s_status = st;
    s_k = (cnt > K) ? K : cnt;
    g.status[b] = st;
    g.jn[b] = s_k;
    g.wmax[b] = wmax;
    for (int i = 0; i < K; ++i) g.jidx[b * K + i] = (i < s_k) ? sJ[i] : 0;
  }
  __syncthreads();
  cd* QE = g.QE + b * (long long)n * K;
  cd* QH = g.QH + b * (long long)n * K;
  for (int m = 0; m < RMAX; ++m) {       // zero by the same row ownership the basis writes with
    const int r = t + NTB * m;
    if (r < n)
      for (int c = 0; c < K; ++c) { QE[(long long)r * K + c] = mkc(0.0, 0.0); QH[(long long)r * K + c] = mkc(0.0, 0.0); }
  }
  if (s_status != WB_GRAZE_ACTIVE) return;
  const int k = s_k;
  const bool ok = orthonormal_basis(n, k, sJ, W + b * (long long)n * n, QE, sh) &&
                  orthonormal_basis(n, k, sJh, Vh + b * (long long)n * n, QH, sh);
  if (!ok && t == 0) g.status[b] = WB_GRAZE_KEPT;
}

// The padded first-order matrix of each pair (one thread per pair), and g_E, g_H cleared.
__global__ void k_graze_small(int B, wb_graze_ws g) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int k = (g.status[b] == WB_GRAZE_ACTIVE) ? g.jn[b] : 0;
  const cd* Lr = g.Lr + (long long)b * K * K;
  const cd* Rr = g.Rr + (long long)b * K * K;
  double big = 0.0;
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j) big = fmax(big, fmax(cabs(Lr[i * K + j]), cabs(Rr[i * K + j])));
  const double pad = 8.0 * (1.0 + 2.0 * K * big);
  g.pad[b] = pad;
  cd* Ms = g.Ms + (long long)b * M * M;
  for (int e = 0; e < M * M; ++e) Ms[e] = mkc(0.0, 0.0);
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j) {
      Ms[i * M + k + j] = Lr[i * K + j];
      Ms[(k + i) * M + j] = Rr[i * K + j];
    }
  for (int d = 2 * k; d < M; ++d) Ms[d * M + d] = mkc(pad * (1.0 + d), 0.0);
  for (int e = 0; e < K * K; ++e) {
    g.GE[(long long)b * K * K + e] = mkc(0.0, 0.0);
    g.GH[(long long)b * K * K + e] = mkc(0.0, 0.0);
  }
}

// Pick, test and branch-fix the rebuilt modes of each active pair (thread 0 of a wave per pair).
__global__ __launch_bounds__(64) void k_graze_finish(wb_graze_ws g) {
  __shared__ cd ge[K][M], gh[K][M], re[K][M], rh[K][M];
  __shared__ int cand[M];
  const long long b = blockIdx.x;
  if (threadIdx.x != 0 || g.status[b] != WB_GRAZE_ACTIVE) return;
  if (g.sinfo[b] != 0) { g.status[b] = WB_GRAZE_KEPT; return; }
  const int k = g.jn[b];
  const cd* mu = g.mu + b * M;
  const cd* vec = g.gv + b * (long long)M * M;
  int nc = 0;
  for (int c = 0; c < M; ++c)
    if (cabs(mu[c]) < 0.5 * g.pad[b]) { if (nc < M) cand[nc] = c; ++nc; }
  if (nc != 2 * k) { g.status[b] = WB_GRAZE_KEPT; return; }
  // The picks take one member of each +-mu twin, so the smallest picked rate is the smallest
  // rate of the block: a block with a rate below the defect floor is refused before any pick
  // (its eigenvectors, at a numerically defective point, need not be usable).
  double mumin = 1e300;
  for (int j = 0; j < nc; ++j) mumin = fmin(mumin, cabs(mu[cand[j]]));
  if (mumin < DEFECTIVE_REL * sqrt(g.wmax[b])) { g.status[b] = WB_GRAZE_DEFECTIVE; return; }
  for (int j = 0; j < nc; ++j) {
    const int c = cand[j];
    double s = 0.0;
    for (int r = 0; r < 2 * k; ++r) s += cnorm2(vec[r * M + c]);
    if (!(s > 0.0)) { g.status[b] = WB_GRAZE_KEPT; return; }
    const double inv = 1.0 / sqrt(s);
    for (int r = 0; r < k; ++r) {
      ge[r][j] = re[r][j] = cscale(vec[r * M + c], inv);
      gh[r][j] = rh[r][j] = cscale(vec[(k + r) * M + c], inv);
    }
  }
  int pick[K];
  bool taken[M];
  for (int j = 0; j < nc; ++j) taken[j] = false;
  for (int it = 0; it < k; ++it) {
    int best = -1;
    double bs = -1.0;
    for (int j = 0; j < nc; ++j) {
      if (taken[j]) continue;
      double s = 0.0;
      for (int r = 0; r < k; ++r) s += cnorm2(re[r][j]) + cnorm2(rh[r][j]);
      if (s > bs) { bs = s; best = j; }
    }
    pick[it] = best;
    taken[best] = true;
    for (int part = 0; part < 2; ++part) {
      cd (*R)[M] = part ? rh : re;
      double nu = 0.0;
      for (int r = 0; r < k; ++r) nu += cnorm2(R[r][best]);
      nu = sqrt(nu);
      if (!(nu > 1e-150)) continue;
      cd u[K];
      for (int r = 0; r < k; ++r) u[r] = cscale(R[r][best], 1.0 / nu);
      for (int j = 0; j < nc; ++j) {
        cd d = mkc(0.0, 0.0);
        for (int r = 0; r < k; ++r) d = cadd(d, cmulc(u[r], R[r][j]));
        for (int r = 0; r < k; ++r) R[r][j] = csub(R[r][j], cmul(u[r], d));
      }
    }
  }
  cd* GE = g.GE + b * K * K;
  cd* GH = g.GH + b * K * K;
  for (int i = 0; i < k; ++i) {
    const int j = pick[i];
    const cd m_ = mu[cand[j]];
    const cd qv = mode_q(cmul(m_, m_));
    const double sign = (cabs(csub(qv, m_)) > cabs(cadd(qv, m_))) ? -1.0 : 1.0;
    for (int r = 0; r < k; ++r) {
      GE[r * K + i] = ge[r][j];
      GH[r * K + i] = cscale(gh[r][j], sign);
    }
    g.qc[b * K + i] = qv;
  }
  g.status[b] = WB_GRAZE_REBUILT;
}

// The rebuilt columns, rotated into the s/p basis as the modes' post-processing rotates the rest
// (solve_2d.py:2584-2616), over columns J of big_W, big_V, and their roots over q.
__global__ void k_graze_patch(int N, int L, const cd* ov, cd* bigW, cd* bigV, cd* q, wb_graze_ws g) {
  const long long b = blockIdx.y;
  if (g.status[b] != WB_GRAZE_REBUILT) return;
  const int n = 2 * N;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)N * K) return;
  const int r = (int)(e / K), i = (int)(e % K);
  if (i >= g.jn[b]) return;
  const int c = g.jidx[b * K + i];
  const int p = (int)(b / L);
  const cd* o = ov + (long long)p * OV_COUNT * N;
  const double fc = o[OV_FC * N + r].x, fs = o[OV_FS * N + r].x;
  const cd* Wc = g.WC + b * (long long)n * K;
  const cd* Vc = g.VC + b * (long long)n * K;
  const cd Wt = Wc[(long long)r * K + i], Wb = Wc[(long long)(N + r) * K + i];
  const cd Vt = Vc[(long long)r * K + i], Vb = Vc[(long long)(N + r) * K + i];
  const long long base = b * (long long)n * n;
  bigW[base + (long long)r * n + c] = csub(cscale(Wb, fc), cscale(Wt, fs));
  bigW[base + (long long)(N + r) * n + c] = cadd(cscale(Wt, fc), cscale(Wb, fs));
  bigV[base + (long long)r * n + c] = cadd(cscale(Vt, fc), cscale(Vb, fs));
  bigV[base + (long long)(N + r) * n + c] = csub(cscale(Vb, fc), cscale(Vt, fs));
  if (r == 0) q[b * n + c] = g.qc[b * K + i];
}

// The rebuilt columns over columns J of the raw (unrotated) W and V of the split solve's modes
// stage (pair b at b n^2), and their roots over q.
__global__ void k_graze_patch_raw(int n, cd* W, cd* V, cd* q, wb_graze_ws g) {
  const long long b = blockIdx.y;
  if (g.status[b] != WB_GRAZE_REBUILT) return;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n * K) return;
  const int r = (int)(e / K), i = (int)(e % K);
  if (i >= g.jn[b]) return;
  const int c = g.jidx[b * K + i];
  const long long base = b * (long long)n * n;
  W[base + (long long)r * n + c] = g.WC[b * (long long)n * K + (long long)r * K + i];
  V[base + (long long)r * n + c] = g.VC[b * (long long)n * K + (long long)r * K + i];
  if (r == 0) q[b * n + c] = g.qc[b * K + i];
}

// Flags: a rebuilt layer is marked WB_FLAG_REBUILT and its pair screen cleared (the reference
// returns the rebuilt modes without the pair gates); a defective cluster is WB_FLAG_DEFECTIVE.
__global__ void k_graze_flags(int B, const int* status, int* flags) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  if (status[b] == WB_GRAZE_REBUILT)
    flags[b] = (flags[b] | WB_FLAG_REBUILT) & ~(WB_FLAG_EP | WB_FLAG_EP_LOUD);
  else if (status[b] == WB_GRAZE_DEFECTIVE)
    flags[b] |= WB_FLAG_DEFECTIVE;
}

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

size_t wb_graze_bytes(int B, int n) {
  const size_t b = (size_t)B;
  size_t s = 0;
  s += align_up(b * sizeof(int)) * 5 + align_up(b * K * sizeof(int));   // jn status einfo sinfo ..
  s += align_up(b * sizeof(double)) * 2;                                 // wmax, pad
  s += align_up(b * (size_t)n * K * sizeof(cd)) * 6;                     // QE QH TE TH WC VC
  s += align_up(b * K * K * sizeof(cd)) * 4;                             // Rr Lr GE GH
  s += align_up(b * M * M * sizeof(cd)) * 4;                             // Ms gv Zs Xs
  s += align_up(b * M * sizeof(cd)) * 2;                                 // mu taus
  s += align_up(b * K * sizeof(cd));                                     // qc
  s += align_up(wb_hqr_ms_rot_bytes(M, B));
  return s + 256;
}

wb_graze_ws wb_graze_carve(void* ws, int B, int n) {
  char* p = (char*)ws;
  auto take = [&](size_t bytes) { char* r = p; p += align_up(bytes); return (void*)r; };
  const size_t b = (size_t)B;
  wb_graze_ws g;
  g.jn = (int*)take(b * sizeof(int));
  g.status = (int*)take(b * sizeof(int));
  g.einfo = (int*)take(b * sizeof(int));
  g.sinfo = (int*)take(b * sizeof(int));
  take(b * sizeof(int));
  g.jidx = (int*)take(b * K * sizeof(int));
  g.wmax = (double*)take(b * sizeof(double));
  g.pad = (double*)take(b * sizeof(double));
  g.QE = (cd*)take(b * (size_t)n * K * sizeof(cd));
  g.QH = (cd*)take(b * (size_t)n * K * sizeof(cd));
  g.TE = (cd*)take(b * (size_t)n * K * sizeof(cd));
  g.TH = (cd*)take(b * (size_t)n * K * sizeof(cd));
  g.WC = (cd*)take(b * (size_t)n * K * sizeof(cd));
  g.VC = (cd*)take(b * (size_t)n * K * sizeof(cd));
  g.Rr = (cd*)take(b * K * K * sizeof(cd));
  g.Lr = (cd*)take(b * K * K * sizeof(cd));
  g.GE = (cd*)take(b * K * K * sizeof(cd));
  g.GH = (cd*)take(b * K * K * sizeof(cd));
  g.Ms = (cd*)take(b * M * M * sizeof(cd));
  g.gv = (cd*)take(b * M * M * sizeof(cd));
  g.Zs = (cd*)take(b * M * M * sizeof(cd));
  g.Xs = (cd*)take(b * M * M * sizeof(cd));
  g.mu = (cd*)take(b * M * sizeof(cd));
  g.taus = (cd*)take(b * M * sizeof(cd));
  g.qc = (cd*)take(b * K * sizeof(cd));
  g.rot = take(wb_hqr_ms_rot_bytes(M, B));
  return g;
}

hipError_t wb_graze_stage(int N, int L, int B, const cd* w, const cd* W, BatchPtr epx,
                          BatchPtr epy, BatchPtr epz, BatchPtr epxy, BatchPtr epyx, int offd,
                          const cd* ov, cd* olr, cd* H, cd* Z, cd* X,
                          cd* Vh, cd* wh, cd* tau, void* rotws, const wb_graze_ws& g,
                          hipStream_t st) {
  const int n = 2 * N;
  if (n > WB_GRAZE_NMAX) return hipErrorInvalidValue;
  const long long nn = (long long)n * n;
  const cd one = mkc(1.0, 0.0), zero = mkc(0.0, 0.0);
  hipError_t e;
  hipLaunchKernelGGL(k_omega_lr, dim3((unsigned)(((long long)N * N + 255) / 256), (unsigned)B),
                     dim3(256), 0, st, N, L, epx, epy, epz, epxy, epyx, offd, ov, olr);
  WB_CHECK_LAUNCH();
  const BatchPtr OLp = bp_strided(olr, 2 * nn), ORp = bp_strided(olr + nn, 2 * nn);
  if ((e = wb_zgemm_launch(n, n, n, one, ORp, n, OLp, n, zero, bp_strided(H, nn), n, B, st))) return e;
  if ((e = wb_zgeev_launch(n, H, wh, Vh, Z, X, tau, rotws, g.einfo, B, st))) return e;
  hipLaunchKernelGGL(k_graze_basis, dim3(B), dim3(NTB), 0, st, n, w, W, wh, Vh, g);
  WB_CHECK_LAUNCH();
  const long long nK = (long long)n * K;
  if ((e = wb_zgemm_launch(n, K, n, one, ORp, n, bp_strided(g.QE, nK), K, zero, bp_strided(g.TE, nK),
                           K, B, st))) return e;
  if ((e = wb_zgemm_launch(n, K, n, one, OLp, n, bp_strided(g.QH, nK), K, zero, bp_strided(g.TH, nK),
                           K, B, st))) return e;
  if ((e = wb_zgemm_ops_launch(1, 0, K, K, n, one, bp_strided(g.QH, nK), K, bp_strided(g.TE, nK), K,
                               zero, bp_strided(g.Rr, K * K), K, B, st))) return e;
  if ((e = wb_zgemm_ops_launch(1, 0, K, K, n, one, bp_strided(g.QE, nK), K, bp_strided(g.TH, nK), K,
                               zero, bp_strided(g.Lr, K * K), K, B, st))) return e;
  hipLaunchKernelGGL(k_graze_small, dim3((B + 63) / 64), dim3(64), 0, st, B, g);
  WB_CHECK_LAUNCH();
  if ((e = wb_zgeev_launch(M, g.Ms, g.mu, g.gv, g.Zs, g.Xs, g.taus, g.rot, g.sinfo, B, st))) return e;
  hipLaunchKernelGGL(k_graze_finish, dim3(B), dim3(64), 0, st, g);
  WB_CHECK_LAUNCH();
  if ((e = wb_zgemm_launch(n, K, K, one, bp_strided(g.QE, nK), K, bp_strided(g.GE, K * K), K, zero,
                           bp_strided(g.WC, nK), K, B, st))) return e;
  if ((e = wb_zgemm_launch(n, K, K, one, bp_strided(g.QH, nK), K, bp_strided(g.GH, K * K), K, zero,
                           bp_strided(g.VC, nK), K, B, st))) return e;
  return hipSuccess;
}

hipError_t wb_graze_patch(int N, int L, int B, const cd* ov, cd* bigW, cd* bigV, cd* q,
                          const wb_graze_ws& g, int* flags, hipStream_t st) {
  hipLaunchKernelGGL(k_graze_patch, dim3((unsigned)(((long long)N * K + 255) / 256), (unsigned)B),
                     dim3(256), 0, st, N, L, ov, bigW, bigV, q, g);
  WB_CHECK_LAUNCH();
  if (flags) {
    hipLaunchKernelGGL(k_graze_flags, dim3((B + 255) / 256), dim3(256), 0, st, B, g.status, flags);
    WB_CHECK_LAUNCH();
  }
  return hipSuccess;
}

hipError_t wb_graze_patch_raw(int n, int B, cd* W, cd* V, cd* q, const wb_graze_ws& g, int* flags,
                              hipStream_t st) {
  hipLaunchKernelGGL(k_graze_patch_raw, dim3((unsigned)(((long long)n * K + 255) / 256), (unsigned)B),
                     dim3(256), 0, st, n, W, V, q, g);
  WB_CHECK_LAUNCH();
  if (flags) {
    hipLaunchKernelGGL(k_graze_flags, dim3((B + 255) / 256), dim3(256), 0, st, B, g.status, flags);
    WB_CHECK_LAUNCH();
  }
  return hipSuccess;
}
