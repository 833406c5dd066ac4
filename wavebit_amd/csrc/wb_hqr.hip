// Windowed small-bulge multishift QR for the complex Hessenberg eigenproblem, with Schur
// vectors (one 512-thread workgroup per matrix).
//
// A multishift sweep chases M single-shift bulges (plane rotations) down the active block
// [lo, hi] in a tightly packed chain: bulge j enters at step 3j and at step s acts at position
// k = lo + s - 3j. Spacing 3 keeps the bulges' rotation inputs disjoint, so each step computes
// all M rotations, then applies all M row operations, then all M column operations, and the
// result equals M consecutive single-shift sweeps (the implicit-Q argument behind
// Braman-Byers-Mathias small-bulge multishift QR; Golub & Van Loan 7.5.7).
//
// The chase runs in a W x W diagonal window held in LDS. Everything outside the window is
// deferred and applied from the recorded rotations by a register "band": in step s, bulge j
// always touches band slots 3(G-1-j) and 3(G-1-j)+1 of its group's band [lo+s-3(G-1), lo+s+1]
// (G = 8 bulges per group), so a strip of H or Z streams through B = 3(G-1)+2 registers with
// compile-time indices. With M = 16 the two groups of a strip run on the two half-waves in
// lockstep (band_apply_pair), so each element is read and written once per window (H right of
// and above the window) or once per sweep (Z) for all 16 bulges. The window's rotations are
// written to global memory during the chase and staged back into the window's LDS for the
// deferred updates. Shifts are the eigenvalues of the trailing M x M block (one wave's QR
// iteration, wave_qr); active blocks of at most nmin rows finish with a single-shift iteration
// held in LDS by one wave (small_block).
//
// Before every sweep, aggressive early deflation (Braman, Byers & Mathias, SIAM J. Matrix Anal.
// Appl. 23 (2002) 948-973; LAPACK zlaqr3) takes the trailing window of the active block to Schur
// form, deflates each eigenvalue whose spike entry s conj(U[0][j]) is negligible, moves the
// others to the top of the window by swaps of the triangular form (ztrexc's rotation), and
// returns the window to Hessenberg form (a reflector that folds the spike into one entry, then
// Householder reduction of the undeflated part); the undeflated eigenvalues are the next
// sweep's shifts. The window's own QR runs on one wave, so its cost grows as its cube: on the
// (9,9) layer matrices (n = 722, 256 per launch) windows of 16-24 rows take the iteration from
// 371 to 271-278 ms per launch, 32 rows to 301, 48 to 358 and 64 to 464 (a CPU model of the
// sweep counts: 93 sweeps without the window, 68 at 24 rows, 46 at 48).
//
// The single-wave iteration (wave_qr, shared with the fused small-order solve) lives in
// wb_waveqr.h with its deflation rules.
#include "wb_internal.h"
#include "wb_timing.h"
#include <float.h>
#include <type_traits>

#include "wb_waveqr.h"

namespace {

#ifndef WB_HQR_NT
#define WB_HQR_NT 512      // threads per matrix
#endif
#ifndef WB_HQR_M
#define WB_HQR_M 16        // bulges (shifts) per sweep: 8 or 16
#endif
#ifndef WB_HQR_NMIN
#define WB_HQR_NMIN ((3 * WB_HQR_M + 8 > 48) ? 3 * WB_HQR_M + 8 : 48)   // largest small block
#endif
#ifndef WB_HQR_WSTEP
#define WB_HQR_WSTEP 8     // window sizes tried, largest first
#endif
#ifndef WB_HQR_BC
#define WB_HQR_BC 2        // rotations per register chunk of the band updates
#endif
#ifndef WB_HQR_AED
#define WB_HQR_AED 20      // aggressive-early-deflation window (<= 64 rows; 0 turns it off)
#endif
static_assert(WB_HQR_AED == 0 || (WB_HQR_AED <= 64 && WB_HQR_AED >= WB_HQR_M),
              "the deflation window fits one wave and holds a sweep's shifts");
constexpr int NT = WB_HQR_NT;
static_assert(NT % 64 == 0, "whole waves");


// Largest k in (lo, hi] with a negligible subdiagonal (Ahues-Tisseur test), else lo. With
// `relaxed` (the iteration has stalled) the conventional test |h(k,k-1)| <= ulp (|h(k-1,k-1)| +
// |h(k,k)|) alone decides: still a normwise backward-stable deflation, and the one that lets an
// exactly degenerate pair coupled at round-off (a uniform layer's two polarizations of one order)
// split, where the Ahues-Tisseur refinement waits for a separation QR cannot produce.
__device__ int find_split(const cd* H, int n, int lo, int hi, int* s_k, bool relaxed = false) {
  const int t = threadIdx.x;
  const double smlnum = SAFMIN * ((double)n / ULP);
  int found = lo;
  for (int k = lo + 1 + t; k <= hi; k += NT) {
    const cd hk = H[(long long)k * n + k - 1];
    bool small = cabs1(hk) <= smlnum;
    if (!small) {
      double tst = cabs1(H[(long long)(k - 1) * n + k - 1]) + cabs1(H[(long long)k * n + k]);
      if (tst == 0.0) {
        if (k - 2 >= 0) tst += cabs1(H[(long long)(k - 1) * n + k - 2]);
        if (k + 1 <= n - 1) tst += cabs1(H[(long long)(k + 1) * n + k]);
      }
      if (cabs1(hk) <= ULP * tst) {
        const double h1 = cabs1(hk), h2 = cabs1(H[(long long)(k - 1) * n + k]);
        const double ab = fmax(h1, h2), ba = fmin(h1, h2);
        const double g1 = cabs1(H[(long long)k * n + k]);
        const double g2 = cabs1(csub(H[(long long)(k - 1) * n + k - 1], H[(long long)k * n + k]));
        const double aa = fmax(g1, g2), bb = fmin(g1, g2);
        const double s = aa + ab;
        small = relaxed || ba * (ab / s) <= fmax(smlnum, ULP * (bb * (aa / s)));
      }
    }
    if (small && k > found) found = k;
  }
  // maximum over the workgroup: within each wave, then over the waves
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(found, off);
    found = (o > found) ? o : found;
  }
  if ((t & 63) == 0) s_k[t >> 6] = found;
  __syncthreads();
  int r = s_k[0];
  for (int w = 1; w < NT / 64; ++w) r = (s_k[w] > r) ? s_k[w] : r;
  __syncthreads();
  return r;
}

// A (rows x m, row stride lda) := A U, U m x m in LDS (ldu): a wave per group of four rows,
// lane j computes column j; the group's outputs are written after a barrier.
__device__ void rows_times_u(cd* A, int lda, int rows, int m, const cd* U, int ldu) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int R = 4;
  for (int q0 = 0; q0 < rows; q0 += R * (NT / 64)) {
    cd acc[R];
#pragma unroll
    for (int rr = 0; rr < R; ++rr) {
      acc[rr] = mkc(0.0, 0.0);
      const int q = q0 + wave * R + rr;
      if (q < rows && lane < m) {
        const cd* a = A + (long long)q * lda;
        for (int i = 0; i < m; ++i) acc[rr] = cadd(acc[rr], cmul(a[i], U[i * ldu + lane]));
      }
    }
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < R; ++rr) {
      const int q = q0 + wave * R + rr;
      if (q < rows && lane < m) A[(long long)q * lda + lane] = acc[rr];
    }
  }
  __syncthreads();
}

// B (m x cols, row stride ldb) := U^H B: lanes on 64 consecutive columns, wave w computes the
// output rows j = w, w + 8, ...; written after a barrier.
__device__ void uh_times_cols(cd* B, int ldb, int m, int cols, const cd* U, int ldu) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int NW = NT / 64;
  constexpr int JM = (64 + NW - 1) / NW;   // output rows per wave (m <= 64)
  for (int c0 = 0; c0 < cols; c0 += 64) {
    const int c = c0 + lane;
    cd acc[JM];
#pragma unroll
    for (int jj = 0; jj < JM; ++jj) acc[jj] = mkc(0.0, 0.0);
    if (c < cols) {
      for (int i = 0; i < m; ++i) {
        const cd b = B[(long long)i * ldb + c];
#pragma unroll
        for (int jj = 0; jj < JM; ++jj) {
          const int j = wave + NW * jj;
          if (j < m) acc[jj] = cadd(acc[jj], cmulc(U[i * ldu + j], b));
        }
      }
    }
    __syncthreads();
    if (c < cols) {
#pragma unroll
      for (int jj = 0; jj < JM; ++jj) {
        const int j = wave + NW * jj;
        if (j < m) B[(long long)j * ldb + c] = acc[jj];
      }
    }
  }
  __syncthreads();
}

static_assert(WB_HQR_NMIN <= 64 && WB_HQR_NMIN >= WB_HQR_M + 2, "small blocks (at most nmin rows) fit one wave");
__device__ __attribute__((noinline)) int small_block(cd* H, cd* Z, int zh, cd* wout, int n, int lo, int hi, cd* lds,
                                                     int* s_info, double dfl) {
  ASSUME_LDS(lds);
  ASSUME_LDS(s_info);
  const int t = threadIdx.x;
  const int m = hi - lo + 1;
  const int ld = (m % 2) ? m : m + 1;      // odd row stride: column accesses spread over banks
  cd* T = lds;
  cd* U = lds + m * ld;
  for (int e = t; e < m * m; e += NT) {
    const int r = e / m, c = e % m;
    T[r * ld + c] = H[(long long)(lo + r) * n + lo + c];
    U[r * ld + c] = mkc(r == c ? 1.0 : 0.0, 0.0);
  }
  if (t == 0) *s_info = 0;
  __syncthreads();
  if (t < 64) {
    const int info = wave_qr<true>(T, U, m, ld, n, wout + lo, dfl);
    if (t == 0) *s_info = info ? lo + info : 0;
  }
  __syncthreads();
  const int info = *s_info;
  for (int e = t; e < m * m; e += NT) {
    const int r = e / m, c = e % m;
    H[(long long)(lo + r) * n + lo + c] = T[r * ld + c];
  }
  rows_times_u(H + lo, n, lo, m, U, ld);                              // rows above the block
  uh_times_cols(H + (long long)lo * n + hi + 1, n, m, n - hi - 1, U, ld);   // right of the block
  if (zh) uh_times_cols(Z + (long long)lo * n, n, m, n, U, ld);        // Z^H rows lo..hi
  else rows_times_u(Z + lo, n, n, m, U, ld);                          // Z columns lo..hi
  return info;
}

// --------------------------------------------------------------------- aggressive early deflation
// The routines below run on one wave over a window of m <= 64 rows in LDS: T (row stride ld) and
// U its accumulated unitary. Lane j owns column j (or row j) of a step; LDS accesses of one wave
// stay in program order, so no barriers are needed.

// Swap the diagonal entries k, k+1 of the upper-triangular T (ztrexc's step): the rotation with
// Q [T(k,k+1); T(k+1,k+1) - T(k,k)] = [r; 0] applied to rows k, k+1 right of the pair, columns
// k, k+1 above it and the columns of U.
__device__ void wave_swap(cd* T, cd* U, int m, int ld, int k) {
  const int lane = threadIdx.x & 63;
  const cd t11 = T[k * ld + k], t22 = T[(k + 1) * ld + k + 1];
  cd rr;
  const Rot r = givens(T[k * ld + k + 1], csub(t22, t11), &rr);
  if (lane >= k + 2 && lane < m) rot_left(r, T[k * ld + lane], T[(k + 1) * ld + lane]);
  if (lane < k) rot_right(r, T[lane * ld + k], T[lane * ld + k + 1]);
  if (lane < m) rot_right(r, U[lane * ld + k], U[lane * ld + k + 1]);
  if (lane == 0) {
    T[k * ld + k] = t22;
    T[(k + 1) * ld + k + 1] = t11;
  }
}

// Householder reflector (zlarfg) of x = (alpha, x[1..len-1]): (I - tau v v^H)^H x = beta e1 with
// v[0] = 1; writes v (length len) to vout and returns tau (0 when x is already beta e1).
__device__ cd wave_reflector(cd* vout, const cd* x, int stride, int len, double* beta_out) {
  const int lane = threadIdx.x & 63;
  cd xi = (lane >= 1 && lane < len) ? x[lane * stride] : mkc(0.0, 0.0);
  double ss = cnorm2(xi);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) ss += __shfl_xor(ss, off);
  const cd alpha = x[0];
  if (ss == 0.0 && alpha.y == 0.0) {
    if (lane < len) vout[lane] = (lane == 0) ? mkc(1.0, 0.0) : mkc(0.0, 0.0);
    *beta_out = alpha.x;
    return mkc(0.0, 0.0);
  }
  const double beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + ss), alpha.x);
  const cd tau = mkc((beta - alpha.x) / beta, -alpha.y / beta);
  const cd scale = crecip(mkc(alpha.x - beta, alpha.y));
  if (lane < len) vout[lane] = (lane == 0) ? mkc(1.0, 0.0) : cmul(xi, scale);
  *beta_out = beta;
  return tau;
}

// rows r0..r0+len-1 of T, columns c0..m-1: T <- (I - conj(tau) v v^H) T (lane per column)
__device__ void wave_refl_left(cd* T, int ld, int m, int r0, int len, int c0, const cd* v, cd tau) {
  const int lane = threadIdx.x & 63;
  if (lane >= c0 && lane < m) {
    cd w = mkc(0.0, 0.0);
    for (int j = 0; j < len; ++j) w = cadd(w, cmulc(v[j], T[(r0 + j) * ld + lane]));
    const cd f = cmul(cconj(tau), w);
    for (int j = 0; j < len; ++j) T[(r0 + j) * ld + lane] = csub(T[(r0 + j) * ld + lane], cmul(v[j], f));
  }
}

// rows 0..rows-1 of A, columns c0..c0+len-1: A <- A (I - tau v v^H) (lane per row)
__device__ void wave_refl_right(cd* A, int ld, int rows, int c0, int len, const cd* v, cd tau) {
  const int lane = threadIdx.x & 63;
  if (lane < rows) {
    cd w = mkc(0.0, 0.0);
    for (int j = 0; j < len; ++j) w = cadd(w, cmul(A[lane * ld + c0 + j], v[j]));
    const cd f = cmul(tau, w);
    for (int j = 0; j < len; ++j) A[lane * ld + c0 + j] = csub(A[lane * ld + c0 + j], cmul(f, cconj(v[j])));
  }
}

// The deflation of one window (one wave): T, U hold the window's Schur form and unitary, s its
// coupling H(kw, kw-1). Deflates from the bottom every eigenvalue with |s| |U(0, j)| <=
// max(smlnum, ulp |T(j, j)|) and moves the others to the top; returns ns, the undeflated count,
// with their eigenvalues copied to eig[0..ns-1]. When some deflated (ns < m) and s != 0, the
// undeflated part is returned to Hessenberg form: a reflector folds the spike into entry 0, then
// Householder reduction of T(0:ns, 0:ns), every transformation also applied to U.
__device__ int wave_aed(cd* T, cd* U, int m, int ld, cd s, double smlnum, cd* eig, cd* v) {
  const int lane = threadIdx.x & 63;
  int ns = m, ilst = 0;
  const double as = cabs1(s);
  while (ilst < ns) {
    double foo = cabs1(T[(ns - 1) * ld + ns - 1]);
    if (foo == 0.0) foo = as;
    if (as * cabs1(U[ns - 1]) <= fmax(smlnum, ULP * foo)) {
      --ns;                                   // deflatable: stays at the bottom
    } else {
      for (int k = ns - 2; k >= ilst; --k) wave_swap(T, U, m, ld, k);
      ++ilst;                                 // undeflatable: moved to the top
    }
  }
  if (lane < ns) eig[lane] = T[lane * ld + lane];
  if (ns == m || as == 0.0 || ns <= 1) return ns;
  // fold the spike s conj(U[0][0..ns-1]) into its first entry
  double beta;
  {
    const cd x = (lane < ns) ? cconj(U[lane]) : mkc(0.0, 0.0);
    if (lane < ns) v[64 + lane] = x;           // scratch copy of the spike direction
  }
  const cd tau0 = wave_reflector(v, v + 64, 1, ns, &beta);
  wave_refl_left(T, ld, m, 0, ns, 0, v, tau0);
  wave_refl_right(T, ld, ns, 0, ns, v, tau0);
  wave_refl_right(U, ld, m, 0, ns, v, tau0);
  // Hessenberg reduction of T(0:ns, 0:ns) (columns ns.. see the left reflectors only)
  for (int i = 0; i + 2 < ns; ++i) {
    const int len = ns - i - 1;
    const cd tau = wave_reflector(v, T + (i + 1) * ld + i, ld, len, &beta);
    wave_refl_left(T, ld, m, i + 1, len, i + 1, v, tau);
    if (lane == 0) T[(i + 1) * ld + i] = mkc(beta, 0.0);
    if (lane >= i + 2 && lane < ns) T[lane * ld + i] = mkc(0.0, 0.0);
    wave_refl_right(T, ld, ns, i + 1, len, v, tau);
    wave_refl_right(U, ld, m, i + 1, len, v, tau);
  }
  return ns;
}

// Aggressive early deflation on the trailing jw rows of the active block [lo, hi] (all threads;
// the window in the dynamic LDS). Returns the number of eigenvalues deflated (their rows end
// upper triangular, split off by zero subdiagonals), or -1 when the window's QR did not converge
// (H untouched); the undeflated eigenvalues (ns of them) are written to eig (LDS) and *ns_out.
__device__ __attribute__((noinline)) int aed_window(cd* H, cd* Z, int zh, int n, int lo, int hi, int jw,
                                                    cd* lds, cd* eig, int* s_int, double dfl) {
  ASSUME_LDS(lds);
  ASSUME_LDS(eig);
  ASSUME_LDS(s_int);
  const int t = threadIdx.x;
  const int kw = hi - jw + 1;
  const int ld = (jw % 2) ? jw : jw + 1;
  cd* T = lds;
  cd* U = lds + jw * ld;
  cd* v = U + jw * ld;                          // 128 entries of reflector scratch
  cd* wtmp = v + 128;                           // the window's eigenvalues from its QR
  for (int e = t; e < jw * jw; e += NT) {
    const int r = e / jw, c = e % jw;
    T[r * ld + c] = (r > c + 1) ? mkc(0.0, 0.0) : H[(long long)(kw + r) * n + kw + c];
    U[r * ld + c] = mkc(r == c ? 1.0 : 0.0, 0.0);
  }
  const cd s = (kw > lo) ? H[(long long)kw * n + kw - 1] : mkc(0.0, 0.0);
  __syncthreads();
  if (t < 64) {
    const double smlnum = SAFMIN * ((double)n / ULP);
    const int bad = wave_qr<true>(T, U, jw, ld, n, wtmp, dfl);
    int ns = -1;
    if (!bad) ns = wave_aed(T, U, jw, ld, s, smlnum, eig, v);
    if (t == 0) s_int[0] = ns;
  }
  __syncthreads();
  const int ns = s_int[0];
  if (ns < 0) return -1;
  const int nd = jw - ns;
  if (nd == 0) return 0;                        // nothing deflated: H is left as it was
  // the window back into H: T (Hessenberg over the undeflated rows, triangular below), the spike
  for (int e = t; e < jw * jw; e += NT) {
    const int r = e / jw, c = e % jw;
    H[(long long)(kw + r) * n + kw + c] = (r > c + 1) ? mkc(0.0, 0.0) : T[r * ld + c];
  }
  if (t == 0 && kw > lo) H[(long long)kw * n + kw - 1] = cmul(s, cconj(U[0]));
  __syncthreads();
  rows_times_u(H + kw, n, kw, jw, U, ld);                                  // rows above
  uh_times_cols(H + (long long)kw * n + hi + 1, n, jw, n - hi - 1, U, ld);  // right of it
  if (zh) uh_times_cols(Z + (long long)kw * n, n, jw, n, U, ld);
  else rows_times_u(Z + kw, n, n, jw, U, ld);
  return nd;
}

// Compiler fence on values: the asm "reads and writes" x, so everything computing x is
// emitted before it and everything using x after it (no instruction is generated).
__device__ inline void pin(cd& x) { asm volatile("" : "+v"(x.x), "+v"(x.y)); }

// Strip element access through a buffer resource over one matrix: element i of strip k sits at
// byte offset soff0 + k sstride + i estride. Offsets outside the valid range are replaced by
// BUF_OOB, beyond the resource: such loads return zero and such stores are dropped by the
// hardware, so the band needs no branches around its loads and stores (and the waits for its
// loads do not depend on which stores were issued).
typedef unsigned v4u __attribute__((ext_vector_type(4)));
constexpr unsigned BUF_OOB = 0x80000000u;
__device__ inline __amdgpu_buffer_rsrc_t strip_rsrc(const cd* base, int nbytes) {
  const unsigned long long a = (unsigned long long)base;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
  const int nrec = __builtin_amdgcn_readfirstlane(nbytes);
  return __builtin_amdgcn_make_buffer_rsrc((void*)(((unsigned long long)hi << 32) | lo), (short)0, nrec,
                                           0x00020000);
}
__device__ inline cd buf_load(__amdgpu_buffer_rsrc_t r, unsigned off) {
  union { v4u u; double d[2]; } c;
  c.u = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
  return mkc(c.d[0], c.d[1]);
}
__device__ inline void buf_store(__amdgpu_buffer_rsrc_t r, unsigned off, cd v) {
  union { v4u u; double d[2]; } c;
  c.d[0] = v.x;
  c.d[1] = v.y;
  __builtin_amdgcn_raw_buffer_store_b128(c.u, r, off, 0, 0);
}


// Apply the steps [s0, s1) of a group of G consecutive bulges' rotations to one strip through
// a B-register band. left: row operations on a column strip (rows), else column operations on
// a row strip. get(i)/put(i) address strip element i (absolute index); [vlo, vhi) is the valid
// range. `lo` is the group's entry position (lo - 3 g0 for the group starting at bulge g0) and
// `rots` points at the group's first bulge in rows of rot_stride rotations per step. Bulge j
// only conflicts with earlier rotations of bulge j - 1, so whole groups may be applied one after
// another in bulge order: the result equals the step-by-step order of the chase.
template <int M, bool LEFT, int S>
__device__ __attribute__((noinline)) void band_apply(int lo, int s0, int s1, const Rot* rots, int rot_stride,
                                                     int vlo, int vhi, const cd* base, int nbytes, int estride,
                                                     int soff0, int sstride) {
  ASSUME_LDS(rots);
  // Steps run in groups of U = 3 with compile-time band slots (bulge j of step s + u acts on
  // slots 3 (M - 1 - j) + u, + u + 1), so the band shifts by three registers once per group
  // rather than by one every step, and the group's three entering elements are loaded one
  // group ahead, overlapping their memory latency with the rotations of the current group.
  // Within a step the bulges are independent; they are applied BC at a time, the rotations
  // of the next BC loaded while the current ones are applied, and pinned so that no more
  // than two chunks of rotations are held in registers.
  constexpr int B = 3 * (M - 1) + 2;
  constexpr int U = 3;
  constexpr int BC = (M % 4 == 0) ? 4 : M;
  constexpr int CPS = M / BC;          // chunks per step
  constexpr int NCH = U * CPS;         // chunks per group
  const __amdgpu_buffer_rsrc_t rsrc = strip_rsrc(base, nbytes);
  cd band[S][B + U - 1];
  cd pre[S][U];
  const int base0 = lo + s0 - 3 * (M - 1);
  const int top_last = base0 + (s1 - 1 - s0) + B - 1;   // last element the pass reads
  const int hi_read = (top_last + 1 < vhi) ? top_last + 1 : vhi;
  auto off = [&](int k, int idx, int lim) -> unsigned {
    return (idx >= vlo && idx < lim) ? (unsigned)(soff0 + k * sstride + idx * estride) : BUF_OOB;
  };
  auto rotate = [&](const Rot& r, cd& x0, cd& x1) {
    if (LEFT) rot_left(r, x0, x1);
    else rot_right(r, x0, x1);
  };
#pragma unroll
  for (int i = 0; i < B - 1; ++i)
#pragma unroll
    for (int k = 0; k < S; ++k) band[k][i] = buf_load(rsrc, off(k, base0 + i, hi_read));
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int k = 0; k < S; ++k) pre[k][u] = buf_load(rsrc, off(k, base0 + B - 1 + u, hi_read));
  int s = s0;
  int dep = 0;   // always 0; threaded through the pins to order the rotation loads
  const int ng = (s1 - s0) / U;
  for (int g = 0; g < ng; ++g, s += U) {
    const int base = lo + s - 3 * (M - 1);
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int k = 0; k < S; ++k) {
        band[k][B - 1 + u] = pre[k][u];
        pre[k][u] = buf_load(rsrc, off(k, base + U + B - 1 + u, hi_read));
      }
    const Rot* rs = rots + (long long)(s - s0) * rot_stride;
    Rot cur[BC], nxt[BC];
#pragma unroll
    for (int jj = 0; jj < BC; ++jj) cur[jj] = rs[jj];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      const int u = c / CPS, j0 = (c % CPS) * BC;
      if (c + 1 < NCH) {
        const int un = (c + 1) / CPS, jn = ((c + 1) % CPS) * BC;
#pragma unroll
        for (int jj = 0; jj < BC; ++jj) nxt[jj] = rs[(long long)un * rot_stride + jn + jj + dep];
      }
#pragma unroll
      for (int jj = 0; jj < BC; ++jj) {
        const int slot = 3 * (M - 1 - (j0 + jj)) + u;
#pragma unroll
        for (int k = 0; k < S; ++k) rotate(cur[jj], band[k][slot], band[k][slot + 1]);
      }
#pragma unroll
      for (int jj = 0; jj < BC; ++jj) {
        const int slot = 3 * (M - 1 - (j0 + jj)) + u;
#pragma unroll
        for (int k = 0; k < S; ++k) { pin(band[k][slot]); pin(band[k][slot + 1]); }
      }
      asm volatile("" : "+v"(dep));
      if (c % CPS == CPS - 1) {
        // slot u is final once step u is done
#pragma unroll
        for (int k = 0; k < S; ++k) buf_store(rsrc, off(k, base + u, vhi), band[k][u]);
      }
#pragma unroll
      for (int jj = 0; jj < BC; ++jj) cur[jj] = nxt[jj];
    }
#pragma unroll
    for (int i = 0; i < B - 1; ++i)
#pragma unroll
      for (int k = 0; k < S; ++k) band[k][i] = band[k][i + U];
  }
  // the remaining (at most U - 1) steps one at a time; their entering elements are in pre
#pragma unroll
  for (int u = 0; u < U - 1; ++u) {
    if (s >= s1) break;
    const int base = lo + s - 3 * (M - 1);
#pragma unroll
    for (int k = 0; k < S; ++k) band[k][B - 1] = pre[k][u];
    const Rot* rs = rots + (long long)(s - s0) * rot_stride;
#pragma unroll
    for (int j = 0; j < M; ++j) {
      const Rot r = rs[j];
#pragma unroll
      for (int k = 0; k < S; ++k) rotate(r, band[k][3 * (M - 1 - j)], band[k][3 * (M - 1 - j) + 1]);
    }
#pragma unroll
    for (int k = 0; k < S; ++k) buf_store(rsrc, off(k, base, vhi), band[k][0]);
#pragma unroll
    for (int i = 0; i < B - 1; ++i)
#pragma unroll
      for (int k = 0; k < S; ++k) band[k][i] = band[k][i + 1];
    ++s;
  }
  const int base_end = lo + s1 - 3 * (M - 1);
#pragma unroll
  for (int i = 0; i < B - 1; ++i)
#pragma unroll
    for (int k = 0; k < S; ++k) buf_store(rsrc, off(k, base_end + i, vhi), band[k][i]);
}

// Two bulge groups of 8 on one strip, one group per half-wave: lane l (group 0, bulges 0..7)
// and lane l + 32 (group 1, bulges 8..15) hold the two bands of strip l, in lockstep. Group 1's
// band trails group 0's by 24 positions, so the element that leaves group 0's band at step s
// enters group 1's at step s + 2: it crosses to the other half-wave (a lane swap) and waits in
// a two-element queue. The result equals the two groups' passes one after the other (group 1
// only meets elements group 0 no longer touches), with each strip element read and written
// once per pass for all 16 bulges. `active` is false for lanes past the last strip (their
// loads and stores go to BUF_OOB).
// Value select of two complex numbers (a conditional on the values, not on their addresses, so
// the register arrays they come from stay in registers).
__device__ inline cd csel(bool c, cd a, cd b) { return mkc(c ? a.x : b.x, c ? a.y : b.y); }

//
// STAGED: the pass spans more steps than LDS holds. `rots` is then an LDS stage of `chunk`
// steps (a multiple of U) that the workgroup refills from `rglob` (all steps, M rotations
// each, from step 0) whenever the pass crosses a chunk boundary, so each strip element is still
// read and written once for all steps [s0, s1). Every thread of the workgroup must make the
// same calls (the refills synchronize the workgroup).
template <bool LEFT, bool STAGED>
__device__ __attribute__((noinline)) void band_apply_pair(int lo, int s0, int s1, Rot* rots, int rot_stride,
                                                          int vlo, int vhi, const cd* base, int nbytes,
                                                          int estride, int soff0, bool active,
                                                          const Rot* rglob, int chunk) {
  ASSUME_LDS(rots);
  ASSUME_GLOBAL(rglob);
  constexpr int G = 8;                    // bulges per group
  constexpr int B = 3 * (G - 1) + 2;
  constexpr int U = 3;
  constexpr int BC = WB_HQR_BC;
  constexpr int CPS = G / BC;
  constexpr int NCH = U * CPS;
  const int lane = threadIdx.x & 63;
  const bool g1 = lane >= 32;
  const int lo_g = g1 ? lo - 3 * G : lo;
  const Rot* rots_g = rots + (g1 ? G : 0);
  const __amdgpu_buffer_rsrc_t rsrc = strip_rsrc(base, nbytes);
  int cbase = s0;   // first step held by the stage
  auto restage = [&](int sb) {
    __syncthreads();
    const int ns = (s1 - sb < chunk) ? s1 - sb : chunk;
    for (int e = threadIdx.x; e < ns * rot_stride; e += NT) rots[e] = rglob[(long long)sb * rot_stride + e];
    __syncthreads();
    cbase = sb;
  };
  if constexpr (STAGED) {
    restage(s0);
    // a wave with no strip in range only takes part in the refills
    if (__ballot(active) == 0) {
      for (int sb = s0 + chunk; sb < s1; sb += chunk) restage(sb);
      return;
    }
  }
  cd band[B + U - 1];
  cd pre[U];
  cd carried[2];
  const int base0 = lo_g + s0 - 3 * (G - 1);
  const int top_last = base0 + (s1 - 1 - s0) + B - 1;   // last element group 0 reads
  const int hi_read = (top_last + 1 < vhi) ? top_last + 1 : vhi;
  auto off = [&](int idx, int lim) -> unsigned {
    return (active && idx >= vlo && idx < lim) ? (unsigned)(soff0 + idx * estride) : BUF_OOB;
  };
  auto rotate = [&](const Rot& r, cd& x0, cd& x1) {
    if (LEFT) rot_left(r, x0, x1);
    else rot_right(r, x0, x1);
  };
  auto swap_halves = [&](cd x) {
    cd y;
    y.x = __shfl_xor(x.x, 32);
    y.y = __shfl_xor(x.y, 32);
    return y;
  };
  // group 0 reads its band and the entering elements from memory; group 1 its band and the two
  // elements in flight between the groups at the start of the pass
#pragma unroll
  for (int i = 0; i < B - 1; ++i) band[i] = buf_load(rsrc, off(base0 + i, g1 ? vhi : hi_read));
#pragma unroll
  for (int u = 0; u < U; ++u) pre[u] = buf_load(rsrc, g1 ? BUF_OOB : off(base0 + B - 1 + u, hi_read));
#pragma unroll
  for (int q = 0; q < 2; ++q) carried[q] = buf_load(rsrc, g1 ? off(base0 + B - 1 + q, vhi) : BUF_OOB);
#pragma unroll
  for (int i = 0; i < B - 1; ++i) pin(band[i]);
#pragma unroll
  for (int u = 0; u < U; ++u) pin(pre[u]);
  pin(carried[0]);
  pin(carried[1]);
  int s = s0;
  int dep = 0;
  const int ng = (s1 - s0) / U;
  for (int g = 0; g < ng; ++g, s += U) {
    if constexpr (STAGED) {
      if (s - cbase >= chunk) restage(s);
    }
    const int base = lo_g + s - 3 * (G - 1);
    band[B - 1] = csel(g1, carried[0], pre[0]);
    band[B] = csel(g1, carried[1], pre[1]);
    band[B + 1] = pre[2];                      // group 1: replaced after step 0
#pragma unroll
    for (int u = 0; u < U; ++u) pre[u] = buf_load(rsrc, g1 ? BUF_OOB : off(base + U + B - 1 + u, hi_read));
    const Rot* rs = rots_g + (long long)(s - cbase) * rot_stride;
    Rot cur[BC], nxt[BC];
#pragma unroll
    for (int jj = 0; jj < BC; ++jj) cur[jj] = rs[jj];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      const int u = c / CPS, j0 = (c % CPS) * BC;
      if (c + 1 < NCH) {
        const int un = (c + 1) / CPS, jn = ((c + 1) % CPS) * BC;
#pragma unroll
        for (int jj = 0; jj < BC; ++jj) nxt[jj] = rs[(long long)un * rot_stride + jn + jj + dep];
      }
#pragma unroll
      for (int jj = 0; jj < BC; ++jj) {
        const int slot = 3 * (G - 1 - (j0 + jj)) + u;
        rotate(cur[jj], band[slot], band[slot + 1]);
      }
#pragma unroll
      for (int jj = 0; jj < BC; ++jj) {
        const int slot = 3 * (G - 1 - (j0 + jj)) + u;
        pin(band[slot]);
        pin(band[slot + 1]);
      }
      asm volatile("" : "+v"(dep));
      if (c % CPS == CPS - 1) {
        // slot u is final: group 0 hands it to group 1, group 1 writes it back
        const cd x = swap_halves(band[u]);
        if (u == 0) band[B + 1] = csel(g1, x, band[B + 1]);
        if (u == 1) carried[0] = x;
        if (u == 2) carried[1] = x;
        buf_store(rsrc, g1 ? off(base + u, vhi) : BUF_OOB, band[u]);
      }
#pragma unroll
      for (int jj = 0; jj < BC; ++jj) cur[jj] = nxt[jj];
    }
#pragma unroll
    for (int i = 0; i < B - 1; ++i) band[i] = band[i + U];
  }
  // the remaining (at most U - 1) steps one at a time
#pragma unroll
  for (int u = 0; u < U - 1; ++u) {
    if (s >= s1) break;
    if constexpr (STAGED) {
      if (s - cbase >= chunk) restage(s);
    }
    const int base = lo_g + s - 3 * (G - 1);
    band[B - 1] = csel(g1, carried[0], pre[u]);
    const Rot* rs = rots_g + (long long)(s - cbase) * rot_stride;
#pragma unroll
    for (int j = 0; j < G; ++j) {
      const Rot r = rs[j];
      rotate(r, band[3 * (G - 1 - j)], band[3 * (G - 1 - j) + 1]);
    }
    const cd x = swap_halves(band[0]);
    carried[0] = carried[1];
    carried[1] = x;
    buf_store(rsrc, g1 ? off(base, vhi) : BUF_OOB, band[0]);
#pragma unroll
    for (int i = 0; i < B - 1; ++i) band[i] = band[i + 1];
    ++s;
  }
  // flush: each group its band; group 1 also the two elements in flight
  const int base_end = lo_g + s1 - 3 * (G - 1);
#pragma unroll
  for (int i = 0; i < B - 1; ++i) buf_store(rsrc, off(base_end + i, vhi), band[i]);
#pragma unroll
  for (int q = 0; q < 2; ++q) buf_store(rsrc, g1 ? off(base_end + B - 1 + q, vhi) : BUF_OOB, carried[q]);
}

// Reflector of bulge j at chase position k (k == lo: the bulge enters with shift sh).
// Writes r / 0 into the window for k > lo. Inactive positions give the identity.
__device__ inline Rot bulge_rotation(cd* Hw, int ldw, int a, int lo, int hi, int k, cd sh) {
  Rot r;
  r.c = 1.0;
  r.sx = 0.0;
  r.sy = 0.0;
  if (k < lo || k > hi - 1) return r;
  cd x1, x2;
  if (k == lo) {
    x1 = csub(Hw[(k - a) * ldw + (k - a)], sh);
    x2 = Hw[(k + 1 - a) * ldw + (k - a)];
    const double sc = cabs1(x1) + cabs1(x2);
    if (sc != 0.0) { x1 = cscale(x1, 1.0 / sc); x2 = cscale(x2, 1.0 / sc); }
  } else {
    x1 = Hw[(k - a) * ldw + (k - 1 - a)];
    x2 = Hw[(k + 1 - a) * ldw + (k - 1 - a)];
  }
  cd rr;
  r = givens(x1, x2, &rr);
  if (k > lo) {
    Hw[(k - a) * ldw + (k - 1 - a)] = rr;
    Hw[(k + 1 - a) * ldw + (k - 1 - a)] = mkc(0.0, 0.0);
  }
  return r;
}

// One multishift sweep over the active block [lo, hi] with shifts sh[0..M).
// Per chase step two barriers: row operations of all bulges, then their column operations
// on waves 1.., while lane j of wave 0 finishes rows k+1, k+2 of bulge j's column operation in
// registers and forms bulge j's rotation for the next step from them (the only inputs it
// needs), the step's serial chain. The
// current and next steps' rotations live in an LDS ring; every step's rotations are also
// logged to rot_all for the deferred updates outside the window and of Z.
// Deferred updates of one window's steps [s0, s1) (rotations rw in LDS, M per step) on the
// strips [first, end) of a matrix: M = 8 one band per thread, M = 16 a band pair per two lanes.
template <int M, bool LEFT>
__device__ inline void deferred(int lo, int s0, int s1, Rot* rw, int vlo, int vhi, const cd* base,
                                int nbytes, int estride, int sstride, int first, int end) {
  const int t = threadIdx.x;
  if constexpr (M == 16) {
    for (int xb = first + (t >> 6) * 32; xb < end; xb += NT / 2) {
      const int x = xb + (t & 31);
      band_apply_pair<LEFT, false>(lo, s0, s1, rw, M, vlo, vhi, base, nbytes, estride, sstride * x, x < end,
                                   nullptr, 0);
    }
  } else {
    static_assert(M == 8, "8 or 16 bulges per sweep");
    for (int x = first + t; x < end; x += NT)
      band_apply<M, LEFT, 1>(lo, s0, s1, rw, M, vlo, vhi, base, nbytes, estride, sstride * x, 0);
  }
}

// Deferred updates of a whole sweep's steps [0, nsteps) on the strips [first, end), the
// rotations read from rot_all (step-major, M per step) through the LDS stage `rz` of `cap`
// steps. M = 16: one pass per strip (the stage is refilled inside the pass, every thread takes
// part); M = 8: one pass per stage-full of steps.
template <int M, bool LEFT>
__device__ inline void deferred_sweep(int lo, int nsteps, const Rot* rot_all, Rot* rz, int cap, int vlo,
                                      int vhi, const cd* base, int nbytes, int estride, int sstride,
                                      int first, int end) {
  const int t = threadIdx.x;
  if (end <= first) return;
  if constexpr (M == 16) {
    const int chunk = (cap / 3) * 3;
    const int iters = (end - first + NT / 2 - 1) / (NT / 2);
    for (int it = 0; it < iters; ++it) {
      const int x = first + it * (NT / 2) + (t >> 6) * 32 + (t & 31);
      band_apply_pair<LEFT, true>(lo, 0, nsteps, rz, M, vlo, vhi, base, nbytes, estride, sstride * x, x < end,
                                  rot_all, chunk);
    }
    __syncthreads();
  } else {
    for (int c0 = 0; c0 < nsteps; c0 += cap) {
      const int c1 = (c0 + cap < nsteps) ? c0 + cap : nsteps;
      __syncthreads();
      for (int e = t; e < (c1 - c0) * M; e += NT) rz[e] = rot_all[(long long)c0 * M + e];
      __syncthreads();
      deferred<M, LEFT>(lo, c0, c1, rz, vlo, vhi, base, nbytes, estride, sstride, first, end);
    }
    __syncthreads();
  }
}

template <int M>
__device__ void multishift_sweep(cd* H, cd* Z, int zh, int n, int lo, int hi, const cd* sh, int W,
                                 cd* Hw, Rot (*s_rot)[M], Rot* rot_all) {
  const int nbytes = 16 * n * n;   // one matrix (H or Z): byte offsets fit in 32 bits
  ASSUME_LDS(Hw);
  ASSUME_LDS(s_rot);
  ASSUME_LDS(sh);
  ASSUME_GLOBAL(H);
  ASSUME_GLOBAL(Z);
  ASSUME_GLOBAL(rot_all);
  const int t = threadIdx.x;
  const int T = (hi - 1 - lo) + 3 * (M - 1);       // last step
  const int ldw = W + 1;
  int s_begin = 0;
  bool first = true;
  while (s_begin <= T) {
    int kmin = lo + s_begin - 3 * (M - 1);
    if (kmin < lo) kmin = lo;
    const int a = (kmin == lo) ? lo : kmin - 1;
    const int wend = (a + W < hi + 1) ? a + W : hi + 1;
    const int ww = wend - a;
    int s_end;
    if (wend == hi + 1) s_end = T + 1;
    else s_end = wend - 3 - lo + 1;
    if (s_end > T + 1) s_end = T + 1;
    for (int e = t; e < ww * ww; e += NT) {
      const int r = e / ww, c = e % ww;
      Hw[r * ldw + c] = H[(long long)(a + r) * n + a + c];
    }
    __syncthreads();
    if (first) {
      // rotations of step 0 (only bulge 0 can be active)
      if (t < M) {
        const Rot r = bulge_rotation(Hw, ldw, a, lo, hi, lo + s_begin - 3 * t, sh[t]);
        s_rot[s_begin & 1][t] = r;
        rot_all[(long long)s_begin * M + t] = r;
      }
      first = false;
      __syncthreads();
    }
    for (int s = s_begin; s < s_end; ++s) {
      const Rot* rs = s_rot[s & 1];
      // row operations: bulge j (wave j mod the wave count) on rows k, k+1, columns k..wend-1
      for (int j = t >> 6; j < M; j += NT / 64) {
        const int k = lo + s - 3 * j;
        if (k < lo || k > hi - 1) continue;
        const Rot r = rs[j];
        for (int c = k + (t & 63); c < wend; c += 64)
          rot_left(r, Hw[(k - a) * ldw + (c - a)], Hw[(k + 1 - a) * ldw + (c - a)]);
      }
      __syncthreads();
      // column operations: bulge j on columns k, k+1, rows a..k (waves 1.. ); rows k+1, k+2
      // and the next rotation by lane j of wave 0, the serial part of the step
      if (t >= 64) {
        for (int j = (t >> 6) - 1; j < M; j += NT / 64 - 1) {
          const int k = lo + s - 3 * j;
          if (k < lo || k > hi - 1) continue;
          const Rot r = rs[j];
          for (int q = a + (t & 63); q <= k; q += 64)
            rot_right(r, Hw[(q - a) * ldw + (k - a)], Hw[(q - a) * ldw + (k + 1 - a)]);
        }
      } else if (t < M) {
        const int j = t;
        const int k = lo + s - 3 * j;
        Rot nr;
        nr.c = 1.0;
        nr.sx = 0.0;
        nr.sy = 0.0;
        if (k >= lo && k <= hi - 1) {
          // rows k+1 and k+2 of columns k, k+1 stay in registers: rotated, then the entries
          // (k+1, k) and (k+2, k) give the next rotation (position k+1) directly
          const Rot r = rs[j];
          cd* p1 = &Hw[(k + 1 - a) * ldw + (k - a)];
          cd a10 = p1[0], a11 = p1[1];
          rot_right(r, a10, a11);
          p1[1] = a11;
          if (k + 2 <= hi) {
            cd* p2 = p1 + ldw;
            cd a20 = p2[0], a21 = p2[1];
            rot_right(r, a20, a21);
            p2[1] = a21;
            if (s + 1 <= T) {
              cd rr;
              nr = givens(a10, a20, &rr);
              a10 = rr;
              a20 = mkc(0.0, 0.0);
            }
            p2[0] = a20;
          }
          p1[0] = a10;
        } else if (s + 1 <= T) {
          nr = bulge_rotation(Hw, ldw, a, lo, hi, k + 1, sh[j]);   // entering or inactive
        }
        if (s + 1 <= T) {
          s_rot[(s + 1) & 1][j] = nr;
          rot_all[(long long)(s + 1) * M + j] = nr;
        }
      }
      __syncthreads();
    }
    // store window
    for (int e = t; e < ww * ww; e += NT) {
      const int r = e / ww, c = e % ww;
      H[(long long)(a + r) * n + a + c] = Hw[r * ldw + c];
    }
    // the window's rotations, from rot_all into the LDS the window no longer needs
    __syncthreads();
    Rot* rw = (Rot*)Hw;
    for (int e = t; e < (s_end - s_begin) * M; e += NT) rw[e] = rot_all[(long long)s_begin * M + e];
    __syncthreads();
    // deferred row operations on the active columns right of the window (rows a..wend-1), then
    // column operations on the active rows above it (columns a..wend-1): later windows read them
    deferred<M, true>(lo, s_begin, s_end, rw, a, wend, H, nbytes, 16 * n, 16, wend, hi + 1);
    deferred<M, false>(lo, s_begin, s_end, rw, a, wend, H, nbytes, 16, 16 * n, lo, a);
    __syncthreads();
    s_begin = s_end;
  }
  // The strips no window reads get the whole sweep in one pass each: the columns right of the
  // active block (row operations on rows lo..hi), the rows above it (column operations on
  // columns lo..hi) and the Schur vectors (every row of Z: column operations; Z^H: row
  // operations on its column strips)
  Rot* rz = (Rot*)Hw;
  const int cap = (int)((sizeof(cd) * (size_t)W * (W + 1)) / (sizeof(Rot) * M));
  deferred_sweep<M, true>(lo, T + 1, rot_all, rz, cap, lo, hi + 1, H, nbytes, 16 * n, 16, hi + 1, n);
  deferred_sweep<M, false>(lo, T + 1, rot_all, rz, cap, lo, hi + 1, H, nbytes, 16, 16 * n, 0, lo);
  if (zh) deferred_sweep<M, true>(lo, T + 1, rot_all, rz, cap, lo, hi + 1, Z, nbytes, 16 * n, 16, 0, n);
  else deferred_sweep<M, false>(lo, T + 1, rot_all, rz, cap, lo, hi + 1, Z, nbytes, 16, 16 * n, 0, n);
  __syncthreads();
}

template <int M>
__global__ __launch_bounds__(NT) void k_hqr_ms(int n, cd* Hb, cd* Zb, int zh, cd* wb, int* info,
                                               Rot* rws, int W, int nmin) {
  const long long b = blockIdx.x;
  cd* H = (cd*)(Hb + b * (long long)n * n);
  cd* Z = (cd*)(Zb + b * (long long)n * n);
  cd* wout = wb + b * (long long)n;
  Rot* rot_all = (Rot*)(rws + b * (long long)(n + 3 * M) * M);
  extern __shared__ unsigned char smem[];
  cd* Hw = (cd*)smem;
  __shared__ Rot s_rot[2][M];
  __shared__ int s_k[NT / 64];
  __shared__ cd s_sh[M];
  __shared__ cd s_small[M * M];
  const int t = threadIdx.x;
  int hi = n - 1;
  int lo = 0;
  int stall = 0;
  int sweeps = 0;
  const int max_sweeps = 40 * (n / M + 10);
  // the normwise deflation floor ulp ||H||_F (a wave per row of the Hessenberg part)
  double dfl;
  {
    __shared__ double s_nrm[NT / 64];
    const int lane = t & 63, wave = t >> 6;
    double acc = 0.0;
    for (int r = wave; r < n; r += NT / 64) {
      const cd* row = H + (long long)r * n;
      for (int c = (r > 0 ? r - 1 : 0) + lane; c < n; c += 64) acc += cnorm2(row[c]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) s_nrm[wave] = acc;
    __syncthreads();
    double sum = 0.0;
    for (int w = 0; w < NT / 64; ++w) sum += s_nrm[w];
    dfl = ULP * sqrt(sum);
  }
  while (hi >= 0) {
    lo = find_split(H, n, 0, hi, s_k, stall >= 3);
    if (lo > 0 && t == 0) H[(long long)lo * n + lo - 1] = mkc(0.0, 0.0);
    __syncthreads();
    if (lo == hi) {
      if (t == 0) wout[hi] = H[(long long)hi * n + hi];
      --hi;
      stall = 0;
      __syncthreads();
      continue;
    }
    if (hi - lo + 1 <= nmin) {
      const int r = small_block(H, Z, zh, wout, n, lo, hi, Hw, &s_k[0], dfl);
      if (r != 0) {
        if (t == 0) info[b] = r;
        return;
      }
      hi = lo - 1;
      stall = 0;
      __syncthreads();
      continue;
    }
    int hs = hi;                 // last row of the sweep
    bool have = false;           // shifts from the deflation window
#if WB_HQR_AED
    {
      const int jw = (WB_HQR_AED < hi - lo + 1) ? WB_HQR_AED : hi - lo + 1;
      const int nd = aed_window(H, Z, zh, n, lo, hi, jw, Hw, s_small, s_k, dfl);
      if (nd > 0) {
        stall = 0;
        hs = hi - nd;
        // a large deflation (LAPACK's nibble, 14% of the block) or a block left small: split
        // off what deflated and look again before sweeping
        if (100 * nd > 14 * (hi - lo + 1) || hs - lo + 1 <= nmin) continue;
      }
      if (nd >= 0 && jw - nd >= M) {
        if (t < M) s_sh[t] = s_small[jw - nd - M + t];
        have = true;
      }
    }
#endif
    if (++sweeps > max_sweeps) {
      if (t == 0) info[b] = hi + 1;
      return;
    }
    // shifts: the deflation window's undeflated eigenvalues, else those of the trailing M x M
    // block (one wave); exceptional shifts when stalled or when that small iteration fails
    ++stall;
    if (!have || stall % 6 == 0) {
      const int b0 = hs - M + 1;
      __syncthreads();
      for (int e = t; e < M * M; e += NT) {
        const int r = e / M, c = e % M;
        s_small[e] = (r > c + 1) ? mkc(0.0, 0.0) : H[(long long)(b0 + r) * n + b0 + c];
      }
      __syncthreads();
      if (t < 64) {
        const int bad = wave_qr<false>(s_small, nullptr, M, M, M, s_sh, dfl);
        if (t == 0 && (bad || stall % 6 == 0)) {
          const double ex = 0.75 * cabs1(H[(long long)hs * n + hs - 1]);
          for (int j = 0; j < M; ++j)
            s_sh[j] = cadd(H[(long long)(hs - (j % 2)) * n + hs - (j % 2)], mkc(ex * (1.0 + 0.1 * j), 0.0));
        }
      }
    }
    __syncthreads();
    multishift_sweep<M>(H, Z, zh, n, lo, hs, s_sh, W, Hw, s_rot, rot_all);
  }
}

}  // namespace

// smallest active block handled by the multishift sweep - 1 (blocks of at most nmin rows go to
// small_block)
static int hqr_nmin() { return WB_HQR_NMIN; }
static size_t hqr_lds_bytes(int W) {
  constexpr int M = WB_HQR_M;
  const size_t window = sizeof(cd) * (size_t)W * (W + 1);
  const size_t rots = sizeof(Rot) * (size_t)(W + 3 * M) * M;
  const int m = hqr_nmin(), ld = (m % 2) ? m : m + 1;
  const size_t small = 2 * sizeof(cd) * (size_t)m * ld;
  constexpr int jw = WB_HQR_AED, ldj = (jw % 2) ? jw : jw + 1;
  const size_t aed = sizeof(cd) * (2 * (size_t)jw * ldj + 192);   // T, U, reflector, eigenvalues
  size_t need = window;
  if (rots > need) need = rots;
  if (small > need) need = small;
  if (aed > need) need = aed;
  return need;
}

size_t wb_hqr_ms_rot_bytes(int n, int batch) {
  return sizeof(Rot) * (size_t)(n + 3 * WB_HQR_M) * WB_HQR_M * (size_t)batch;
}

// Multishift QR: H (Hessenberg, below-subdiagonal zero), Z (Q of the reduction, or Q^H when zh)
// -> Schur T in H, Schur vectors in Z (as Z^H when zh), eigenvalues in w.
hipError_t wb_hqr_ms_launch(int n, cd* H, cd* Z, int zh, cd* w, int* info, void* rot_ws, int batch,
                            hipStream_t stream) {
  constexpr int M = WB_HQR_M;
  static int W = 0;
  if (W == 0) {
    int dev = 0, smem = 65536;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&smem, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    // dynamic LDS: the W x W window; after the window is stored the same space holds its
    // rotations (at most W + 3M steps of M), and the small-block iteration uses it for T and
    // U. The rest is the kernel's static LDS.
    const size_t fixed = sizeof(Rot) * 2 * M + sizeof(int) * (NT / 64) + sizeof(cd) * (M + M * M) + 2048;
    for (int c = 128; c >= 32; c -= WB_HQR_WSTEP) {
      if (c < 3 * M + 16) break;
      if (hqr_lds_bytes(c) + fixed <= (size_t)smem) { W = c; break; }
    }
    if (W == 0) W = 3 * M + 16;
  }
  const size_t sh = hqr_lds_bytes(W);
  const int nmin = hqr_nmin();
  WB_TIMED_F(WB_T_HQR, stream, 72.0 * (double)n * n * n * batch,
             hipLaunchKernelGGL(k_hqr_ms<M>, dim3(batch), dim3(NT), sh, stream, n, H, Z, zh, w, info,
                                (Rot*)rot_ws, W, nmin));
  return hipGetLastError();
}
