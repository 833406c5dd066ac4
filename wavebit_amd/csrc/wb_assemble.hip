// Li-rule convolution-matrix assembly for raster layers, sampled coefficients, fixed nesting.
//
// Restates, in device code of its own, reference factorization.py:289-642 (the staircase
// series, the Toeplitz layout, the nested rule with v inside) for a stack of L rasters
// (L, Ny, Nx) complex128, row-major:
//   k_series_v     column series along v of eps and 1/eps: one workgroup per (column, layer,
//                  slot), one lane per order; the direct DFT of (a - a0)/n times the exact
//                  pixel transform sinc(k/n) exp(-i pi k/n), plus a0 at order 0.
//   k_series_2d    the direct 2D coefficients of eps (series along u of the column series).
//   k_inner_inv    every column's (2hv+1)^2 Toeplitz matrix inverted in LDS (Gauss-Jordan
//                  with partial pivoting, one workgroup per column).
//   k_outer_bttb   series along u of each inverse entry, scattered straight into the N x N
//                  block-Toeplitz layout out[(n,m),(n',m')] = A_{m-m'}[n,n'].
//   k_direct_bttb  the 2D direct rule M[(n,m),(n',m')] = c(n-n', m-m').
// Then epx = inv(T_u([T_v(eps)]^-1)) and epz_i = inv(M) on the batched LU path.
// Laurent's rule (reference factorization.py:944-1034) is the direct rule in every slot:
// epx = epy = M and epz_i = inv(M), from the column series of eps alone (k_series_v,
// k_series_2d, k_direct_bttb); its M is the same matrix, from the same arithmetic, as Li's.
// Each series workgroup first stages its n samples and the n twiddles e^{-2 pi i idx/n}
// (sincospi of the exactly reduced phase idx = k*j mod n) in LDS, so the per-order sums read
// LDS only and every twiddle is evaluated once per workgroup instead of once per (order,
// sample). The outer kernel writes its (2hu+1)^2 Toeplitz sub-block with whole-row runs.
// Series longer than SERIES_LDS_MAX samples take the global-memory kernels (same arithmetic).
#include "wb_internal.h"
#include "wb_timing.h"


namespace {

__device__ inline cd pixel_series(const cd* vals, long long stride, int n, int k, bool recip) {
  cd base = vals[0];
  if (recip) base = crecip(base);
  double sr = 0.0, si = 0.0;
  const double inv_n = 1.0 / (double)n;
  for (int j = 1; j < n; ++j) {
    cd a = vals[(long long)j * stride];
    if (recip) a = crecip(a);
    const cd d = cscale(csub(a, base), inv_n);
    long long idx = ((long long)k * j) % n;
    if (idx < 0) idx += n;
    double s, c;
    sincospi(-2.0 * (double)idx / (double)n, &s, &c);
    sr += d.x * c - d.y * s;
    si += d.x * s + d.y * c;
  }
  cd out = mkc(sr, si);
  if (n > 1) {
    const double tt = (double)k / (double)n;
    const double sinc = (k == 0) ? 1.0 : sinpi(tt) / (3.141592653589793 * tt);
    double s, c;
    sincospi(-tt, &s, &c);
    out = cscale(cmul(out, mkc(c, s)), sinc);
  } else {
    out = mkc(0.0, 0.0);
  }
  if (k == 0) out = cadd(out, base);
  return out;
}

// ------------------------------------------------------------------ LDS-staged series
constexpr int SERIES_LDS_MAX = 2048;   // samples per staged series: 2 x 32 KiB of LDS

// Stage samples s[j] = g[j * stride] (their reciprocals if recip) and twiddles
// tw[idx] = e^{-2 pi i idx / n} in LDS; the caller synchronizes before reading them.
__device__ inline void stage_series(const cd* g, long long stride, int n, bool recip, cd* s, cd* tw) {
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    cd a = g[(long long)j * stride];
    s[j] = recip ? crecip(a) : a;
    double sn, c;
    sincospi(-2.0 * (double)j / (double)n, &sn, &c);
    tw[j] = mkc(c, sn);
  }
}

// pixel_series over staged samples and twiddles: the same terms in the same order (the phase
// index k*j mod n is advanced by k mod n per sample).
__device__ inline cd pixel_series_lds(const cd* s, const cd* tw, int n, int k) {
  const cd base = s[0];
  double sr = 0.0, si = 0.0;
  const double inv_n = 1.0 / (double)n;
  const int kmod = ((k % n) + n) % n;
  int idx = 0;
  for (int j = 1; j < n; ++j) {
    idx += kmod;
    if (idx >= n) idx -= n;
    const cd d = cscale(csub(s[j], base), inv_n);
    const cd w = tw[idx];
    sr += d.x * w.x - d.y * w.y;
    si += d.x * w.y + d.y * w.x;
  }
  cd out = mkc(sr, si);
  if (n > 1) {
    const double tt = (double)k / (double)n;
    const double sinc = (k == 0) ? 1.0 : sinpi(tt) / (3.141592653589793 * tt);
    double sn, c;
    sincospi(-tt, &sn, &c);
    out = cscale(cmul(out, mkc(c, sn)), sinc);
  } else {
    out = mkc(0.0, 0.0);
  }
  if (k == 0) out = cadd(out, base);
  return out;
}

// k_series_v with the column staged in LDS (dynamic LDS: 2 Ny complex)
__global__ __launch_bounds__(64) void k_series_v_lds(const cd* eps, int L, int Ny, int Nx, int hv, cd* lines) {
  extern __shared__ cd sm[];
  const int ju = blockIdx.x, l = blockIdx.y, slot = blockIdx.z;
  const int Kv = 4 * hv + 1;
  stage_series(eps + (long long)l * Ny * Nx + ju, Nx, Ny, slot == 1, sm, sm + Ny);
  __syncthreads();
  for (int kk = threadIdx.x; kk < Kv; kk += blockDim.x)
    lines[(((long long)slot * L + l) * Nx + ju) * Kv + kk] = pixel_series_lds(sm, sm + Ny, Ny, kk - 2 * hv);
}

// k_series_2d with the series staged in LDS (dynamic LDS: 2 Nx complex)
__global__ __launch_bounds__(64) void k_series_2d_lds(const cd* lines, int L, int Nx, int hu, int hv, cd* c2d) {
  extern __shared__ cd sm[];
  const int kv = blockIdx.x, l = blockIdx.y;
  const int Kv = 4 * hv + 1, Ku = 4 * hu + 1;
  stage_series(lines + ((long long)l * Nx) * Kv + kv, Kv, Nx, false, sm, sm + Nx);
  __syncthreads();
  for (int kk = threadIdx.x; kk < Ku; kk += blockDim.x)
    c2d[((long long)l * Kv + kv) * Ku + kk] = pixel_series_lds(sm, sm + Nx, Nx, kk - 2 * hu);
}

// k_outer_bttb with the series staged in LDS: the Ku coefficients of inner entry e = (n1, n2)
// go to LDS, then the workgroup writes the nu x nu Toeplitz sub-block (n1, n2) of the output,
// element (m, m') = coefficient m - m', consecutive lanes along each output row.
// Dynamic LDS: 2 Nx + Ku complex.
__global__ __launch_bounds__(256) void k_outer_bttb_lds(const cd* inner, int L, int Nx, int hu, int hv,
                                                        cd* outA, cd* outB) {
  extern __shared__ cd sm[];
  const int e = blockIdx.x, l = blockIdx.y, slot = blockIdx.z;
  const int nv = 2 * hv + 1, nu = 2 * hu + 1, Ku = 4 * hu + 1;
  const int N = nu * nv;
  const int n1 = e / nv, n2 = e % nv;
  cd* coef = sm + 2 * Nx;
  stage_series(inner + (((long long)slot * L + l) * Nx) * nv * nv + e, (long long)nv * nv, Nx, false,
               sm, sm + Nx);
  __syncthreads();
  for (int kk = threadIdx.x; kk < Ku; kk += blockDim.x)
    coef[kk] = pixel_series_lds(sm, sm + Nx, Nx, kk - 2 * hu);
  __syncthreads();
  cd* out = (slot == 0 ? outA : outB) + (long long)l * N * N + (long long)(n1 * nu) * N + n2 * nu;
  for (int t = threadIdx.x; t < nu * nu; t += blockDim.x) {
    const int m = t / nu, mp = t % nu;
    out[(long long)m * N + mp] = coef[m - mp + 2 * hu];
  }
}

// ------------------------------------------------------------------ global-memory series
// lines[slot][l][ju][kv], slot 0: eps, slot 1: 1/eps
__global__ void k_series_v(const cd* eps, int L, int Ny, int Nx, int hv, cd* lines) {
  const int ju = blockIdx.x, l = blockIdx.y, slot = blockIdx.z;
  const int Kv = 4 * hv + 1;
  const cd* col = eps + (long long)l * Ny * Nx + ju;
  for (int kk = threadIdx.x; kk < Kv; kk += blockDim.x) {
    const int k = kk - 2 * hv;
    lines[(((long long)slot * L + l) * Nx + ju) * Kv + kk] = pixel_series(col, Nx, Ny, k, slot == 1);
  }
}

// c2d[l][kv][ku] = series_u over ju of lines[0][l][ju][kv]
__global__ void k_series_2d(const cd* lines, int L, int Nx, int hu, int hv, cd* c2d) {
  const int kv = blockIdx.x, l = blockIdx.y;
  const int Kv = 4 * hv + 1, Ku = 4 * hu + 1;
  const cd* seq = lines + ((long long)l * Nx) * Kv + kv;
  for (int kk = threadIdx.x; kk < Ku; kk += blockDim.x)
    c2d[((long long)l * Kv + kv) * Ku + kk] = pixel_series(seq, Kv, Nx, kk - 2 * hu, false);
}

// inner[slot][l][ju][nv*nv] = inv(Toeplitz(lines[slot][l][ju]))
__global__ __launch_bounds__(64) void k_inner_inv(const cd* lines, int L, int Nx, int hv, cd* inner) {
  const int ju = blockIdx.x, l = blockIdx.y, slot = blockIdx.z;
  const int nv = 2 * hv + 1, Kv = 4 * hv + 1;
  const cd* ln = lines + (((long long)slot * L + l) * Nx + ju) * Kv;
  __shared__ cd M[31][63];
  __shared__ int s_piv;
  const int t = threadIdx.x;
  for (int e = t; e < nv * 2 * nv; e += 64) {
    const int r = e / (2 * nv), c = e % (2 * nv);
    M[r][c] = (c < nv) ? ln[r - c + 2 * hv] : mkc((c - nv) == r ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  for (int j = 0; j < nv; ++j) {
    if (t == 0) {
      int p = j;
      double best = cabs1(M[j][j]);
      for (int r = j + 1; r < nv; ++r) {
        const double v = cabs1(M[r][j]);
        if (v > best) { best = v; p = r; }
      }
      s_piv = p;
    }
    __syncthreads();
    const int p = s_piv;
    if (p != j) {
      for (int c = t; c < 2 * nv; c += 64) { cd x = M[j][c]; M[j][c] = M[p][c]; M[p][c] = x; }
    }
    __syncthreads();
    const cd rd = crecip(M[j][j]);
    __syncthreads();
    for (int c = t; c < 2 * nv; c += 64) M[j][c] = cmul(M[j][c], rd);
    __syncthreads();
    for (int e = t; e < nv * 2 * nv; e += 64) {
      const int r = e / (2 * nv), c = e % (2 * nv);
      if (r != j && c != j) M[r][c] = csub(M[r][c], cmul(M[r][j], M[j][c]));
    }
    __syncthreads();
    for (int r = t; r < nv; r += 64)
      if (r != j) M[r][j] = mkc(0.0, 0.0);
    __syncthreads();
  }
  cd* out = inner + (((long long)slot * L + l) * Nx + ju) * nv * nv;
  for (int e = t; e < nv * nv; e += 64) out[e] = M[e / nv][nv + e % nv];
}

// out[slot][l] (N x N) from the u-series of each inner entry.
__global__ void k_outer_bttb(const cd* inner, int L, int Nx, int hu, int hv, cd* outA, cd* outB) {
  const int e = blockIdx.x, l = blockIdx.y, slot = blockIdx.z;
  const int nv = 2 * hv + 1, nu = 2 * hu + 1, Ku = 4 * hu + 1;
  const int N = nu * nv;
  const int n1 = e / nv, n2 = e % nv;
  const cd* seq = inner + (((long long)slot * L + l) * Nx) * nv * nv + e;
  cd* out = (slot == 0 ? outA : outB) + (long long)l * N * N;
  for (int kk = threadIdx.x; kk < Ku; kk += blockDim.x) {
    const cd c = pixel_series(seq, (long long)nv * nv, Nx, kk - 2 * hu, false);
    const int d = kk - 2 * hu;   // m - m'
    for (int m = 0; m < nu; ++m) {
      const int mp = m - d;
      if (mp < 0 || mp >= nu) continue;
      out[(long long)(n1 * nu + m) * N + (n2 * nu + mp)] = c;
    }
  }
}

// out, and out2/out3 when given, receive the same matrix.
__global__ void k_direct_bttb(const cd* c2d, int L, int hu, int hv, cd* out, cd* out2, cd* out3) {
  const int l = blockIdx.y;
  const int nu = 2 * hu + 1, nv = 2 * hv + 1, N = nu * nv, Ku = 4 * hu + 1;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)N * N) return;
  const int r = (int)(e / N), c = (int)(e % N);
  const int n1 = r / nu, m1 = r % nu, n2 = c / nu, m2 = c % nu;
  const cd v = c2d[((long long)l * (4 * hv + 1) + (n1 - n2 + 2 * hv)) * Ku + (m1 - m2 + 2 * hu)];
  const long long o = (long long)l * N * N + e;
  out[o] = v;
  if (out2) out2[o] = v;
  if (out3) out3[o] = v;
}

// ------------------------------------------------------------------------------ adjoint
// The pixel series is linear in the samples: c_k = sum_j T[k][j] a_j with, for j >= 1,
// T[k][j] = g_k w^{kj} / n and T[k][0] = -g_k S_k / n + [k == 0], g_k = sinc(k/n) e^{-i pi k/n},
// w = e^{-2 pi i / n}, S_k = sum_{j>=1} w^{kj} = n - 1 if n | k else -1. Its adjoint:
// abar_j = sum_k conj(T[k][j]) cbar_k, over the 4 half + 1 orders k = -2 half .. 2 half.
__device__ inline cd pixel_series_adj(const cd* cbar, int cstride, int n, int j, int half) {
  const int K = 4 * half + 1;
  if (n <= 1) return (j == 0) ? cbar[(long long)(2 * half) * cstride] : mkc(0.0, 0.0);
  double sr = 0.0, si = 0.0;
  const double inv_n = 1.0 / (double)n;
  for (int kk = 0; kk < K; ++kk) {
    const int k = kk - 2 * half;
    const cd cb = cbar[(long long)kk * cstride];
    const double tt = (double)k / (double)n;
    const double sinc = (k == 0) ? 1.0 : sinpi(tt) / (3.141592653589793 * tt);
    double gs, gc;
    sincospi(-tt, &gs, &gc);
    const cd gconj = mkc(sinc * gc, -sinc * gs);
    cd coef;
    if (j >= 1) {
      long long idx = ((long long)k * j) % n;
      if (idx < 0) idx += n;
      double s, c;
      sincospi(2.0 * (double)idx / (double)n, &s, &c);
      coef = cscale(cmul(gconj, mkc(c, s)), inv_n);
    } else {
      const double Sk = (k % n == 0) ? (double)(n - 1) : -1.0;
      coef = cscale(gconj, -Sk * inv_n);
      if (k == 0) coef = cadd(coef, mkc(1.0, 0.0));
    }
    sr += coef.x * cb.x - coef.y * cb.y;
    si += coef.x * cb.y + coef.y * cb.x;
  }
  return mkc(sr, si);
}

// c2d_bar[l][kv][ku] = sum of zpre_bar over the entries the direct rule fills from c2d[kv][ku]
__global__ void k_direct_adj(const cd* zpre_bar, int L, int hu, int hv, cd* c2d_bar) {
  const int nu = 2 * hu + 1, nv = 2 * hv + 1, N = nu * nv, Ku = 4 * hu + 1, Kv = 4 * hv + 1;
  const int l = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= Kv * Ku) return;
  const int kv = e / Ku, ku = e % Ku;
  const int dn = kv - 2 * hv, dm = ku - 2 * hu;
  const cd* Z = zpre_bar + (long long)l * N * N;
  cd acc = mkc(0.0, 0.0);
  for (int n1 = 0; n1 < nv; ++n1) {
    const int n2 = n1 - dn;
    if (n2 < 0 || n2 >= nv) continue;
    for (int m1 = 0; m1 < nu; ++m1) {
      const int m2 = m1 - dm;
      if (m2 < 0 || m2 >= nu) continue;
      acc = cadd(acc, Z[(long long)(n1 * nu + m1) * N + (n2 * nu + m2)]);
    }
  }
  c2d_bar[((long long)l * Kv + kv) * Ku + ku] = acc;
}

// inner_bar[slot][l][ju][e] from the block-Toeplitz outputs' cotangents (slot 0: pre_bar of
// epx, slot 1: epy-bar): first the per-offset sums, then the u-series adjoint over columns.
__global__ __launch_bounds__(64) void k_outer_adj(const cd* pre_bar, const cd* epy_bar, int L, int Nx,
                                                  int hu, int hv, cd* inner_bar) {
  const int e = blockIdx.x, l = blockIdx.y, slot = blockIdx.z;
  const int nv = 2 * hv + 1, nu = 2 * hu + 1, Ku = 4 * hu + 1, N = nu * nv;
  const int n1 = e / nv, n2 = e % nv;
  const cd* ob = (slot == 0 ? pre_bar : epy_bar) + (long long)l * N * N;
  __shared__ cd cu[61];
  for (int kk = threadIdx.x; kk < Ku; kk += blockDim.x) {
    const int d = kk - 2 * hu;
    cd acc = mkc(0.0, 0.0);
    for (int m = 0; m < nu; ++m) {
      const int mp = m - d;
      if (mp < 0 || mp >= nu) continue;
      acc = cadd(acc, ob[(long long)(n1 * nu + m) * N + (n2 * nu + mp)]);
    }
    cu[kk] = acc;
  }
  __syncthreads();
  cd* out = inner_bar + (((long long)slot * L + l) * Nx) * nv * nv + e;
  for (int ju = threadIdx.x; ju < Nx; ju += blockDim.x)
    out[(long long)ju * nv * nv] = pixel_series_adj(cu, 1, Nx, ju, hu);
}

// Per column: Toeplitz-bar = -inner^H inner_bar inner^H, then its diagonals summed into the
// series cotangent lines_bar[slot][l][ju][kk] (kk = r - c + 2hv).
__global__ __launch_bounds__(64) void k_inner_adj(const cd* inner, const cd* inner_bar, int L, int Nx,
                                                  int hv, cd* lines_bar) {
  const int ju = blockIdx.x, l = blockIdx.y, slot = blockIdx.z;
  const int nv = 2 * hv + 1, Kv = 4 * hv + 1;
  const long long off = (((long long)slot * L + l) * Nx + ju) * nv * nv;
  __shared__ cd Ai[31][32], Gb[31][32], Tm[31][32];
  const int t = threadIdx.x;
  for (int e = t; e < nv * nv; e += 64) {
    Ai[e / nv][e % nv] = inner[off + e];
    Gb[e / nv][e % nv] = inner_bar[off + e];
  }
  __syncthreads();
  // Tm = Gb Ai^H
  for (int e = t; e < nv * nv; e += 64) {
    const int r = e / nv, c = e % nv;
    cd acc = mkc(0.0, 0.0);
    for (int k = 0; k < nv; ++k) acc = cadd(acc, cmul(Gb[r][k], cconj(Ai[c][k])));
    Tm[r][c] = acc;
  }
  __syncthreads();
  // Gb <- -Ai^H Tm
  for (int e = t; e < nv * nv; e += 64) {
    const int r = e / nv, c = e % nv;
    cd acc = mkc(0.0, 0.0);
    for (int k = 0; k < nv; ++k) acc = cadd(acc, cmul(cconj(Ai[k][r]), Tm[k][c]));
    Gb[r][c] = cneg(acc);
  }
  __syncthreads();
  cd* lb = lines_bar + (((long long)slot * L + l) * Nx + ju) * Kv;
  for (int kk = t; kk < Kv; kk += 64) {
    const int dd = kk - 2 * hv;   // r - c
    cd acc = mkc(0.0, 0.0);
    for (int r = 0; r < nv; ++r) {
      const int c = r - dd;
      if (c >= 0 && c < nv) acc = cadd(acc, Gb[r][c]);
    }
    lb[kk] = acc;
  }
}

// lines_bar[0][l][ju][kv] += u-series adjoint of c2d_bar[l][kv][:] (= when !accumulate)
__global__ __launch_bounds__(64) void k_series2d_adj(const cd* c2d_bar, int L, int Nx, int hu, int hv,
                                                     cd* lines_bar, int accumulate) {
  const int kv = blockIdx.x, l = blockIdx.y;
  const int Kv = 4 * hv + 1, Ku = 4 * hu + 1;
  const cd* cb = c2d_bar + ((long long)l * Kv + kv) * Ku;
  for (int ju = threadIdx.x; ju < Nx; ju += blockDim.x) {
    cd* x = lines_bar + ((long long)l * Nx + ju) * Kv + kv;
    const cd g = pixel_series_adj(cb, 1, Nx, ju, hu);
    *x = accumulate ? cadd(*x, g) : g;
  }
}

// eps_bar[l][jv][ju] = v-series adjoint of lines_bar[0] + conj(-1/eps^2) of lines_bar[1]
// (the second term only when recip: Laurent's rule has no 1/eps series)
__global__ __launch_bounds__(64) void k_series_v_adj(const cd* eps, const cd* lines_bar, int L, int Ny,
                                                     int Nx, int hv, cd* eps_bar, int recip) {
  const int ju = blockIdx.x, l = blockIdx.y;
  const int Kv = 4 * hv + 1;
  const cd* lb0 = lines_bar + ((long long)l * Nx + ju) * Kv;
  const cd* lb1 = lines_bar + (((long long)L + l) * Nx + ju) * Kv;
  for (int jv = threadIdx.x; jv < Ny; jv += blockDim.x) {
    const long long o = ((long long)l * Ny + jv) * Nx + ju;
    const cd g0 = pixel_series_adj(lb0, 1, Ny, jv, hv);
    if (!recip) {
      eps_bar[o] = g0;
      continue;
    }
    const cd a = eps[o];
    const cd r = crecip(a);
    const cd dr = cneg(cmul(r, r));          // d(1/a)/da
    const cd g1 = pixel_series_adj(lb1, 1, Ny, jv, hv);
    eps_bar[o] = cadd(g0, cmulc(dr, g1));
  }
}

// Column series of eps and (slots == 2) 1/eps (lines), staged in LDS up to SERIES_LDS_MAX rows.
hipError_t series_v_launch(const cd* eps, int L, int Ny, int Nx, int hv, cd* lines, hipStream_t stream,
                           int slots = 2) {
  if (Ny <= SERIES_LDS_MAX)
    hipLaunchKernelGGL(k_series_v_lds, dim3(Nx, L, slots), dim3(64), 2 * Ny * sizeof(cd), stream, eps, L,
                       Ny, Nx, hv, lines);
  else
    hipLaunchKernelGGL(k_series_v, dim3(Nx, L, slots), dim3(64), 0, stream, eps, L, Ny, Nx, hv, lines);
  return hipGetLastError();
}

// The direct 2D coefficients c2d (L, Kv, Ku) of eps from its column series (slot 0 of lines).
hipError_t series_2d_launch(const cd* lines, int L, int Nx, int hu, int hv, cd* c2d, hipStream_t stream) {
  const int Kv = 4 * hv + 1;
  if (Nx <= SERIES_LDS_MAX)
    hipLaunchKernelGGL(k_series_2d_lds, dim3(Kv, L), dim3(64), 2 * Nx * sizeof(cd), stream, lines, L, Nx,
                       hu, hv, c2d);
  else
    hipLaunchKernelGGL(k_series_2d, dim3(Kv, L), dim3(64), 0, stream, lines, L, Nx, hu, hv, c2d);
  return hipGetLastError();
}

// z[i] += x[i] + y[i] over n complex entries
__global__ void k_add2(cd* z, const cd* x, const cd* y, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = cadd(z[i], cadd(x[i], y[i]));
}

}  // namespace

size_t wb_assemble_workspace_bytes(int L, int Ny, int Nx, int hu, int hv) {
  const size_t nv = 2 * hv + 1, nu = 2 * hu + 1, N = nu * nv;
  const size_t Kv = 4 * hv + 1, Ku = 4 * hu + 1;
  size_t b = 0;
  b += sizeof(cd) * 2 * L * Nx * Kv;       // lines
  b += sizeof(cd) * L * Kv * Ku;           // c2d
  b += sizeof(cd) * 2 * L * Nx * nv * nv;  // inner inverses
  b += sizeof(cd) * 2 * L * N * N;         // pre-inversion epx, epz
  b += sizeof(int) * 2 * L * N + sizeof(int) * 2 * L;
  return b + 1024;
}

hipError_t wb_assemble_li_launch(const cd* eps, int L, int Ny, int Nx, int hu, int hv, cd* epx,
                                 cd* epy, cd* epz_i, void* ws, int* info_out, hipStream_t stream) {
  const int nv = 2 * hv + 1, nu = 2 * hu + 1, N = nu * nv;
  const int Kv = 4 * hv + 1, Ku = 4 * hu + 1;
  char* p = (char*)ws;
  cd* lines = (cd*)p; p += sizeof(cd) * 2 * (size_t)L * Nx * Kv;
  cd* c2d = (cd*)p; p += sizeof(cd) * (size_t)L * Kv * Ku;
  cd* inner = (cd*)p; p += sizeof(cd) * 2 * (size_t)L * Nx * nv * nv;
  cd* pre = (cd*)p; p += sizeof(cd) * 2 * (size_t)L * N * N;     // [0]: epx pre, [1]: epz pre
  int* ipiv = (int*)p; p += sizeof(int) * 2 * (size_t)L * N;
  int* info = info_out;
  hipError_t e;
  cd* epz_pre = pre + (size_t)L * N * N;
  const long long NN = (long long)N * N;
  {
    // the gathers (raster -> epx pre-inversion, epy, epz pre-inversion): one timed stage
    WbStageTimer timer(WB_T_ASSEMBLE, stream, 0.0, WB_STAGE_BOTH);
    if ((e = series_v_launch(eps, L, Ny, Nx, hv, lines, stream)) != hipSuccess) return e;
    if ((e = series_2d_launch(lines, L, Nx, hu, hv, c2d, stream)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_inner_inv, dim3(Nx, L, 2), dim3(64), 0, stream, lines, L, Nx, hv, inner);
    WB_CHECK_LAUNCH();
    // slot 0 (eps) -> epx pre-inversion, slot 1 (1/eps) -> epy directly
    if ((2 * Nx + Ku) * sizeof(cd) <= 65536)
      hipLaunchKernelGGL(k_outer_bttb_lds, dim3(nv * nv, L, 2), dim3(256), (2 * Nx + Ku) * sizeof(cd), stream,
                         inner, L, Nx, hu, hv, pre, epy);
    else
      hipLaunchKernelGGL(k_outer_bttb, dim3(nv * nv, L, 2), dim3(64), 0, stream, inner, L, Nx, hu, hv,
                         pre, epy);
    WB_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_direct_bttb, dim3((unsigned)((NN + 255) / 256), L), dim3(256), 0, stream, c2d,
                       L, hu, hv, epz_pre, (cd*)nullptr, (cd*)nullptr);
    WB_CHECK_LAUNCH();
  }
  // epx = inv(pre), epz_i = inv(epz_pre): one batched inversion of 2L matrices
  // output slots are separate buffers: use a grid BatchPtr over (slot, layer)
  BatchPtr A = bp_strided(pre, NN);
  e = wb_inverse_launch(N, A, N, bp_strided(epx, NN), N, ipiv, info, L, stream);
  if (e != hipSuccess) return e;
  e = wb_inverse_launch(N, bp_strided(epz_pre, NN), N, bp_strided(epz_i, NN), N, ipiv + (size_t)L * N,
                        info + L, L, stream);
  return e;
}

size_t wb_assemble_adjoint_workspace_bytes(int L, int Ny, int Nx, int hu, int hv) {
  const size_t nv = 2 * hv + 1, nu = 2 * hu + 1, N = nu * nv;
  const size_t Kv = 4 * hv + 1, Ku = 4 * hu + 1;
  size_t b = 0;
  b += sizeof(cd) * 2 * L * Nx * Kv * 2;        // lines, lines_bar
  b += sizeof(cd) * 2 * L * Nx * nv * nv * 2;   // inner, inner_bar
  b += sizeof(cd) * L * Kv * Ku;                // c2d_bar
  b += sizeof(cd) * 3 * L * N * N;              // pre_bar, zpre_bar, GEMM temporary
  return b + 1024;
}

// d_eps (L, Ny, Nx) from the cotangents of epx, epy, epz_i (L, N, N each); epx, epz_i are the
// forward's outputs for the same eps.
hipError_t wb_assemble_adjoint_launch(const cd* eps, int L, int Ny, int Nx, int hu, int hv,
                                      const cd* epx, const cd* epz_i, const cd* d_epx,
                                      const cd* d_epy, const cd* d_epz, cd* d_eps, void* ws,
                                      hipStream_t stream) {
  const int nv = 2 * hv + 1, nu = 2 * hu + 1, N = nu * nv;
  const int Kv = 4 * hv + 1, Ku = 4 * hu + 1;
  const long long NN = (long long)N * N;
  char* p = (char*)ws;
  cd* lines = (cd*)p; p += sizeof(cd) * 2 * (size_t)L * Nx * Kv;
  cd* lines_bar = (cd*)p; p += sizeof(cd) * 2 * (size_t)L * Nx * Kv;
  cd* inner = (cd*)p; p += sizeof(cd) * 2 * (size_t)L * Nx * nv * nv;
  cd* inner_bar = (cd*)p; p += sizeof(cd) * 2 * (size_t)L * Nx * nv * nv;
  cd* c2d_bar = (cd*)p; p += sizeof(cd) * (size_t)L * Kv * Ku;
  cd* pre_bar = (cd*)p; p += sizeof(cd) * (size_t)L * NN;
  cd* zpre_bar = (cd*)p; p += sizeof(cd) * (size_t)L * NN;
  cd* tmp = (cd*)p; p += sizeof(cd) * (size_t)L * NN;
  const cd one = mkc(1.0, 0.0), zero = mkc(0.0, 0.0), mone = mkc(-1.0, 0.0);
  hipError_t e;
  // recompute the column series and their Toeplitz inverses
  if ((e = series_v_launch(eps, L, Ny, Nx, hv, lines, stream)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_inner_inv, dim3(Nx, L, 2), dim3(64), 0, stream, lines, L, Nx, hv, inner);
  WB_CHECK_LAUNCH();
  // outer inverses: pre_bar = -epx^H d_epx epx^H, zpre_bar = -epz^H d_epz epz^H
  if ((e = wb_zgemm_ops_launch(0, 1, N, N, N, one, bp_strided(d_epx, NN), N, bp_strided(epx, NN), N,
                               zero, bp_strided(tmp, NN), N, L, stream))) return e;
  if ((e = wb_zgemm_ops_launch(1, 0, N, N, N, mone, bp_strided(epx, NN), N, bp_strided(tmp, NN), N,
                               zero, bp_strided(pre_bar, NN), N, L, stream))) return e;
  if ((e = wb_zgemm_ops_launch(0, 1, N, N, N, one, bp_strided(d_epz, NN), N, bp_strided(epz_i, NN), N,
                               zero, bp_strided(tmp, NN), N, L, stream))) return e;
  if ((e = wb_zgemm_ops_launch(1, 0, N, N, N, mone, bp_strided(epz_i, NN), N, bp_strided(tmp, NN), N,
                               zero, bp_strided(zpre_bar, NN), N, L, stream))) return e;
  hipLaunchKernelGGL(k_direct_adj, dim3((Kv * Ku + 255) / 256, L), dim3(256), 0, stream, zpre_bar, L,
                     hu, hv, c2d_bar);
  WB_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_outer_adj, dim3(nv * nv, L, 2), dim3(64), 0, stream, pre_bar, d_epy, L, Nx, hu,
                     hv, inner_bar);
  WB_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_inner_adj, dim3(Nx, L, 2), dim3(64), 0, stream, inner, inner_bar, L, Nx, hv,
                     lines_bar);
  WB_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_series2d_adj, dim3(Kv, L), dim3(64), 0, stream, c2d_bar, L, Nx, hu, hv,
                     lines_bar, 1);
  WB_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_series_v_adj, dim3(Nx, L), dim3(64), 0, stream, eps, lines_bar, L, Ny, Nx, hv,
                     d_eps, 1);
  WB_CHECK_LAUNCH();
  return hipSuccess;
}

// ------------------------------------------------------------------------ Laurent's rule
size_t wb_assemble_laurent_workspace_bytes(int L, int Ny, int Nx, int hu, int hv) {
  const size_t nv = 2 * hv + 1, nu = 2 * hu + 1, N = nu * nv;
  const size_t Kv = 4 * hv + 1, Ku = 4 * hu + 1;
  size_t b = 0;
  b += sizeof(cd) * L * Nx * Kv;   // column series of eps
  b += sizeof(cd) * L * Kv * Ku;   // c2d
  b += sizeof(cd) * L * N * N;     // the direct rule, factored in place by the inversion
  b += sizeof(int) * L * N;        // pivots
  return b + 1024;
}

hipError_t wb_assemble_laurent_launch(const cd* eps, int L, int Ny, int Nx, int hu, int hv, cd* epx,
                                      cd* epy, cd* epz_i, void* ws, int* info, hipStream_t stream) {
  const int nv = 2 * hv + 1, nu = 2 * hu + 1, N = nu * nv;
  const int Kv = 4 * hv + 1, Ku = 4 * hu + 1;
  char* p = (char*)ws;
  cd* lines = (cd*)p; p += sizeof(cd) * (size_t)L * Nx * Kv;
  cd* c2d = (cd*)p; p += sizeof(cd) * (size_t)L * Kv * Ku;
  cd* pre = (cd*)p; p += sizeof(cd) * (size_t)L * N * N;
  int* ipiv = (int*)p;
  const long long NN = (long long)N * N;
  hipError_t e;
  {
    WbStageTimer timer(WB_T_ASSEMBLE, stream, 0.0, WB_STAGE_BOTH);
    if ((e = series_v_launch(eps, L, Ny, Nx, hv, lines, stream, 1)) != hipSuccess) return e;
    if ((e = series_2d_launch(lines, L, Nx, hu, hv, c2d, stream)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_direct_bttb, dim3((unsigned)((NN + 255) / 256), L), dim3(256), 0, stream, c2d,
                       L, hu, hv, epx, epy, pre);
    WB_CHECK_LAUNCH();
  }
  return wb_inverse_launch(N, bp_strided(pre, NN), N, bp_strided(epz_i, NN), N, ipiv, info, L, stream);
}

size_t wb_assemble_laurent_adjoint_workspace_bytes(int L, int Ny, int Nx, int hu, int hv) {
  const size_t nv = 2 * hv + 1, nu = 2 * hu + 1, N = nu * nv;
  const size_t Kv = 4 * hv + 1, Ku = 4 * hu + 1;
  size_t b = 0;
  b += sizeof(cd) * L * Nx * Kv;   // lines_bar
  b += sizeof(cd) * L * Kv * Ku;   // c2d_bar
  b += sizeof(cd) * 2 * L * N * N; // cotangent of the direct rule, GEMM temporary
  return b + 1024;
}

// d_eps from the cotangents of epx = epy = M and epz_i = inv(M):
// M_bar = d_epx + d_epy - epz_i^H d_epz epz_i^H, then the direct rule's and the series' adjoints.
hipError_t wb_assemble_laurent_adjoint_launch(const cd* eps, int L, int Ny, int Nx, int hu, int hv,
                                              const cd* epz_i, const cd* d_epx, const cd* d_epy,
                                              const cd* d_epz, cd* d_eps, void* ws, hipStream_t stream) {
  const int nv = 2 * hv + 1, nu = 2 * hu + 1, N = nu * nv;
  const int Kv = 4 * hv + 1, Ku = 4 * hu + 1;
  const long long NN = (long long)N * N;
  char* p = (char*)ws;
  cd* lines_bar = (cd*)p; p += sizeof(cd) * (size_t)L * Nx * Kv;
  cd* c2d_bar = (cd*)p; p += sizeof(cd) * (size_t)L * Kv * Ku;
  cd* zbar = (cd*)p; p += sizeof(cd) * (size_t)L * NN;
  cd* tmp = (cd*)p;
  const cd one = mkc(1.0, 0.0), zero = mkc(0.0, 0.0), mone = mkc(-1.0, 0.0);
  hipError_t e;
  if ((e = wb_zgemm_ops_launch(0, 1, N, N, N, one, bp_strided(d_epz, NN), N, bp_strided(epz_i, NN), N,
                               zero, bp_strided(tmp, NN), N, L, stream))) return e;
  if ((e = wb_zgemm_ops_launch(1, 0, N, N, N, mone, bp_strided(epz_i, NN), N, bp_strided(tmp, NN), N,
                               zero, bp_strided(zbar, NN), N, L, stream))) return e;
  const long long tot = (long long)L * NN;
  hipLaunchKernelGGL(k_add2, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, stream, zbar, d_epx, d_epy,
                     tot);
  WB_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_direct_adj, dim3((Kv * Ku + 255) / 256, L), dim3(256), 0, stream, zbar, L, hu, hv,
                     c2d_bar);
  WB_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_series2d_adj, dim3(Kv, L), dim3(64), 0, stream, c2d_bar, L, Nx, hu, hv, lines_bar,
                     0);
  WB_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_series_v_adj, dim3(Nx, L), dim3(64), 0, stream, eps, lines_bar, L, Ny, Nx, hv, d_eps,
                     0);
  WB_CHECK_LAUNCH();
  return hipSuccess;
}

// ------------------------------------------------------------------- normal-vector-field rule
// NVF with the default construction (reference factorization.py:811-878 the floor lift,
// :1054-1404 the normal field from the periodic distance map, :1934-1966 the projector series,
// :2461-2540 the closing, :2796-2815 the layer's builder): per layer
//   E = [[eps]], A = [[1/eps]]^-1, Delta = E - A, C_ab = [[n_a n_b]] (point-sample series of the
//   normal field on the raster lifted to 4 h + 1 samples per axis),
//   epx = E - (Delta C_xx + C_xx Delta)/2, epy = E - (Delta C_yy + C_yy Delta)/2,
//   epxy = epyx = (Delta C_xy + C_xy Delta)/2, epz_i = E^-1.
// The normal field: the interface mask (pixels differing from a 4-neighbour, wrapping), the
// exact periodic Euclidean distance to it in two separable passes (per column the nearest mask
// row in either direction around the period, then per row the lower envelope over every
// column, brute force from LDS), its wrapping central differences carried to Cartesian
// components by the cell's J^-T, normalised where the length is at least 1e-9. A raster with a
// one-cell axis (after the lift) has the constant normal the host passes.
namespace {

constexpr double NVF_FLAT_PIXEL = 1e-9;      // factorization.py:1462-1464

// Point-sample series along one axis, orders -2h..2h: sum_j (a_j / n) w^{kj} times the
// half-pixel phase e^{-i pi k / n} (factorization.py:322-338, 398-440): no box transform and no
// base split.
__device__ inline cd point_series_lds(const cd* s, const cd* tw, int n, int k) {
  double sr = 0.0, si = 0.0;
  const double inv_n = 1.0 / (double)n;
  const int kmod = ((k % n) + n) % n;
  int idx = 0;
  for (int j = 0; j < n; ++j) {
    const cd d = cscale(s[j], inv_n);
    const cd w = tw[idx];
    sr += d.x * w.x - d.y * w.y;
    si += d.x * w.y + d.y * w.x;
    idx += kmod;
    if (idx >= n) idx -= n;
  }
  double sn, c;
  sincospi(-(double)k / (double)n, &sn, &c);
  return cmul(mkc(sr, si), mkc(c, sn));
}

// The same sum with the twiddles evaluated per term (series longer than SERIES_LDS_MAX).
__device__ inline cd point_series_global(const cd* g, long long stride, int n, int k) {
  double sr = 0.0, si = 0.0;
  const double inv_n = 1.0 / (double)n;
  for (int j = 0; j < n; ++j) {
    const cd d = cscale(g[(long long)j * stride], inv_n);
    long long idx = ((long long)k * j) % n;
    if (idx < 0) idx += n;
    double s, c;
    sincospi(-2.0 * (double)idx / (double)n, &s, &c);
    sr += d.x * c - d.y * s;
    si += d.x * s + d.y * c;
  }
  double sn, c;
  sincospi(-(double)k / (double)n, &sn, &c);
  return cmul(mkc(sr, si), mkc(c, sn));
}

// lines[g][ju][kv] = v-series of column ju of grid g (g over the S grids of Ny x Nx samples)
__global__ __launch_bounds__(64) void k_point_v(const cd* f, int Ny, int Nx, int hv, cd* lines) {
  extern __shared__ cd sm[];
  const int ju = blockIdx.x, g = blockIdx.y;
  const int Kv = 4 * hv + 1;
  const cd* col = f + (long long)g * Ny * Nx + ju;
  cd* out = lines + ((long long)g * Nx + ju) * Kv;
  if (Ny <= SERIES_LDS_MAX) {
    stage_series(col, Nx, Ny, false, sm, sm + Ny);
    __syncthreads();
    for (int kk = threadIdx.x; kk < Kv; kk += blockDim.x)
      out[kk] = point_series_lds(sm, sm + Ny, Ny, kk - 2 * hv);
  } else {
    for (int kk = threadIdx.x; kk < Kv; kk += blockDim.x)
      out[kk] = point_series_global(col, Nx, Ny, kk - 2 * hv);
  }
}

// c2d[g][kv][ku] = u-series over ju of lines[g][ju][kv]
__global__ __launch_bounds__(64) void k_point_u(const cd* lines, int Nx, int hu, int hv, cd* c2d) {
  extern __shared__ cd sm[];
  const int kv = blockIdx.x, g = blockIdx.y;
  const int Kv = 4 * hv + 1, Ku = 4 * hu + 1;
  const cd* row = lines + ((long long)g * Nx) * Kv + kv;
  cd* out = c2d + ((long long)g * Kv + kv) * Ku;
  if (Nx <= SERIES_LDS_MAX) {
    stage_series(row, Kv, Nx, false, sm, sm + Nx);
    __syncthreads();
    for (int kk = threadIdx.x; kk < Ku; kk += blockDim.x)
      out[kk] = point_series_lds(sm, sm + Nx, Nx, kk - 2 * hu);
  } else {
    for (int kk = threadIdx.x; kk < Ku; kk += blockDim.x)
      out[kk] = point_series_global(row, Kv, Nx, kk - 2 * hu);
  }
}

__device__ inline bool cne(cd a, cd b) { return a.x != b.x || a.y != b.y; }

// Interface mask of the lifted raster (Ny rv x Nx ru; sample (J, I) holds eps[J / rv][I / ru]):
// a sample differing from one of its four neighbours, neighbours wrapping
// (factorization.py:1467-1481).
__global__ void k_nvf_mask(const cd* eps, int L, int Ny, int Nx, int rv, int ru, unsigned char* mask) {
  const int Y = Ny * rv, X = Nx * ru;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)L * Y * X) return;
  const int l = (int)(t / ((long long)Y * X));
  const int J = (int)((t / X) % Y), I = (int)(t % X);
  const cd* e = eps + (long long)l * Ny * Nx;
  auto at = [&](int j, int i) { return e[(long long)(j / rv) * Nx + i / ru]; };
  const cd v = at(J, I);
  const bool m = cne(v, at((J + 1) % Y, I)) || cne(v, at((J + Y - 1) % Y, I)) ||
                 cne(v, at(J, (I + 1) % X)) || cne(v, at(J, (I + X - 1) % X));
  mask[t] = m ? 1 : 0;
}

// Column pass: g[J][I] = (sv d)^2 with d the distance, around the period, from row J to the
// nearest mask row of column I (+inf for a column without one). One thread per (layer, column):
// a forward and a backward sweep over two laps of the column.
__global__ void k_nvf_edt_cols(const unsigned char* mask, int L, int Y, int X, double sv, double* g) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)L * X) return;
  const int l = (int)(t / X), I = (int)(t % X);
  const unsigned char* m = mask + (long long)l * Y * X + I;
  double* out = g + (long long)l * Y * X + I;
  const int NONE = -1;
  int last = NONE;
  for (int s = 0; s < 2 * Y; ++s) {                 // forward: nearest mask at or above
    const int J = s < Y ? s : s - Y;
    if (m[(long long)J * X]) last = s;
    if (s >= Y) out[(long long)J * X] = (last == NONE) ? HUGE_VAL : (double)(s - last);
  }
  last = NONE;
  for (int s = 2 * Y - 1; s >= 0; --s) {            // backward: nearest mask at or below
    const int J = s < Y ? s : s - Y;
    if (m[(long long)J * X]) last = s;
    if (s < Y) {
      const double up = (last == NONE) ? HUGE_VAL : (double)(last - s);
      const double d = fmin(out[(long long)J * X], up);
      if (d == HUGE_VAL) {
        out[(long long)J * X] = HUGE_VAL;
      } else {
        const double a = d * sv;
        out[(long long)J * X] = __dmul_rn(a, a);
      }
    }
  }
}

// Row pass: dist[J][I] = sqrt(min over I' of g[J][I'] + (su d(I, I'))^2), d around the period,
// the sum in the order (v term) + (u term) without contraction. One workgroup per (layer, row),
// the row's g in LDS when it fits.
__global__ __launch_bounds__(256) void k_nvf_edt_rows(const double* g, int L, int Y, int X, double su,
                                                      double* dist) {
  extern __shared__ double gs[];
  const int J = blockIdx.x, l = blockIdx.y;
  const double* row = g + ((long long)l * Y + J) * X;
  const bool lds = X <= 8192;
  if (lds)
    for (int i = threadIdx.x; i < X; i += blockDim.x) gs[i] = row[i];
  __syncthreads();
  const double* src = lds ? gs : row;
  for (int I = threadIdx.x; I < X; I += blockDim.x) {
    double best = HUGE_VAL;
    for (int i2 = 0; i2 < X; ++i2) {
      const double gv = src[i2];
      if (gv == HUGE_VAL) continue;
      int d = I - i2;
      if (d < 0) d = -d;
      if (X - d < d) d = X - d;
      const double a = (double)d * su;
      const double v = __dadd_rn(gv, __dmul_rn(a, a));
      best = fmin(best, v);
    }
    dist[((long long)l * Y + J) * X + I] = (best == HUGE_VAL) ? HUGE_VAL : sqrt(best);
  }
}

// The normal field and its three products n_x^2, n_x n_y, n_y^2 as complex samples:
// prod[slot][l][J][I]. cst: the constant normal (ux, uy) of a raster with a one-cell axis.
// A layer with no interface (every distance infinite) gets the zero field.
__global__ void k_nvf_normal(const double* dist, int L, int Y, int X, double jt00, double jt01,
                             double jt10, double jt11, int cst, double ux, double uy, cd* prod) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long plane = (long long)L * Y * X;
  if (t >= plane) return;
  double nx, ny;
  if (cst) {
    nx = ux; ny = uy;
  } else {
    const int l = (int)(t / ((long long)Y * X));
    const int J = (int)((t / X) % Y), I = (int)(t % X);
    const double* d = dist + (long long)l * Y * X;
    const double dJp = d[(long long)((J + 1) % Y) * X + I], dJm = d[(long long)((J + Y - 1) % Y) * X + I];
    const double dIp = d[(long long)J * X + (I + 1) % X], dIm = d[(long long)J * X + (I + X - 1) % X];
    if (dJp == HUGE_VAL || dJm == HUGE_VAL || dIp == HUGE_VAL || dIm == HUGE_VAL) {
      nx = 0.0; ny = 0.0;
    } else {
      const double gj = (dJp - dJm) / 2.0, gi = (dIp - dIm) / 2.0;
      const double gx = __dadd_rn(__dmul_rn(jt00, gi), __dmul_rn(jt01, gj));
      const double gy = __dadd_rn(__dmul_rn(jt10, gi), __dmul_rn(jt11, gj));
      double norm = hypot(gx, gy);
      if (norm < NVF_FLAT_PIXEL) norm = 1.0;
      nx = gx / norm; ny = gy / norm;
    }
  }
  prod[t] = mkc(__dmul_rn(nx, nx), 0.0);
  prod[plane + t] = mkc(__dmul_rn(nx, ny), 0.0);
  prod[2 * plane + t] = mkc(__dmul_rn(ny, ny), 0.0);
}

// D = E - A in place over n complex entries (D holds A)
__global__ void k_sub_from(const cd* E, cd* D, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) D[i] = csub(E[i], D[i]);
}

// out = E - (X + Y) / 2 (sub), or (X + Y) / 2 (!sub), over n complex entries
__global__ void k_nvf_close(const cd* E, const cd* X, const cd* Y, int sub, cd* out, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const cd h = cscale(cadd(X[i], Y[i]), 0.5);
  out[i] = sub ? csub(E[i], h) : h;
}

// active[l] = 1 when layer l holds a value that is not a passive dielectric (Re eps <= 0 or
// Im eps < 0): the reference floors the singular values of [[1/eps]] there
// (factorization.py:1968-2160), which this builder does not.
__global__ void k_nvf_passive(const cd* eps, int L, long long per, int* active) {
  const int l = blockIdx.y;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= per) return;
  const cd v = eps[(long long)l * per + i];
  if (!(v.x > 0.0 && v.y >= 0.0)) atomicOr(active + l, 1);
}

// f[l] = 1 / (INVERSE_RULE_BOUND max|eps_l|) on a layer flagged active (not a passive
// dielectric), 0 elsewhere (factorization.py:1968-2043: the floor of [[1/eps]]'s singular
// values, bound 2). One workgroup per layer.
__global__ __launch_bounds__(256) void k_nvf_floor_value(const cd* eps, long long per, const int* active,
                                                         double* f) {
  __shared__ double red[256];
  const int l = blockIdx.x;
  double m = 0.0;
  for (long long i = threadIdx.x; i < per; i += blockDim.x) {
    const cd v = eps[(long long)l * per + i];
    m = fmax(m, hypot(v.x, v.y));
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) f[l] = active[l] ? 1.0 / (2.0 * red[0]) : 0.0;
}

// c[l][k] = f / s_k - 1 for every singular value s_k = sqrt(eigenvalue k of B^H B) below the
// layer's floor f, else 0: B + (B Z) diag(c) Z^H is U max(S, f) V^H with V = Z.
__global__ void k_nvf_floor_gain(int N, int L, const cd* w, const double* f, double* c) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)L * N) return;
  const int l = (int)(t / N);
  const double s = sqrt(fmax(w[t].x, 0.0));
  const double fl = f[l];
  c[t] = (s < fl) ? fl / s - 1.0 : 0.0;
}

// Y[l][:, k] *= c[l][k]
__global__ void k_nvf_scale_cols(int N, int L, cd* Y, const double* c) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)L * N * N) return;
  const long long l = t / ((long long)N * N);
  const int k = (int)(t % N);
  Y[t] = cscale(Y[t], c[l * N + k]);
}

}  // namespace

// The NVF builder's workspace: the series, the normal field's buffers, the layer matrices
// (E, Delta, the three projectors C_xx | C_xy | C_yy, two products) and the floor's scratch.
struct NvfWs {
  cd *lines, *c2d, *prod, *plines, *pc2d, *E, *D, *C, *T1, *T2, *fw, *ftau;
  unsigned char* mask;
  double *gcol, *dist, *fval, *fgain;
  int *ipiv, *finfo;
  void* frot;
};
static NvfWs nvf_carve(void* ws, int L, int Ny, int Nx, int hu, int hv) {
  const int nv = 2 * hv + 1, nu = 2 * hu + 1, N = nu * nv;
  const int Kv = 4 * hv + 1, Ku = 4 * hu + 1;
  const int rv = (Ny < Kv) ? Kv / Ny + 1 : 1, ru = (Nx < Ku) ? Ku / Nx + 1 : 1;
  const int Y = Ny * rv, X = Nx * ru;
  const long long P = (long long)L * Y * X;
  const long long NN = (long long)N * N;
  char* p = (char*)ws;
  auto take = [&](size_t bytes) { char* r = p; p += (bytes + 255) / 256 * 256; return r; };
  NvfWs w;
  w.lines = (cd*)take(sizeof(cd) * 2 * (size_t)L * Nx * Kv);
  w.c2d = (cd*)take(sizeof(cd) * 2 * (size_t)L * Kv * Ku);
  w.mask = (unsigned char*)take((size_t)P);
  w.gcol = (double*)take(sizeof(double) * (size_t)P);
  w.dist = (double*)take(sizeof(double) * (size_t)P);
  w.prod = (cd*)take(sizeof(cd) * 3 * (size_t)P);
  w.plines = (cd*)take(sizeof(cd) * 3 * (size_t)L * X * Kv);
  w.pc2d = (cd*)take(sizeof(cd) * 3 * (size_t)L * Kv * Ku);
  w.E = (cd*)take(sizeof(cd) * (size_t)L * NN);
  w.D = (cd*)take(sizeof(cd) * (size_t)L * NN);
  w.C = (cd*)take(sizeof(cd) * 3 * (size_t)L * NN);
  w.T1 = (cd*)take(sizeof(cd) * (size_t)L * NN);
  w.T2 = (cd*)take(sizeof(cd) * (size_t)L * NN);
  w.ipiv = (int*)take(sizeof(int) * 2 * (size_t)L * N);
  w.fw = (cd*)take(sizeof(cd) * (size_t)L * N);
  w.ftau = (cd*)take(sizeof(cd) * (size_t)L * N);
  w.frot = (void*)take(wb_hqr_ms_rot_bytes(N, L));
  w.finfo = (int*)take(sizeof(int) * (size_t)L);
  w.fval = (double*)take(sizeof(double) * (size_t)L);
  w.fgain = (double*)take(sizeof(double) * (size_t)L * N);
  return w;
}

size_t wb_assemble_nvf_workspace_bytes(int L, int Ny, int Nx, int hu, int hv) {
  const size_t nv = 2 * hv + 1, nu = 2 * hu + 1, N = nu * nv;
  const size_t Kv = 4 * hv + 1, Ku = 4 * hu + 1;
  const size_t rv = (Ny < (int)Kv) ? Kv / Ny + 1 : 1, ru = (Nx < (int)Ku) ? Ku / Nx + 1 : 1;
  const size_t Y = Ny * rv, X = Nx * ru, P = (size_t)L * Y * X;
  size_t b = 0;
  b += sizeof(cd) * 2 * L * Nx * Kv + 256;      // column series of eps and 1/eps
  b += sizeof(cd) * 2 * L * Kv * Ku + 256;      // their 2D coefficients
  b += P + 256;                                 // interface mask
  b += sizeof(double) * 2 * P + 512;            // column pass, distances
  b += sizeof(cd) * 3 * P + 256;                // the three projector products
  b += sizeof(cd) * 3 * L * X * Kv + 256;       // their column series
  b += sizeof(cd) * 3 * L * Kv * Ku + 256;      // their 2D coefficients
  b += sizeof(cd) * 7 * L * N * N + 7 * 256;    // E, Delta, C_xx, C_xy, C_yy, two products
  b += sizeof(int) * 2 * L * N + 256;           // pivots
  // the singular-value floor: eigenvalues, reflector scalars, rotation staging, statuses, the
  // floor and the gains (its n x n scratch is the closing's buffers, free at that point)
  b += sizeof(cd) * 2 * L * N + 512;
  b += wb_hqr_ms_rot_bytes((int)N, L) + 256;
  b += sizeof(int) * L + sizeof(double) * L + sizeof(double) * L * N + 768;
  return b + 1024;
}

hipError_t wb_assemble_nvf_launch(const cd* eps, int L, int Ny, int Nx, int hu, int hv,
                                  const double* metric, int cst, double ux, double uy, int floor,
                                  cd* epx, cd* epy, cd* epz_i, cd* epxy, void* ws, int* info,
                                  hipStream_t stream) {
  const double* jt = metric;
  const double step_v = metric[4], step_u = metric[5];
  const int nv = 2 * hv + 1, nu = 2 * hu + 1, N = nu * nv;
  const int Kv = 4 * hv + 1, Ku = 4 * hu + 1;
  const int rv = (Ny < Kv) ? Kv / Ny + 1 : 1, ru = (Nx < Ku) ? Ku / Nx + 1 : 1;
  const int Y = Ny * rv, X = Nx * ru;
  const long long P = (long long)L * Y * X;
  const long long NN = (long long)N * N;
  const NvfWs w = nvf_carve(ws, L, Ny, Nx, hu, hv);
  cd *lines = w.lines, *c2d = w.c2d, *prod = w.prod, *plines = w.plines, *pc2d = w.pc2d;
  unsigned char* mask = w.mask;
  double *gcol = w.gcol, *dist = w.dist, *fval = w.fval, *fgain = w.fgain;
  cd *E = w.E, *D = w.D, *C = w.C, *T1 = w.T1, *T2 = w.T2, *fw = w.fw, *ftau = w.ftau;
  int *ipiv = w.ipiv, *finfo = w.finfo;
  void* frot = w.frot;
  const cd one = mkc(1.0, 0.0), zero = mkc(0.0, 0.0);
  const dim3 blk(256);
  auto grid1 = [](long long n) { return dim3((unsigned)((n + 255) / 256)); };
  hipError_t e;
  {
    WbStageTimer timer(WB_T_ASSEMBLE, stream, 0.0, WB_STAGE_BOTH);
    // E and [[1/eps]] from the raster itself (a block repeat has the same staircase series)
    if ((e = series_v_launch(eps, L, Ny, Nx, hv, lines, stream)) != hipSuccess) return e;
    if ((e = series_2d_launch(lines, L, Nx, hu, hv, c2d, stream)) != hipSuccess) return e;
    if ((e = series_2d_launch(lines + (size_t)L * Nx * Kv, L, Nx, hu, hv, c2d + (size_t)L * Kv * Ku,
                              stream)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_direct_bttb, dim3((unsigned)((NN + 255) / 256), L), blk, 0, stream, c2d, L, hu,
                       hv, E, T1, (cd*)nullptr);
    WB_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_direct_bttb, dim3((unsigned)((NN + 255) / 256), L), blk, 0, stream,
                       c2d + (size_t)L * Kv * Ku, L, hu, hv, T2, (cd*)nullptr, (cd*)nullptr);
    WB_CHECK_LAUNCH();
    // the normal field on the lifted raster and its three products
    if (!cst) {
      hipLaunchKernelGGL(k_nvf_mask, grid1(P), blk, 0, stream, eps, L, Ny, Nx, rv, ru, mask);
      WB_CHECK_LAUNCH();
      hipLaunchKernelGGL(k_nvf_edt_cols, grid1((long long)L * X), blk, 0, stream, mask, L, Y, X, step_v,
                         gcol);
      WB_CHECK_LAUNCH();
      hipLaunchKernelGGL(k_nvf_edt_rows, dim3(Y, L), blk, X <= 8192 ? X * sizeof(double) : 0, stream, gcol,
                         L, Y, X, step_u, dist);
      WB_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_nvf_normal, grid1(P), blk, 0, stream, dist, L, Y, X, jt[0], jt[1], jt[2], jt[3],
                       cst, ux, uy, prod);
    WB_CHECK_LAUNCH();
    // C_xx, C_xy, C_yy: point-sample series, then the direct rule's layout
    const size_t lv = Y <= SERIES_LDS_MAX ? 2 * Y * sizeof(cd) : 0;
    hipLaunchKernelGGL(k_point_v, dim3(X, 3 * L), dim3(64), lv, stream, prod, Y, X, hv, plines);
    WB_CHECK_LAUNCH();
    const size_t lu = X <= SERIES_LDS_MAX ? 2 * X * sizeof(cd) : 0;
    hipLaunchKernelGGL(k_point_u, dim3(Kv, 3 * L), dim3(64), lu, stream, plines, X, hu, hv, pc2d);
    WB_CHECK_LAUNCH();
  }
  // info: [0, L) the inversions of E, [L, 2L) those of [[1/eps]], [2L, 3L) non-passive layers
  if ((e = wb_zero_int_launch(info + 2 * L, L, stream)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_nvf_passive, dim3((unsigned)(((long long)Ny * Nx + 255) / 256), L), blk, 0,
                     stream, eps, L, (long long)Ny * Nx, info + 2 * L);
  WB_CHECK_LAUNCH();
  if (floor) {
    // [[1/eps]] (T2) with its singular values below 1 / (2 max|eps|) raised to that floor on the
    // non-passive layers (factorization.py:2046-2160): the Schur vectors Z of the Hermitian
    // B^H B are B's right singular vectors, and B + (B Z) diag(f / s - 1 where s < f) Z^H is
    // U max(S, f) V^H; a layer without a value below the floor (every passive one) gets B back.
    // Scratch: D (B^H B, then T), C (the Schur stage's W and X, then B Z), C + 2 L NN (Z).
    hipLaunchKernelGGL(k_nvf_floor_value, dim3(L), dim3(256), 0, stream, eps, (long long)Ny * Nx,
                       info + 2 * L, fval);
    WB_CHECK_LAUNCH();
    if ((e = wb_zgemm_ops_launch(1, 0, N, N, N, one, bp_strided(T2, NN), N, bp_strided(T2, NN), N, zero,
                                 bp_strided(D, NN), N, L, stream))) return e;
    cd* Zs = C + 2 * (size_t)L * NN;
    int zh = 0;
    if ((e = wb_schur_launch(N, D, fw, C, Zs, C + (size_t)L * NN, ftau, frot, finfo, L, &zh, stream)))
      return e;
    hipLaunchKernelGGL(k_nvf_floor_gain, grid1((long long)L * N), blk, 0, stream, N, L, fw, fval, fgain);
    WB_CHECK_LAUNCH();
    if ((e = wb_zgemm_ops_launch(0, zh ? 1 : 0, N, N, N, one, bp_strided(T2, NN), N, bp_strided(Zs, NN), N,
                                 zero, bp_strided(C, NN), N, L, stream))) return e;
    hipLaunchKernelGGL(k_nvf_scale_cols, grid1((long long)L * NN), blk, 0, stream, N, L, C, fgain);
    WB_CHECK_LAUNCH();
    if ((e = wb_zgemm_ops_launch(0, zh ? 0 : 1, N, N, N, one, bp_strided(C, NN), N, bp_strided(Zs, NN), N,
                                 one, bp_strided(T2, NN), N, L, stream))) return e;
  }
  hipLaunchKernelGGL(k_direct_bttb, dim3((unsigned)((NN + 255) / 256), 3 * L), blk, 0, stream, pc2d,
                     3 * L, hu, hv, C, (cd*)nullptr, (cd*)nullptr);
  WB_CHECK_LAUNCH();
  // epz_i = E^-1 (from the copy in T1), A = [[1/eps]]^-1 into D, Delta = E - A
  if ((e = wb_inverse_launch(N, bp_strided(T1, NN), N, bp_strided(epz_i, NN), N, ipiv, info, L, stream)))
    return e;
  if ((e = wb_inverse_launch(N, bp_strided(T2, NN), N, bp_strided(D, NN), N, ipiv + (size_t)L * N,
                             info + L, L, stream)))
    return e;
  hipLaunchKernelGGL(k_sub_from, grid1((long long)L * NN), blk, 0, stream, E, D, (long long)L * NN);
  WB_CHECK_LAUNCH();
  // the closing, one projector at a time: T1 = Delta C, T2 = C Delta, then the slot
  cd* outs[3] = {epx, epxy, epy};
  const int subs[3] = {1, 0, 1};
  for (int s = 0; s < 3; ++s) {
    const cd* Cs = C + (size_t)s * L * NN;
    if ((e = wb_zgemm_launch(N, N, N, one, bp_strided(D, NN), N, bp_strided(Cs, NN), N, zero,
                             bp_strided(T1, NN), N, L, stream))) return e;
    if ((e = wb_zgemm_launch(N, N, N, one, bp_strided(Cs, NN), N, bp_strided(D, NN), N, zero,
                             bp_strided(T2, NN), N, L, stream))) return e;
    hipLaunchKernelGGL(k_nvf_close, grid1((long long)L * NN), blk, 0, stream, E, T1, T2, subs[s], outs[s],
                       (long long)L * NN);
    WB_CHECK_LAUNCH();
  }
  return hipSuccess;
}

// ------------------------------------------------------------------- NVF adjoint
// Cotangent of the raster from those of the NVF builder's outputs, the normal field held fixed
// (it is geometry: the interface mask and the distance map do not move under an infinitesimal
// change of the values, which is how the reference's gradient treats it; only a layer whose
// floor stays inactive -- a passive dielectric -- is differentiated). With the forward's
// E, A = [[1/eps]]^-1, Delta = E - A and the projectors C_ab rebuilt in the workspace:
//   Delta-bar = -(Xb C_xx^H + C_xx^H Xb + Yb C_yy^H + C_yy^H Yb)/2 + (Ob C_xy^H + C_xy^H Ob)/2,
//   E-bar = Xb + Yb + Delta-bar - Z^H Zb Z^H (Z = E^-1),  [[1/eps]]-bar = A^H Delta-bar A^H,
// (Xb, Yb, Zb the cotangents of epx, epy, epz_i; Ob the sum of the two off-diagonal slots'),
// then the direct rule's and the two series' adjoints (eps and 1/eps) as for Laurent's rule.
size_t wb_assemble_nvf_adjoint_workspace_bytes(int L, int Ny, int Nx, int hu, int hv) {
  const size_t N = (size_t)(2 * hu + 1) * (2 * hv + 1), NN = N * N;
  const size_t Kv = 4 * hv + 1, Ku = 4 * hu + 1;
  size_t b = wb_assemble_nvf_workspace_bytes(L, Ny, Nx, hu, hv) + 256;
  b += sizeof(cd) * 7 * (size_t)L * NN + 7 * 256;            // forward outputs, three bars
  b += sizeof(int) * 3 * (size_t)L + 256;                     // the builder's statuses
  b += sizeof(cd) * 2 * (size_t)L * Nx * Kv + 256;            // series cotangents
  b += sizeof(cd) * 2 * (size_t)L * Kv * Ku + 256;
  return b + 1024;
}

namespace {
// out += a + b + c over n complex entries
__global__ void k_acc3(const cd* a, const cd* b, const cd* c, cd* out, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = cadd(out[i], cadd(cadd(a[i], b[i]), c[i]));
}
}  // namespace

hipError_t wb_assemble_nvf_adjoint_launch(const cd* eps, int L, int Ny, int Nx, int hu, int hv,
                                          const double* metric, int cst, double ux, double uy,
                                          const cd* d_epx, const cd* d_epy, const cd* d_epz,
                                          const cd* d_off, cd* d_eps, void* ws, hipStream_t stream) {
  const int nv = 2 * hv + 1, nu = 2 * hu + 1, N = nu * nv;
  const int Kv = 4 * hv + 1, Ku = 4 * hu + 1;
  const long long NN = (long long)N * N, LNN = (long long)L * NN;
  char* p = (char*)ws;
  auto take = [&](size_t bytes) { char* r = p; p += (bytes + 255) / 256 * 256; return r; };
  void* fws = take(wb_assemble_nvf_workspace_bytes(L, Ny, Nx, hu, hv));
  cd* fx = (cd*)take(sizeof(cd) * LNN);
  cd* fy = (cd*)take(sizeof(cd) * LNN);
  cd* fz = (cd*)take(sizeof(cd) * LNN);
  cd* fxy = (cd*)take(sizeof(cd) * LNN);
  cd* Db = (cd*)take(sizeof(cd) * LNN);
  cd* Eb = (cd*)take(sizeof(cd) * LNN);
  cd* Tm = (cd*)take(sizeof(cd) * LNN);
  int* info = (int*)take(sizeof(int) * 3 * (size_t)L);
  cd* lines_bar = (cd*)take(sizeof(cd) * 2 * (size_t)L * Nx * Kv);
  cd* c2d_bar = (cd*)take(sizeof(cd) * 2 * (size_t)L * Kv * Ku);
  hipError_t e;
  // the forward without the floor: E, Delta and the projectors stay in its workspace
  if ((e = wb_assemble_nvf_launch(eps, L, Ny, Nx, hu, hv, metric, cst, ux, uy, 0, fx, fy, fz, fxy, fws,
                                  info, stream)))
    return e;
  const NvfWs w = nvf_carve(fws, L, Ny, Nx, hu, hv);
  const cd* Cxx = w.C;
  const cd* Cxy = w.C + LNN;
  const cd* Cyy = w.C + 2 * LNN;
  auto S = [&](const cd* x) { return bp_strided(x, NN); };
  const cd half = mkc(0.5, 0.0), mhalf = mkc(-0.5, 0.0), one = mkc(1.0, 0.0), mone = mkc(-1.0, 0.0),
           zero = mkc(0.0, 0.0);
  // Delta-bar
  if ((e = wb_zgemm_ops_launch(0, 1, N, N, N, mhalf, S(d_epx), N, S(Cxx), N, zero, S(Db), N, L, stream))) return e;
  if ((e = wb_zgemm_ops_launch(1, 0, N, N, N, mhalf, S(Cxx), N, S(d_epx), N, one, S(Db), N, L, stream))) return e;
  if ((e = wb_zgemm_ops_launch(0, 1, N, N, N, mhalf, S(d_epy), N, S(Cyy), N, one, S(Db), N, L, stream))) return e;
  if ((e = wb_zgemm_ops_launch(1, 0, N, N, N, mhalf, S(Cyy), N, S(d_epy), N, one, S(Db), N, L, stream))) return e;
  if ((e = wb_zgemm_ops_launch(0, 1, N, N, N, half, S(d_off), N, S(Cxy), N, one, S(Db), N, L, stream))) return e;
  if ((e = wb_zgemm_ops_launch(1, 0, N, N, N, half, S(Cxy), N, S(d_off), N, one, S(Db), N, L, stream))) return e;
  // E-bar = Xb + Yb + Delta-bar - Z^H Zb Z^H
  if ((e = wb_zgemm_ops_launch(0, 1, N, N, N, one, S(d_epz), N, S(fz), N, zero, S(Tm), N, L, stream))) return e;
  if ((e = wb_zgemm_ops_launch(1, 0, N, N, N, mone, S(fz), N, S(Tm), N, zero, S(Eb), N, L, stream))) return e;
  hipLaunchKernelGGL(k_acc3, dim3((unsigned)((LNN + 255) / 256)), dim3(256), 0, stream, d_epx, d_epy, Db, Eb,
                     LNN);
  WB_CHECK_LAUNCH();
  // [[1/eps]]-bar = A^H Delta-bar A^H with A = E - Delta (into fy, fx as scratch)
  hipLaunchKernelGGL(k_sub_from, dim3((unsigned)((LNN + 255) / 256)), dim3(256), 0, stream, w.E, w.D, LNN);
  WB_CHECK_LAUNCH();
  if ((e = wb_zgemm_ops_launch(0, 1, N, N, N, one, S(Db), N, S(w.D), N, zero, S(fx), N, L, stream))) return e;
  if ((e = wb_zgemm_ops_launch(1, 0, N, N, N, one, S(w.D), N, S(fx), N, zero, S(fy), N, L, stream))) return e;
  // the direct rule's, then the two series' adjoints
  hipLaunchKernelGGL(k_direct_adj, dim3((Kv * Ku + 255) / 256, L), dim3(256), 0, stream, Eb, L, hu, hv, c2d_bar);
  WB_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_direct_adj, dim3((Kv * Ku + 255) / 256, L), dim3(256), 0, stream, fy, L, hu, hv,
                     c2d_bar + (size_t)L * Kv * Ku);
  WB_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_series2d_adj, dim3(Kv, L), dim3(64), 0, stream, c2d_bar, L, Nx, hu, hv, lines_bar, 0);
  WB_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_series2d_adj, dim3(Kv, L), dim3(64), 0, stream, c2d_bar + (size_t)L * Kv * Ku, L, Nx,
                     hu, hv, lines_bar + (size_t)L * Nx * Kv, 0);
  WB_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_series_v_adj, dim3(Nx, L), dim3(64), 0, stream, eps, lines_bar, L, Ny, Nx, hv, d_eps, 1);
  WB_CHECK_LAUNCH();
  return hipSuccess;
}
