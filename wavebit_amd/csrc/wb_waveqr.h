// Single-wave complex Schur iteration shared by the batched QR (wb_hqr.hip) and the fused
// small-order solve (wb_small.hip): plane rotations, their generation, and the single-shift
// iteration of an active block held in LDS by one wave (the zlahqr scheme; Golub & Van Loan 7.5).
//
// The iteration (wave_qr) also deflates a subdiagonal at or below dfl = ulp ||H||_F (the
// Frobenius norm taken once before the iteration; it is invariant under the unitary similarity)
// once the bottom subdiagonal it drives down has stagnated there. Setting such an entry to zero
// perturbs H by ulp ||H|| in norm, the size of the rounding every rotation already commits, so
// the Schur form stays normwise backward stable. Without it an eigenvalue far below ||H|| in
// magnitude must reach a subdiagonal below ulp times its own size, which rounding at ulp ||H||
// does not allow, and the iteration stalls until exceptional shifts (a 21-harmonic grating
// operator with eigenvalues from -11 to 159: 205 single-shift iterations without the floor, 42
// with it, eigenvalues within 1.4e-13 of the Hermitian solver's). Only a stagnated entry takes
// the floor: while the entry still converges (a tiny eigenvalue, or a cluster at a grazing order,
// computed to relative accuracy) the LAPACK tests alone decide.
#pragma once
#include "wb_common.h"
#include <float.h>

namespace {

constexpr double ULP = DBL_EPSILON;

// The sweep and the band updates are separate (non-inlined) functions: these assumptions tell
// the compiler which pointers are LDS or global, so their accesses use LDS / global
// instructions rather than flat ones.
#if defined(__HIP_DEVICE_COMPILE__)
#define ASSUME_LDS(p) __builtin_assume(__builtin_amdgcn_is_shared((const void*)(p)))
#define ASSUME_GLOBAL(p)                                              \
  do {                                                                \
    __builtin_assume(!__builtin_amdgcn_is_shared((const void*)(p)));  \
    __builtin_assume(!__builtin_amdgcn_is_private((const void*)(p))); \
  } while (0)
#else
#define ASSUME_LDS(p) ((void)0)
#define ASSUME_GLOBAL(p) ((void)0)
#endif

constexpr double SAFMIN = DBL_MIN;

#include "wb_givens.h"

// left application (rows k, k+1): [a0; a1] <- Q [a0; a1]: a0 = c a0 + s a1, a1 = c a1 - conj(s) a0
__device__ inline void rot_left(const Rot& q, cd& a0, cd& a1) {
  const double c = q.c, sx = q.sx, sy = q.sy;
  const double b0x = fma(sx, a1.x, fma(-sy, a1.y, c * a0.x));
  const double b0y = fma(sx, a1.y, fma(sy, a1.x, c * a0.y));
  const double b1x = fma(-sx, a0.x, fma(-sy, a0.y, c * a1.x));
  const double b1y = fma(-sx, a0.y, fma(sy, a0.x, c * a1.y));
  a0 = mkc(b0x, b0y);
  a1 = mkc(b1x, b1y);
}
// right application (columns k, k+1): [a0, a1] <- [a0, a1] Q^H: a0 = c a0 + conj(s) a1,
// a1 = c a1 - s a0
__device__ inline void rot_right(const Rot& q, cd& a0, cd& a1) {
  const double c = q.c, sx = q.sx, sy = q.sy;
  const double b0x = fma(sx, a1.x, fma(sy, a1.y, c * a0.x));
  const double b0y = fma(sx, a1.y, fma(-sy, a1.x, c * a0.y));
  const double b1x = fma(-sx, a0.x, fma(sy, a0.y, c * a1.x));
  const double b1y = fma(-sx, a0.y, fma(-sy, a0.x, c * a1.y));
  a0 = mkc(b0x, b0y);
  a1 = mkc(b1x, b1y);
}

// givens() without its scaling where none is needed: with max(|Re f|, |Im f|) in
// [2^-250, 2^250] and g's components at most 2^250, |f|^2 (|f|^2 + |g|^2) stays a normal double, so
// c = |f|^2 q, s = f conj(g) q, r = f d^2 q with q = 1 / (|f| d) need no exponent handling; other
// inputs (f = 0, extreme magnitudes) take givens(). f and g are the same on every lane, so the
// wave takes one branch.
__device__ inline Rot givens_fast(cd f, cd g, cd* r) {
  const double fm = fmax(fabs(f.x), fabs(f.y)), gm = fmax(fabs(g.x), fabs(g.y));
  if (fm >= 0x1p-250 && fm <= 0x1p250 && gm <= 0x1p250) {
    const double f2 = cnorm2(f), d2 = f2 + cnorm2(g);
    const double iq = rsqrt_pos(f2 * d2);
    Rot q;
    q.c = f2 * iq;
    q.sx = (f.x * g.x + f.y * g.y) * iq;
    q.sy = (f.y * g.x - f.x * g.y) * iq;
    const double t = d2 * iq;
    *r = mkc(f.x * t, f.y * t);
    return q;
  }
  return givens(f, g, r);
}


// Lane `src` (wave-uniform) of a complex value, to every lane of the wave.
__device__ inline cd readlane_c(cd x, int src) {
  const unsigned long long rb = __double_as_longlong(x.x), ib = __double_as_longlong(x.y);
  const int r0 = __builtin_amdgcn_readlane((int)(rb & 0xffffffffull), src);
  const int r1 = __builtin_amdgcn_readlane((int)(rb >> 32), src);
  const int i0 = __builtin_amdgcn_readlane((int)(ib & 0xffffffffull), src);
  const int i1 = __builtin_amdgcn_readlane((int)(ib >> 32), src);
  return mkc(__longlong_as_double(((unsigned long long)(unsigned)r1 << 32) | (unsigned)r0),
             __longlong_as_double(((unsigned long long)(unsigned)i1 << 32) | (unsigned)i0));
}

// Active blocks of at most nmin rows. The block's single-shift QR (the zlahqr scheme) runs in
// LDS on one wave with no barriers, every rotation also applied to U (the block's accumulated
// unitary); the rows of H above the block, its columns right of the block and the Schur vectors
// then get U in three small products with all threads. Returns 0, or i + 1 if eigenvalue i of
// the block did not converge. Once the bottom subdiagonal T(i, i-1) has stagnated (after two
// iterations it is at or below dfl, the matrix's normwise floor, and fell by less than half in
// the last one), subdiagonals at or below dfl deflate too.
// WZ: U (LDS, like T) takes every rotation; without it U is not read.
template <bool WZ>
__device__ int wave_qr(cd* T, cd* U, int m, int ld, int n, cd* w, double dfl) {
  ASSUME_LDS(T);
  if constexpr (WZ) ASSUME_LDS(U);
  const int lane = threadIdx.x & 63;
  const double smlnum = SAFMIN * ((double)n / ULP);
  int i = m - 1;
  int kdefl = 0;
  const int itmax = 30 * (m > 10 ? m : 10);
#define TT(r, c) T[(r) * ld + (c)]
  while (i >= 0) {
    int l = 0;
    bool converged = false;
    double prev = 0.0;   // |T(i, i-1)| one iteration back
    for (int its = 0; its <= itmax; ++its) {
      const double hb = (i > 0) ? cabs1(TT(i, i - 1)) : 0.0;
      const double floor_now = (its >= 2 && hb <= dfl && hb > 0.5 * prev) ? dfl : 0.0;
      prev = hb;
      // largest k in (l, i] with a negligible subdiagonal (Ahues-Tisseur), lane per k
      const int k = l + 1 + lane;
      bool small = false;
      if (k <= i) {
        const cd hk = TT(k, k - 1);
        small = cabs1(hk) <= fmax(smlnum, floor_now);
        if (!small) {
          double tst = cabs1(TT(k - 1, k - 1)) + cabs1(TT(k, k));
          if (tst == 0.0) {
            if (k - 2 >= 0) tst += cabs1(TT(k - 1, k - 2));
            if (k + 1 <= m - 1) tst += cabs1(TT(k + 1, k));
          }
          if (cabs1(hk) <= ULP * tst) {
            const double h1 = cabs1(hk), h2 = cabs1(TT(k - 1, k));
            const double ab = fmax(h1, h2), ba = fmin(h1, h2);
            const double g1 = cabs1(TT(k, k));
            const double g2 = cabs1(csub(TT(k - 1, k - 1), TT(k, k)));
            const double aa = fmax(g1, g2), bb = fmin(g1, g2);
            const double sc = aa + ab;
            // after 10 iterations without a deflation the conventional test alone (find_split)
            small = its >= 10 || ba * (ab / sc) <= fmax(smlnum, ULP * (bb * (aa / sc)));
          }
        }
      }
      const unsigned long long mask = __ballot(small);
      l = mask ? l + 1 + (63 - __clzll(mask)) : l;
      if (l > 0 && lane == 0) TT(l, l - 1) = mkc(0.0, 0.0);
      if (l >= i) { converged = true; break; }
      ++kdefl;
      cd sh;
      if (kdefl % 20 == 0) {
        sh = cadd(mkc(0.75 * cabs1(TT(i, i - 1)), 0.0), TT(i, i));
      } else if (kdefl % 10 == 0) {
        sh = cadd(mkc(0.75 * cabs1(TT(l + 1, l)), 0.0), TT(l, l));
      } else {
        sh = TT(i, i);
        const cd u = cmul(csqrt_p(TT(i - 1, i)), csqrt_p(TT(i, i - 1)));
        double su = cabs1(u);
        if (su != 0.0) {
          const cd x = cscale(csub(TT(i - 1, i - 1), sh), 0.5);
          const double sx = cabs1(x);
          su = fmax(su, sx);
          const cd xs = cscale(x, 1.0 / su), us = cscale(u, 1.0 / su);
          cd y = cscale(csqrt_p(cadd(cmul(xs, xs), cmul(us, us))), su);
          if (sx > 0.0) {
            const cd xn = cscale(x, 1.0 / sx);
            if (xn.x * y.x + xn.y * y.y < 0.0) y = cneg(y);
          }
          sh = csub(sh, cmul(u, cdiv(u, cadd(x, y))));
        }
      }
      cd v1 = csub(TT(l, l), sh), v2 = TT(l + 1, l);
      {
        const double sc = cabs1(v1) + cabs1(v2);
        if (sc != 0.0) { v1 = cscale(v1, 1.0 / sc); v2 = cscale(v2, 1.0 / sc); }
      }
      if (m <= 32) {
        // Blocks of at most 32 rows: one rotation application serves both operations. Lanes
        // 0..31 take the row operation (lane j: rows kk, kk+1 of column j >= kk), lanes 32..63 the
        // column operation (lane 32 + r: columns kk, kk+1 of row r <= kk+2); the column operation
        // is the row operation's formula with conj(s). The four entries both touch (rows and
        // columns kk, kk+1) are taken from the row lanes, which finish them with the column
        // rotation after a readlane exchange between lanes kk and kk+1; the column lanes of rows
        // kk, kk+1 are not stored. Loads are unconditional at clamped in-block indices.
        const bool rowh = lane < 32;
        const int q = lane & 31;
        const int qu = q < m ? q : m - 1;
        for (int kk = l; kk < i; ++kk) {
          cd rr;
          const Rot r = givens_fast(v1, v2, &rr);
          const int rmax = (kk + 2 < i) ? kk + 2 : i;
          cd u0, u1;
          if (WZ) { u0 = U[qu * ld + kk]; u1 = U[qu * ld + kk + 1]; }
          const bool act = rowh ? (q >= kk && q < m) : (q <= rmax);
          const int qc = rowh ? qu : (q <= rmax ? q : rmax);
          const int o0 = rowh ? kk * ld + qc : qc * ld + kk;
          const int o1 = rowh ? o0 + ld : o0 + 1;
          cd x0 = T[o0], x1 = T[o1];
          Rot g = r;
          if (!rowh) g.sy = -g.sy;
          rot_left(g, x0, x1);
          // the 2 x 2 block: lanes kk, kk+1 hold columns kk, kk+1 of G B; (G B) G^H takes each
          // column as c times itself plus (conj(s) | -s) times the other
          {
            const cd p00 = readlane_c(x0, kk), p10 = readlane_c(x1, kk);
            const cd p01 = readlane_c(x0, kk + 1), p11 = readlane_c(x1, kk + 1);
            const bool left = (lane == kk);
            const cd o_0 = left ? p01 : p00, o_1 = left ? p11 : p10;
            // h = conj(s) on lane kk, -s on lane kk+1
            const double hx = left ? r.sx : -r.sx, hy = -r.sy;
            const cd y0 = mkc(fma(r.c, x0.x, hx * o_0.x - hy * o_0.y), fma(r.c, x0.y, hx * o_0.y + hy * o_0.x));
            const cd y1 = mkc(fma(r.c, x1.x, hx * o_1.x - hy * o_1.y), fma(r.c, x1.y, hx * o_1.y + hy * o_1.x));
            if (rowh && (lane == kk || lane == kk + 1)) { x0 = y0; x1 = y1; }
          }
          if (act && (rowh || (q != kk && q != kk + 1))) { T[o0] = x0; T[o1] = x1; }
          if (kk > l && lane == 0) {
            TT(kk, kk - 1) = rr;
            TT(kk + 1, kk - 1) = mkc(0.0, 0.0);
          }
          if (WZ) {
            rot_right(r, u0, u1);
            if (lane < m) { U[q * ld + kk] = u0; U[q * ld + kk + 1] = u1; }
          }
          // the next rotation's inputs: T(kk+1, kk) from the row lane kk, T(kk+2, kk) from the
          // column lane of row kk+2
          if (kk + 1 < i) {
            v1 = readlane_c(x1, kk);
            v2 = readlane_c(x0, 32 + kk + 2);
          }
        }
        continue;
      }
      // The chase keeps what it rotates in registers: lane j holds ra = T(kk, j), rb = T(kk+1, j)
      // for the row operation (columns j >= kk), lane r holds ca = T(r, kk), cb = T(r, kk+1) for
      // the column operation (rows r <= kk+2) and u0, u1 = U(r, kk), U(r, kk+1). The 2 x 2 block
      // both operations touch passes from the row to the column layout by readlane, so the
      // serial path of a step is the rotation, two small updates and the readlanes. Each finished
      // row kk and column kk is written back as the chase moves on, and the next row kk+2 and
      // column kk+2 are read from LDS at the end of a step (LDS accesses of one wave stay in
      // program order: a read issued after a write returns it) into trow / tcol / tu, taken into
      // rb / cb / u1 only after the next rotation is formed, so the reads' latency overlaps it.
      // The loads are unconditional at clamped in-block indices; the values are selected after.
      const cd z0 = mkc(0.0, 0.0);
      cd ra = z0, ca = z0, u0 = z0, rb, cb, u1 = z0, trow, tcol, tu = z0, b21;
      {
        const int rmax0 = (l + 2 < i) ? l + 2 : i;
        const int lr = lane < m ? lane : m - 1;
        const int lc = lane <= rmax0 ? lane : rmax0;
        ra = TT(l, lr);
        trow = TT(l + 1, lr);
        b21 = TT(l + 1, l);
        ca = TT(lc, l);
        tcol = TT(lc, l + 1);
        if (WZ) { u0 = U[lr * ld + l]; tu = U[lr * ld + l + 1]; }
      }
      for (int kk = l; kk < i; ++kk) {
        cd rr;
        const Rot r = givens(v1, v2, &rr);
        const int rmax = (kk + 2 < i) ? kk + 2 : i;
        // rows kk+1 and columns kk+1 read at the end of the last step (row kk+1's column kk
        // entry and the column's rows beyond rmax as the layouts need them)
        rb = (lane > kk && lane < m) ? trow : b21;
        cb = (lane <= rmax) ? tcol : z0;
        if (WZ) u1 = tu;
        // row operation on columns kk..m-1
        if (lane >= kk && lane < m) rot_left(r, ra, rb);
        // rows kk, kk+1 of columns kk, kk+1 into the column layout
        {
          const cd a00 = readlane_c(ra, kk), a01 = readlane_c(ra, kk + 1);
          const cd a10 = readlane_c(rb, kk), a11 = readlane_c(rb, kk + 1);
          if (lane == kk) { ca = a00; cb = a01; }
          if (lane == kk + 1) { ca = a10; cb = a11; }
        }
        // column operation on rows 0..rmax, and on U
        if (lane <= rmax) rot_right(r, ca, cb);
        if (WZ && lane < m) rot_right(r, u0, u1);
        // the next rotation's inputs (rows kk+1, kk+2 of column kk)
        if (kk + 1 < i) {
          v1 = readlane_c(ca, kk + 1);
          v2 = readlane_c(ca, kk + 2);
        }
        // row kk beyond the block (its columns kk, kk+1 are in the column layout), the chased
        // entries of column kk-1, then column kk
        if (lane >= kk + 2 && lane < m) TT(kk, lane) = ra;
        if (kk > l && lane == 0) {
          TT(kk, kk - 1) = rr;
          TT(kk + 1, kk - 1) = z0;
        }
        if (lane <= rmax) TT(lane, kk) = ca;
        if (WZ && lane < m) U[lane * ld + kk] = u0;
        if (kk + 1 < i) {
          // step to rows / columns kk+1, kk+2: row kk+1 with its column kk+1 entry from the
          // column operation, column kk+1 (row kk+3 is zero in Hessenberg form); row kk+2 (its
          // column kk+1 entry, b21, likewise from the column operation) and column kk+2 are read
          const cd b11 = readlane_c(cb, kk + 1);
          b21 = readlane_c(cb, kk + 2);
          const int rmax2 = (kk + 3 < i) ? kk + 3 : i;
          trow = TT(kk + 2, lane < m ? lane : m - 1);
          tcol = TT(lane <= rmax2 ? lane : rmax2, kk + 2);
          if (WZ) tu = U[(lane < m ? lane : m - 1) * ld + kk + 2];
          ra = (lane == kk + 1) ? b11 : rb;
          ca = (lane <= rmax) ? cb : z0;
          u0 = u1;
        } else {
          // the last step: row i beyond the block, column i, U column i
          if (lane >= kk + 2 && lane < m) TT(kk + 1, lane) = rb;
          if (lane <= rmax) TT(lane, kk + 1) = cb;
          if (WZ && lane < m) U[lane * ld + kk + 1] = u1;
        }
      }
    }
    if (!converged) return i + 1;
    if (lane == 0) w[i] = TT(i, i);
    kdefl = 0;
    i = l - 1;
  }
#undef TT
  return 0;
}

}  // namespace
