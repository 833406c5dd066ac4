// Complex plane rotations of the QR iteration (zlartg's scheme), shared by the device code and a
// host-compiled unit check of its edge cases.
#pragma once
#include <float.h>
#include <math.h>

#include "wb_common.h"

// 1 / sqrt(x) for a positive normal x: on the device the hardware estimate refined by two
// Newton steps (y <- y + y (1 - x y^2) / 2; the estimate's ~2^-26 error falls to the last bit),
// a chain of seven dependent operations where a square root and a division take some thirty.
__host__ __device__ inline double rsqrt_pos(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double y = __builtin_amdgcn_rsq(x);
  double e = fma(-x, y * y, 1.0);
  y = fma(0.5 * y, e, y);
  e = fma(-x, y * y, 1.0);
  return fma(0.5 * y, e, y);
#else
  return 1.0 / sqrt(x);
#endif
}

// Plane rotation Q = [c s; -conj(s) c] (c real, c^2 + |s|^2 = 1) with Q [f; g] = [r; 0]. A QR
// step applies it as the similarity H <- Q H Q^H (rows k, k+1 from the left, columns k, k+1
// from the right) and Z <- Z Q^H. Three doubles: a rotation costs 12 FP64 operations per pair
// of entries (a 2-element complex reflector with complex tau costs 14), and the subdiagonal
// entries it produces are complex in general.
struct Rot {
  double c;
  double sx, sy;
};

// Q [f; g] = [r; 0] (zlartg): c = |f| / d, s = (f / |f|) conj(g) / d, r = (f / |f|) d with
// d = sqrt(|f|^2 + |g|^2). f and g are scaled by a power of two (exactly) so the larger
// component lies in [1, 2); then with q = 1 / (|f| d): c = |f|^2 q, s = f conj(g) q and
// r = f d^2 q, one reciprocal square root (rsqrt_pos) on the rotation's serial path. When |f|^2
// underflows at that common scale (|f| below ~1e-154 of |g|), f is scaled on its own: its phase
// f / |f|, c = |f| / |g| (d = |g| to working precision), s = (f / |f|) conj(g) / |g|,
// r = (f / |f|) |g|.
__host__ __device__ inline Rot givens(cd f, cd g, cd* r) {
  Rot q;
  if (g.x == 0.0 && g.y == 0.0) {
    q.c = 1.0; q.sx = 0.0; q.sy = 0.0;
    *r = f;
    return q;
  }
  if (f.x == 0.0 && f.y == 0.0) {
    const double gn = cabs(g);
    q.c = 0.0; q.sx = g.x / gn; q.sy = -g.y / gn;
    *r = mkc(gn, 0.0);
    return q;
  }
  const int e = ilogb(fmax(fmax(fabs(f.x), fabs(f.y)), fmax(fabs(g.x), fabs(g.y))));
  const cd fs = mkc(ldexp(f.x, -e), ldexp(f.y, -e)), gs = mkc(ldexp(g.x, -e), ldexp(g.y, -e));
  const double f2 = cnorm2(fs);
  if (f2 < DBL_MIN) {
    const int ef = ilogb(fmax(fabs(f.x), fabs(f.y)));
    const cd ff = mkc(ldexp(f.x, -ef), ldexp(f.y, -ef));
    const double fn = sqrt(cnorm2(ff)), gn = sqrt(cnorm2(gs));
    const cd ph = mkc(ff.x / fn, ff.y / fn);                 // f / |f|
    q.c = ldexp(fn / gn, ef - e);
    q.sx = (ph.x * gs.x + ph.y * gs.y) / gn;                 // (f / |f|) conj(g) / |g|
    q.sy = (ph.y * gs.x - ph.x * gs.y) / gn;
    *r = mkc(ldexp(ph.x * gn, e), ldexp(ph.y * gn, e));
    return q;
  }
  const double d2 = f2 + cnorm2(gs);
  const double iq = rsqrt_pos(f2 * d2);         // 1 / (|f| d)
  q.c = f2 * iq;
  q.sx = (fs.x * gs.x + fs.y * gs.y) * iq;      // f conj(g) q
  q.sy = (fs.y * gs.x - fs.x * gs.y) * iq;
  const double t = d2 * iq;                      // d / |f|
  *r = mkc(ldexp(fs.x * t, e), ldexp(fs.y * t, e));
  return q;
}
