// Host-compiled check of the QR rotation generator (wavebit_amd/csrc/wb_givens.h) on edge cases:
// prints, per case, |c^2 + |s|^2 - 1|, the residual of Q [f; g] = [r; 0] relative to |(f, g)|,
// and whether every output is finite. Built and run by tests/test_givens_host.py.
#include <cstdio>
#include <cmath>

#include "../wavebit_amd/csrc/wb_givens.h"

int main() {
  const double cases[][4] = {
      {1e-200, 0.0, 1.0, 0.0},      {1e-200, -3e-201, 0.5, 2.0}, {3e-160, 1e-160, 1.0, 0.0},
      {1e-300, 1e-300, 1e10, -1e10}, {0.3, -0.2, 1e-200, 0.0},    {1.0, 2.0, -3.0, 0.5},
      {2.5e-308, 0.0, 1.0, 1.0},    {1e-20, 0.0, 1e140, 0.0}};
  for (const auto& v : cases) {
    const cd f = mkc(v[0], v[1]), g = mkc(v[2], v[3]);
    cd r;
    const Rot q = givens(f, g, &r);
    const double unit = std::fabs(q.c * q.c + q.sx * q.sx + q.sy * q.sy - 1.0);
    // row 1: c f + s g ; row 2: -conj(s) f + c g
    const cd s = mkc(q.sx, q.sy), sc = mkc(q.sx, -q.sy);
    const cd top = cadd(cscale(f, q.c), cmul(s, g));
    const cd bot = csub(cscale(g, q.c), cmul(sc, f));
    const double scale = std::hypot(cabs(f), cabs(g));
    const double res = (cabs(csub(top, r)) + cabs(bot)) / scale;
    const bool finite = std::isfinite(q.c) && std::isfinite(q.sx) && std::isfinite(q.sy) &&
                        std::isfinite(r.x) && std::isfinite(r.y);
    std::printf("%.3e %.3e %d\n", unit, res, finite ? 1 : 0);
  }
  return 0;
}
