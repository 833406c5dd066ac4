// Shared device helpers for the wavebit_amd HIP kernels (gfx950, complex FP64).
//
// Complex numbers are interleaved double2 {re, im}; matrices are row-major with an explicit
// leading dimension (row stride, in elements). Batches address their matrices through
// BatchPtr: element b lives at base + (b / inner) * s_outer + (b % inner) * s_inner (or
// base + idx[b * idx_stride] * s_outer with a gather index), which covers contiguous batches,
// shared operands (stride 0), layer-major (point, layer) grids and one layer's column of such a
// grid, without device pointer arrays.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double2 cd;

struct BatchPtr {
  double2* base;
  long long s_outer;
  long long s_inner;
  int inner;
  const int* idx;          // optional gather: element b is base + idx[b * idx_stride] * s_outer
  int idx_stride;
  __host__ __device__ inline double2* ptr(long long b) const {
    if (idx) return base + (long long)idx[b * idx_stride] * s_outer;
    return base + (b / inner) * s_outer + (b % inner) * s_inner;
  }
};

static inline BatchPtr bp_strided(const void* base, long long stride) {
  BatchPtr p; p.base = (double2*)base; p.s_outer = stride; p.s_inner = 0; p.inner = 1; p.idx = nullptr;
  p.idx_stride = 1;
  return p;
}
static inline BatchPtr bp_grid(const void* base, long long s_outer, long long s_inner, int inner) {
  BatchPtr p; p.base = (double2*)base; p.s_outer = s_outer; p.s_inner = s_inner; p.inner = inner;
  p.idx = nullptr; p.idx_stride = 1;
  return p;
}
static inline BatchPtr bp_gather(const void* base, long long stride, const int* idx,
                                 int idx_stride = 1) {
  BatchPtr p = bp_strided(base, stride); p.idx = idx; p.idx_stride = idx_stride; return p;
}
static inline BatchPtr bp_offset(BatchPtr p, long long off) { p.base += off; return p; }

__host__ __device__ inline cd mkc(double r, double i) { cd z; z.x = r; z.y = i; return z; }
__host__ __device__ inline cd cadd(cd a, cd b) { return mkc(a.x + b.x, a.y + b.y); }
__host__ __device__ inline cd csub(cd a, cd b) { return mkc(a.x - b.x, a.y - b.y); }
__host__ __device__ inline cd cmul(cd a, cd b) { return mkc(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__host__ __device__ inline cd cconj(cd a) { return mkc(a.x, -a.y); }
__host__ __device__ inline cd cscale(cd a, double s) { return mkc(a.x * s, a.y * s); }
__host__ __device__ inline cd cneg(cd a) { return mkc(-a.x, -a.y); }
// conj(a) * b
__host__ __device__ inline cd cmulc(cd a, cd b) { return mkc(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x); }
__host__ __device__ inline double cabs1(cd a) { return fabs(a.x) + fabs(a.y); }
__host__ __device__ inline double cabs(cd a) { return hypot(a.x, a.y); }
__host__ __device__ inline double cnorm2(cd a) { return a.x * a.x + a.y * a.y; }

// Smith's algorithm for a / b.
__host__ __device__ inline cd cdiv(cd a, cd b) {
  if (fabs(b.y) <= fabs(b.x)) {
    double r = b.y / b.x, d = b.x + b.y * r;
    return mkc((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  }
  double r = b.x / b.y, d = b.y + b.x * r;
  return mkc((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}
__host__ __device__ inline cd crecip(cd b) { return cdiv(mkc(1.0, 0.0), b); }

// Principal square root with C99 csqrt semantics on finite inputs (the sign of the imaginary
// part follows Im(z), signed zero included).
__host__ __device__ inline cd csqrt_p(cd z) {
  double x = z.x, y = z.y;
  if (x == 0.0 && y == 0.0) return mkc(0.0, y);
  double t = sqrt((fabs(x) + hypot(x, y)) * 0.5);
  if (x >= 0.0) return mkc(t, y / (2.0 * t));
  return mkc(fabs(y) / (2.0 * t), copysign(t, y));
}

__host__ __device__ inline cd cexp_(cd z) {
  double e = exp(z.x);
  double s, c;
  sincos(z.y, &s, &c);
  return mkc(e * c, e * s);
}

#define WB_CHECK_LAUNCH() do { hipError_t e__ = hipGetLastError(); if (e__ != hipSuccess) return e__; } while (0)
