// Float-float ("double-float") arithmetic for the few discrete predicates that
// the reference decides in double precision on continuous values (the
// "sea in the sky" and ceiling demotions of _markFluidCells,
// js/solver.js:964-996). A float filter decides almost every case; only
// values inside its error band are re-evaluated here, with ~48 significant
// bits, so the decisions match the reference's double arithmetic.
#pragma once
#include <hip/hip_runtime.h>

namespace ws {

struct df {
  float hi, lo;
};

__device__ __forceinline__ df df_make(float a) { return df{a, 0.f}; }

__device__ __forceinline__ df two_sum(float a, float b) {
  float s = a + b;
  float bb = s - a;
  float e = (a - (s - bb)) + (b - bb);
  return df{s, e};
}

__device__ __forceinline__ df quick_two_sum(float a, float b) {
  float s = a + b;
  return df{s, b - (s - a)};
}

__device__ __forceinline__ df two_prod(float a, float b) {
  float p = a * b;
  return df{p, __builtin_fmaf(a, b, -p)};
}

__device__ __forceinline__ df df_add(df x, df y) {
  df s = two_sum(x.hi, y.hi);
  df t = two_sum(x.lo, y.lo);
  s.lo += t.hi;
  s = quick_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return quick_two_sum(s.hi, s.lo);
}

__device__ __forceinline__ df df_neg(df x) { return df{-x.hi, -x.lo}; }
__device__ __forceinline__ df df_sub(df x, df y) { return df_add(x, df_neg(y)); }

__device__ __forceinline__ df df_mul(df x, df y) {
  df p = two_prod(x.hi, y.hi);
  p.lo += x.hi * y.lo + x.lo * y.hi;
  return quick_two_sum(p.hi, p.lo);
}

__device__ __forceinline__ df df_div(df x, df y) {
  float q1 = x.hi / y.hi;
  df r = df_sub(x, df_mul(df_make(q1), y));
  float q2 = r.hi / y.hi;
  r = df_sub(r, df_mul(df_make(q2), y));
  float q3 = r.hi / y.hi;
  df q = quick_two_sum(q1, q2);
  return df_add(q, df_make(q3));
}

__device__ __forceinline__ df df_sqrt(df a) {
  if (a.hi <= 0.f) return df{0.f, 0.f};
  float x = sqrtf(a.hi);
  df r = df_sub(a, two_prod(x, x));
  return quick_two_sum(x, r.hi / (2.f * x));
}

// floor of hi+lo
__device__ __forceinline__ int df_floor(df a) {
  float f = floorf(a.hi);
  if (f == a.hi && a.lo < 0.f) f -= 1.f;
  return (int)f;
}

__device__ __forceinline__ bool df_gt0(df a) { return a.hi > 0.f || (a.hi == 0.f && a.lo > 0.f); }

}  // namespace ws
