// fastmath.h -- reciprocal, square root and power for the fused kernels.
//
// The IEEE-exact forms of 1/x, sqrt and pow cost 10-12, ~12 and ~200
// instructions on gfx950.  These replacements start from the hardware
// estimates (v_rcp_f64, v_rsq_f64: about 23 correct bits) and refine them
// with Newton / Goldschmidt steps to within a few ulp; pow(x, y) for x > 0
// is exp(y ln x) with ln by the atanh series on the mantissa and exp by a
// Taylor polynomial after reduction by ln 2.  Relative error of every
// function is below 1e-15 over the argument range of the hydro step (the
// cube-root powers of g1 below 1e-14), at least two orders below the 1e-12
// parity bar.  FP32 variants use the single-precision
// hardware instructions directly.
#pragma once

#include <hip/hip_runtime.h>

namespace evh {
namespace fm {

// ---- double
__device__ __forceinline__ double rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}

// The IEEE quotient a/d, rounded to nearest, from y = rcp(d) above or from
// the correctly rounded 1/d: q = a y, the exact residual r = a - d q (one
// fma), and the correction q + r y (Markstein's final step).  It is the
// sequence the compiler emits for a double division, without the operand
// scaling (v_div_scale) and special-case fix-up (v_div_fixup) that only
// operands near the ends of the exponent range need, so the result is the
// IEEE one for normal a, d with |a| > 2^-969 and a normal quotient; a zero
// a gives +0 (IEEE: the sign of a/d).  One division by d then costs three
// instructions once y is known, and several divisions share y.
__device__ __forceinline__ double div_cr(double a, double d, double y) {
  const double q = a * y;
  const double r = __builtin_fma(-d, q, a);
  return __builtin_fma(r, y, q);
}

// the correctly rounded 1/d (the IEEE 1.0/d for normal d, |d| < 2^1022)
__device__ __forceinline__ double rcp_cr(double d) {
  const double y = rcp(d);
  return __builtin_fma(__builtin_fma(-d, y, 1.0), y, y);
}

// sqrt of a positive, normal x (Goldschmidt on the rsq estimate)
__device__ __forceinline__ double sqrt_pos(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = __builtin_fma(-g, h, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  return __builtin_fma(d, h, g);
}

// natural log of a positive, normal x
__device__ __forceinline__ double log_pos(double x) {
  int e = __builtin_amdgcn_frexp_exp(x);
  double m = __builtin_amdgcn_frexp_mant(x);  // [0.5, 1)
  if (m < 0.70710678118654752440) {
    m = m + m;
    e = e - 1;
  }
  // ln m = 2 atanh(t), t = (m-1)/(m+1), |t| <= 0.1716
  const double t = (m - 1.0) * rcp(m + 1.0);
  const double t2 = t * t;
  double p = 1.0 / 21.0;
  p = __builtin_fma(p, t2, 1.0 / 19.0);
  p = __builtin_fma(p, t2, 1.0 / 17.0);
  p = __builtin_fma(p, t2, 1.0 / 15.0);
  p = __builtin_fma(p, t2, 1.0 / 13.0);
  p = __builtin_fma(p, t2, 1.0 / 11.0);
  p = __builtin_fma(p, t2, 1.0 / 9.0);
  p = __builtin_fma(p, t2, 1.0 / 7.0);
  p = __builtin_fma(p, t2, 1.0 / 5.0);
  p = __builtin_fma(p, t2, 1.0 / 3.0);
  const double tt = t + t;
  const double lnm = __builtin_fma(tt * t2, p, tt);
  const double de = (double)e;
  // ln 2 split so that de * LN2_HI is exact for |e| < 2^11
  constexpr double LN2_HI = 6.93147180369123816490e-01;
  constexpr double LN2_LO = 1.90821492927058770002e-10;
  return __builtin_fma(de, LN2_HI, __builtin_fma(de, LN2_LO, lnm));
}

// exp(z) for |z| < 700
__device__ __forceinline__ double exp_fin(double z) {
  constexpr double INV_LN2 = 1.44269504088896338700e+00;
  constexpr double LN2_HI = 6.93147180369123816490e-01;
  constexpr double LN2_LO = 1.90821492927058770002e-10;
  const double n = __builtin_rint(z * INV_LN2);
  double r = __builtin_fma(-n, LN2_HI, z);
  r = __builtin_fma(-n, LN2_LO, r);  // |r| <= 0.3466
  // Taylor series to r^13/13!
  double p = 1.0 / 6227020800.0;
  p = __builtin_fma(p, r, 1.0 / 479001600.0);
  p = __builtin_fma(p, r, 1.0 / 39916800.0);
  p = __builtin_fma(p, r, 1.0 / 3628800.0);
  p = __builtin_fma(p, r, 1.0 / 362880.0);
  p = __builtin_fma(p, r, 1.0 / 40320.0);
  p = __builtin_fma(p, r, 1.0 / 5040.0);
  p = __builtin_fma(p, r, 1.0 / 720.0);
  p = __builtin_fma(p, r, 1.0 / 120.0);
  p = __builtin_fma(p, r, 1.0 / 24.0);
  p = __builtin_fma(p, r, 1.0 / 6.0);
  p = __builtin_fma(p, r, 0.5);
  p = __builtin_fma(p, r, 1.0);
  p = __builtin_fma(p, r, 1.0);
  return __builtin_amdgcn_ldexp(p, (int)n);
}

// x**y for x > 0
__device__ __forceinline__ double pow_pos(double x, double y) { return exp_fin(y * log_pos(x)); }

// Powers x**(-g1) and x**g1 when g1 = 2/3 - delta with |delta| <= 1e-6 (the
// single-precision 5/3 of an atomic gas gives delta = 3.97e-8), for
// positive x = m 2^e with |e| < 200:
//   y = x**(-1/3): single-precision estimate (relative error below 4e-6:
//       the single-precision sum e + log2 m carries |e| <= 200), then one
//       Halley step: with t = 1 - x y^3, x**(-1/3) = y (1-t)**(-1/3)
//       = y (1 + t/3 + 2t^2/9 + 14t^3/81 + ...), and the dropped terms are
//       below 3e-16 for |t| <= 1.2e-5 (HALLEY = false: two Newton steps
//       y += (y/3)(1 - x y^3) instead, each squaring the relative error;
//       they hold fewer values at once);
//   x**(-g1) = y^2 exp(c),  x**g1 = (x y) exp(-c),  c = delta ln x,
//   with exp(c) = 1 + c + c^2/2 + c^3/6 (|c| < 1.4e-4, the dropped term is
//   below 2e-17).  ln x enters only through c, so log2 of the mantissa in
//   single precision (absolute error ~1e-7, times delta ln 2) is enough:
//   relative error below 1e-14 (measured 8e-15 over 2^(+-200)).
// third_root returns y and c; the caller checks the exponent range.
template <bool HALLEY = true>
__device__ __forceinline__ void third_root(double x, int e, double dln2, double &y, double &c) {
  const double m = __builtin_amdgcn_frexp_mant(x);  // [0.5, 1)
  const float l2m = __builtin_amdgcn_logf(float(m));
  y = double(__builtin_amdgcn_exp2f((float(e) + l2m) * (-1.0f / 3.0f)));
  if (HALLEY) {
    const double t = __builtin_fma(-x, y * y * y, 1.0);
    y = __builtin_fma(y * t, __builtin_fma(t, 2.0 / 9.0, 1.0 / 3.0), y);
  } else {  // two Newton steps y += (y/3)(1 - x y^3): fewer live values
    y = __builtin_fma(y * (1.0 / 3.0), __builtin_fma(-x, y * y * y, 1.0), y);
    y = __builtin_fma(y * (1.0 / 3.0), __builtin_fma(-x, y * y * y, 1.0), y);
  }
  c = dln2 * (double(e) + double(l2m));
}

// exp(c) - 1 for |c| < 1.4e-4
__device__ __forceinline__ double expm1_tiny(double c) {
  return __builtin_fma(__builtin_fma(c, 1.0 / 6.0, 0.5), c * c, c);
}

// 1/sqrt(x) for positive, normal x (two Newton steps on the rsq estimate)
__device__ __forceinline__ double rsqrt_pos(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  double e = __builtin_fma(-hx * y, y, 0.5);
  y = __builtin_fma(y, e, y);
  e = __builtin_fma(-hx * y, y, 0.5);
  return __builtin_fma(y, e, y);
}

// a*b + c.  FP32: one fused instruction, placed here by hand (the FP32
// build of the march is compiled without contraction, so that every inlined
// call site rounds alike).  FP64: the expression itself, contracted by the
// fast build and left unfused by the exact build.
template <typename T>
__device__ __forceinline__ T fmad(T a, T b, T c) {
  if constexpr (sizeof(T) == 4) return __builtin_fmaf(a, b, c);
  else return a * b + c;
}

// ---- float (hardware transcendentals; ~1-2 ulp)
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float sqrt_pos(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float pow_pos(float x, float y) {
  return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}
__device__ __forceinline__ float rsqrt_pos(float x) { return __builtin_amdgcn_rsqf(x); }

}  // namespace fm
}  // namespace evh
