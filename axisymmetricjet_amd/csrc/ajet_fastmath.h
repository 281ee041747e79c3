// FP64 building blocks of the interior-tile stage kernels: the SBP interior stencils in
// factored form with fused multiply-adds, T**0.666 from a logarithm and an exponential of the
// library's own (series restated below), and a reciprocal.  Each rounds differently from the
// reference's unfused, left-to-right arithmetic by a few ulp of its inputs, far inside the
// parity bars (1e-12 normwise per RHS, 1e-9 on the state after 1000 steps); boundary tiles keep
// the reference's exact operation order (ajet_stencil.h, ajet_physics.h).
#pragma once

#include <hip/hip_runtime.h>

namespace ajet {

// D1 interior row (1/12, -2/3, 0, 2/3, -1/12) (ModOperator.f90:42) as
// (2/3)(c[1] - c[-1]) - (1/12)(c[2] - c[-2]).
__host__ __device__ __forceinline__ double d1_int(double m2, double m1, double p1, double p2) {
  return fma(2.0 / 3.0, p1 - m1, (-1.0 / 12.0) * (p2 - m2));
}

// D2 interior row (-1/12, 4/3, -5/2, 4/3, -1/12) (ModOperator.f90:78) as
// (4/3)(c[-1] + c[1]) - (1/12)(c[-2] + c[2]) - (5/2) c[0].
__host__ __device__ __forceinline__ double d2_int(double m2, double m1, double c0, double p1, double p2) {
  return fma(4.0 / 3.0, m1 + p1, fma(-1.0 / 12.0, m2 + p2, -2.5 * c0));
}

// 1/x to within an ulp: the hardware estimate and two Newton steps.
__host__ __device__ __forceinline__ double rcp(double x) {
#ifdef __HIP_DEVICE_COMPILE__
  double y = __builtin_amdgcn_rcp(x);
#else
  double y = (double)(1.0f / (float)x);
#endif
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}

// x**0.666 for 2^-8 < x < 2^8 (the library pow elsewhere; in this range |y| below stays under 4,
// so its rounding costs at most an ulp of the result), within ~3 ulp:
//   x = m 2^e with m in [1/sqrt 2, sqrt 2);  ln m = 2 atanh(s) = 2 sum_k s^(2k+1) / (2k+1),
//   s = (m - 1)/(m + 1), |s| <= 0.1716, terms to s^21 (the next is below 1e-19 relative);
//   y = 0.666 (e ln 2 + ln m);  exp(y) = 2^k exp(r), k = rint(y / ln 2), r = y - k ln 2 in two
//   parts (|r| <= ln2 / 2), exp(r) = sum_n r^n / n! to n = 14 (the next term is below 1e-18).
__host__ __device__ __forceinline__ double pow0666(double x) {
  if (!(x > 0x1p-8 && x < 0x1p8)) return pow(x, 0.666);
  int e;
  double m = frexp(x, &e);
  if (m < 0.70710678118654752) {
    m = m + m;
    e -= 1;
  }
  const double s = (m - 1.0) * rcp(m + 1.0);
  const double s2 = s * s;
  double p = 1.0 / 21.0;
  p = fma(p, s2, 1.0 / 19.0);
  p = fma(p, s2, 1.0 / 17.0);
  p = fma(p, s2, 1.0 / 15.0);
  p = fma(p, s2, 1.0 / 13.0);
  p = fma(p, s2, 1.0 / 11.0);
  p = fma(p, s2, 1.0 / 9.0);
  p = fma(p, s2, 1.0 / 7.0);
  p = fma(p, s2, 1.0 / 5.0);
  p = fma(p, s2, 1.0 / 3.0);
  const double lnm = fma(2.0 * s, s2 * p, 2.0 * s);
  const double ln2 = 0.6931471805599453, ln2_hi = 0.6931471803691238, ln2_lo = 1.9082149292705877e-10;
  const double y = 0.666 * fma((double)e, ln2, lnm);
  const double k = rint(y * 1.4426950408889634);
  double r = fma(-k, ln2_hi, y);
  r = fma(-k, ln2_lo, r);
  double q = 1.0 / 87178291200.0;  // 1/14!
  q = fma(q, r, 1.0 / 6227020800.0);
  q = fma(q, r, 1.0 / 479001600.0);
  q = fma(q, r, 1.0 / 39916800.0);
  q = fma(q, r, 1.0 / 3628800.0);
  q = fma(q, r, 1.0 / 362880.0);
  q = fma(q, r, 1.0 / 40320.0);
  q = fma(q, r, 1.0 / 5040.0);
  q = fma(q, r, 1.0 / 720.0);
  q = fma(q, r, 1.0 / 120.0);
  q = fma(q, r, 1.0 / 24.0);
  q = fma(q, r, 1.0 / 6.0);
  q = fma(q, r, 0.5);
  q = fma(q, r, 1.0);
  q = fma(q, r, 1.0);
  return ldexp(q, (int)k);
}

}  // namespace ajet
