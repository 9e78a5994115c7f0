// Arithmetic helpers of the stage kernels that are also compiled for the
// host by the tests (tests/math_check.cpp), so that the checks run the same
// source as the device.
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define ECO_HD __host__ __device__ __forceinline__
#else
#define ECO_HD inline
#endif

namespace eco {

// a / b, correctly rounded, from y = RN(1/b): q = RN(a*y) is within one ulp
// of a/b, the residual r = a - b*q is exact in one fma, and RN(q + r*y) is
// the correctly rounded quotient (Markstein's theorem; P. Markstein, "IA-64
// and elementary functions", 2000, Thm. 8.10).  Three operations instead of
// the ten of a division when several numerators share b; bit-identical to
// a / b for finite, normal operands.
ECO_HD double qdiv(double a, double b, double y) {
  const double q = a * y;
  const double r = std::fma(-b, q, a);
  return std::fma(r, y, q);
}

// x^g for x > 0 (the isentrope's density ratio).  ln x = e ln2 + ln m with
// m in [1/sqrt2, sqrt2): ln m = 2 atanh(s), s = (m-1)/(m+1), |s| < 0.172,
// by its series to s^23; ln 2 split in a 32-bit head and a tail so that
// e*ln2_hi is exact; the product g*ln x is carried as a double-double and
// applied as exp(hi)*(1 + lo).  Error against a correctly rounded power:
// <= 2 ulp (g = 1.4), <= 5 ulp (g = 4.4) for x in [1e-3, 1e3].  Zero,
// negative, infinite and NaN bases take the library pow (same special values).
ECO_HD double pow_pos(double x, double g) {
  if (!(x > 0. && x <= 1.7976931348623157e308)) return std::pow(x, g);
  constexpr double LN2_HI = 6.93147180369123816490e-01;  // ln 2 to 32 significant bits
  constexpr double LN2_LO = 1.90821492927058770002e-10;  // ln 2 - LN2_HI
  int e;
  double m = std::frexp(x, &e);  // m in [0.5, 1)
  if (m < 0.70710678118654752440) { m *= 2.; e -= 1; }
  const double s = (m - 1.) / (m + 1.), s2 = s * s;
  double p = 1. / 23.;
  p = std::fma(p, s2, 1. / 21.);
  p = std::fma(p, s2, 1. / 19.);
  p = std::fma(p, s2, 1. / 17.);
  p = std::fma(p, s2, 1. / 15.);
  p = std::fma(p, s2, 1. / 13.);
  p = std::fma(p, s2, 1. / 11.);
  p = std::fma(p, s2, 1. / 9.);
  p = std::fma(p, s2, 1. / 7.);
  p = std::fma(p, s2, 1. / 5.);
  p = std::fma(p, s2, 1. / 3.);
  const double lm = 2. * s + 2. * s * s2 * p;       // ln m
  const double de = (double)e;
  const double A = de * LN2_HI;                     // exact (|e| < 2^11)
  const double B = lm + de * LN2_LO;
  const double p1 = g * A, e1 = std::fma(g, A, -p1);  // p1 + e1 == g*A exactly
  const double p2 = g * B;
  const double th = p1 + p2;
  const double tl = ((p1 - th) + p2) + e1;          // |p2| <= |p1| unless A == 0
  const double E = std::exp(th);
  return std::fma(E, tl, E);
}

}  // namespace eco
