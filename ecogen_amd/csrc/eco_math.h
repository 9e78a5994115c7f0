// Arithmetic helpers of the stage kernels that are also compiled for the
// host by the tests (tests/math_check.cpp), so that the checks run the same
// source as the device.
#pragma once
#include <cmath>
#include <cstring>

#include "eco_pow_tables.h"

#if defined(__HIPCC__)
#define ECO_HD __host__ __device__ __forceinline__
#define ECO_RARE __host__ __device__ inline __attribute__((noinline))
#else
#define ECO_HD inline
#define ECO_RARE inline __attribute__((noinline))
#endif

namespace eco {

// The build target's pow, exp and log for the arguments the table-driven
// routines below leave to them (zero, subnormal, negative, infinite or NaN
// bases, results beyond the exp tables' range), which the isentrope, THINC
// and the saturation laws do not produce.  Kept out of line: inlined, the
// target's pow alone is several hundred instructions at every call site,
// which more than doubled the code of a stage kernel for a path that is
// never taken.
ECO_RARE double pow_rare(double x, double g) { return std::pow(x, g); }
ECO_RARE double exp_rare(double x) { return std::exp(x); }
ECO_RARE double log_rare(double x) { return std::log(x); }

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

// x^g for x > 0 (the isentrope's density ratio), computed as the C library
// computes it, so that the device's power is the reference's bit for bit.
// The reference calls pow from the GNU C library (>= 2.28; 2.35 here), whose
// pow is the table-driven algorithm Szabolcs Nagy published in ARM's
// optimized-routines (math/pow.c), built on x86-64 machines with FMA as the
// variant whose expressions are FMA-contracted.  Restated here from that
// published method:
//   x = 2^k z, z in [0x1.69555p-1, 0x1.69555p0), c from a 128-entry table,
//   log x = k ln2 + log c + log1p(z/c - 1) as a double-double (hi, lo), with
//   r = z/c - 1 exact in one fma and log1p(r) - r + r^2/2 by a degree-8
//   polynomial;
//   y log x = ehi + elo (exact product in one fma);
//   exp(ehi + elo) = 2^(j/128) exp(r'), j = round(ehi 128/ln2),
//   r' = ehi - j ln2/128 + elo, exp(r') - 1 by a degree-5 polynomial,
//   2^(j/128) = H (1 + T) from a 128-entry table.
// The tables are generated from their formulas by gen_pow_tables.py.  Every
// fma below is one of the contractions of the library's build; the kernels
// are compiled without contraction so nothing else fuses.  Agreement with
// the library, bit for bit, is tested on the host (tests/math_check.cpp).
// Outside the isentrope's range (x zero, subnormal, negative, infinite or
// NaN; |g log x| >= 512) the library pow of the build target is called.
namespace powtab {
#if defined(__HIP_DEVICE_COMPILE__)
#define ECO_POW_TABLE __constant__ static const
#else
#define ECO_POW_TABLE static const
#endif
ECO_POW_TABLE double LOG[128][3] = {ECO_POW_LOG_ROWS};
ECO_POW_TABLE unsigned long long EXP[256] = {ECO_POW_EXP_WORDS};
#undef ECO_POW_TABLE
}  // namespace powtab

// The tables as one block of doubles (a copy in fast memory, e.g. LDS, that
// the routines below read in place of the constant tables when given):
// LOG row-major at 0, the EXP words' bit patterns at POWTAB_EXP.
constexpr int POWTAB_EXP = 128 * 3, POWTAB_WORDS = 128 * 3 + 256;

ECO_HD double as_double(unsigned long long u) {
  double d;
  std::memcpy(&d, &u, 8);
  return d;
}
ECO_HD unsigned long long as_bits(double d) {
  unsigned long long u;
  std::memcpy(&u, &d, 8);
  return u;
}

// exp(hi + lo) for |lo| << |hi|: the C library's exp (glibc >= 2.28, the
// same optimized-routines algorithm and FMA build as its pow) evaluated at
// hi with the tail lo added to the reduced argument.  exp(x) = exp_dd(x, 0).
// Outside [2^-54, 512) in magnitude, tiny arguments give 1 + hi as the
// library does and large ones fall back to the build target's pow(px, pg)
// (or exp(hi) when px is 0), which the isentrope and THINC never reach.
ECO_HD double exp_dd(double hi, double lo, double px = 0., double pg = 0., const double *tab = nullptr) {
  const unsigned top = (unsigned)(as_bits(hi) >> 52) & 0x7ffu;
  if (top - 0x3c9u >= 0x408u - 0x3c9u) {  // |hi| < 2^-54 or >= 512
    if (top < 0x3c9u) return 1.0 + hi;
    return px != 0. ? pow_rare(px, pg) : exp_rare(hi);
  }
  constexpr double INVLN2N = 0x1.71547652b82fep0 * 128, SHIFT = 0x1.8p52;
  constexpr double NEGLN2HIN = -0x1.62e42fefa0000p-8, NEGLN2LON = -0x1.cf79abc9e3b3ap-47;
  constexpr double C2 = 0x1.ffffffffffdbdp-2, C3 = 0x1.555555555543cp-3, C4 = 0x1.55555cf172b91p-5,
                   C5 = 0x1.1111167a4d017p-7;
  double jd = std::fma(INVLN2N, hi, SHIFT);
  const unsigned long long ki = as_bits(jd);
  jd -= SHIFT;
  double rr = std::fma(jd, NEGLN2LON, std::fma(jd, NEGLN2HIN, hi));
  rr += lo;
  const unsigned idx = 2u * (unsigned)(ki & 127u);
  const double tail = tab ? tab[POWTAB_EXP + idx] : as_double(powtab::EXP[idx]);
  const double scale = as_double((tab ? as_bits(tab[POWTAB_EXP + idx + 1]) : powtab::EXP[idx + 1]) + (ki << 45));
  const double r2 = rr * rr;
  const double tp = std::fma(r2 * r2, std::fma(rr, C5, C4), std::fma(r2, std::fma(rr, C3, C2), tail + rr));
  return std::fma(scale, tp, scale);
}

ECO_HD double exp_lib(double x) { return exp_dd(x, 0.); }

// log x as lhi + llo (lhi = RN(log x) but for rare hard cases, |llo| within
// half an ulp of it) for a positive normal finite x with bits ix: the log of
// pow_pos below.
ECO_HD void log_hl(unsigned long long ix, double &lhi, double &llo, const double *tab = nullptr) {
  constexpr unsigned long long OFF = 0x3fe6955500000000ull;
  const unsigned long long tmp = ix - OFF;
  const int i = (int)((tmp >> 45) & 127u);
  const int k = (int)((long long)tmp >> 52);
  const double z = as_double(ix - (tmp & (0xfffull << 52)));
  const double kd = (double)k;
  const double invc = tab ? tab[3 * i] : powtab::LOG[i][0];
  const double logc = tab ? tab[3 * i + 1] : powtab::LOG[i][1];
  const double logctail = tab ? tab[3 * i + 2] : powtab::LOG[i][2];
  constexpr double LN2HI = 0x1.62e42fefa3800p-1, LN2LO = 0x1.ef35793c76730p-45;
  constexpr double A0 = -0x1p-1, A1 = 0x1.555555555556p-2 * -2, A2 = -0x1.0000000000006p-2 * -2,
                   A3 = 0x1.999999959554ep-3 * 4, A4 = -0x1.555555529a47ap-3 * 4,
                   A5 = 0x1.2495b9b4845e9p-3 * -8, A6 = -0x1.0002b8b263fc3p-3 * -8;
  const double r = std::fma(z, invc, -1.0);
  const double t1 = std::fma(kd, LN2HI, logc);
  const double t2 = t1 + r;
  const double lo1 = std::fma(kd, LN2LO, logctail);
  const double lo2 = t1 - t2 + r;
  const double ar = A0 * r, ar2 = r * ar, ar3 = r * ar2;
  const double hi = t2 + ar2;
  const double lo3 = std::fma(ar, r, -ar2);
  const double lo4 = t2 - hi + ar2;
  const double q = std::fma(ar2, std::fma(ar2, std::fma(r, A6, A5), std::fma(r, A4, A3)), std::fma(r, A2, A1));
  const double lo = std::fma(ar3, q, lo1 + lo2 + lo3 + lo4);
  lhi = hi + lo;
  llo = hi - lhi + lo;
}

// The natural logarithm of the saturation laws (Mixture::computeTsat) where
// the C library's log is called: the double-double logarithm above rounded
// once.  Its error before rounding is about 2^-68 relative, so it is the
// correctly rounded log but where log x lies within about 2^-15 ulp of a
// rounding boundary; the library's log (within 0.52 ulp) agrees with it there
// in all but rare cases (tests/math_check.cpp counts the mismatches).  Zero,
// subnormal, negative, infinite and NaN arguments take the build target's log.
ECO_HD double log_lib(double x) {
  const unsigned long long ix = as_bits(x);
  if (!((ix >> 52) - 1u < 0x7feu)) return log_rare(x);
  double hi, lo;
  log_hl(ix, hi, lo);
  return hi;
}

ECO_HD double pow_pos(double x, double g, const double *tab = nullptr) {
  const unsigned long long ix = as_bits(x);
  // positive normal finite x only (top 12 bits 0x001..0x7fe)
  if (!((ix >> 52) - 1u < 0x7feu)) return pow_rare(x, g);
  double lhi, llo;
  log_hl(ix, lhi, llo, tab);
  // y log x = ehi + elo
  const double ehi = g * lhi;
  const double elo = std::fma(g, llo, std::fma(g, lhi, -ehi));
  return exp_dd(ehi, elo, x, g, tab);
}

}  // namespace eco
