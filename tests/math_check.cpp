// Host build of the kernels' arithmetic helpers (ecogen_amd/csrc/eco_math.h)
// checked against the C library: qdiv bit for bit against IEEE division,
// pow_pos bit for bit against pow.
// Build: g++ -O2 -ffp-contract=off -std=c++17.  Usage:
//   math_check qdiv <pairs> <seed>          prints the mismatch count
//   math_check pow <samples> <gamma> <seed> prints the mismatch count
//   math_check exp <samples> <seed>         mismatches of exp_lib against exp
//   math_check log <samples> <seed>         mismatches of log_lib against log
//   math_check plain <pairs> <seed>         mismatches of a*RN(1/b) alone
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../ecogen_amd/csrc/eco_math.h"

static uint64_t st;
static uint64_t next64() {
  st += 0x9E3779B97F4A7C15ull;
  uint64_t z = st;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double u01() { return (double)(next64() >> 11) * (1.0 / 9007199254740992.0); }
static double draw(int emin, int emax) {
  uint64_t m = next64() & 0xFFFFFFFFFFFFFull;
  int e = emin + (int)(next64() % (uint64_t)(emax - emin + 1));
  uint64_t b = ((uint64_t)(e + 1023) << 52) | m;
  if (next64() & 1) b |= 1ull << 63;
  double d;
  std::memcpy(&d, &b, 8);
  return d;
}
static double set_low_bits(double x, uint64_t mask) {
  uint64_t b;
  std::memcpy(&b, &x, 8);
  b |= mask;
  std::memcpy(&x, &b, 8);
  return x;
}

static long check_div(long n, bool plain) {
  long bad = 0;
  for (long i = 0; i < n; i++) {
    double b = draw(-40, 40), a = draw(-60, 60);
    if (i % 4 == 1) b = set_low_bits(b, 0xFFFFFFFFFFF00ull);  // near all-ones significands
    if (i % 4 == 2) a = set_low_bits(a, 0xFFFFFFFFFFFF0ull);
    if (i % 8 == 3) b = std::ldexp(1.0, (int)(next64() % 40) - 20) * (1.0 + std::ldexp((double)(next64() % 16), -52));
    volatile double y = 1.0 / b;
    const double t = a / b;
    const double q = plain ? a * y : eco::qdiv(a, b, y);
    if (std::memcmp(&t, &q, 8) != 0) bad++;
  }
  return bad;
}

// Bit mismatches of pow_pos against the C library's pow over x log-uniform
// in [1e-3, 1e3], a quarter of the draws in the weak-wave band 1 +- 0.3 and
// an eighth within 1e-6 of 1 (tiny logarithms).
static long check_pow(long n, double g) {
  long bad = 0;
  for (long i = 0; i < n; i++) {
    double x = std::exp((u01() * 2. - 1.) * std::log(1e3));
    if (i % 4 == 0) x = 1. + (u01() * 2. - 1.) * 0.3;
    if (i % 8 == 1) x = 1. + (u01() * 2. - 1.) * 1e-6;
    const double a = eco::pow_pos(x, g), r = std::pow(x, g);
    if (std::memcmp(&a, &r, 8) != 0) {
      if (bad < 3) std::fprintf(stderr, "x=%a g=%a lib=%a pow_pos=%a\n", x, g, r, a);
      bad++;
    }
  }
  return bad;
}

// Bit mismatches of exp_lib against the C library's exp: THINC's argument
// range [-1.6, 1.6] and [-700, 700].
static long check_exp(long n) {
  long bad = 0;
  for (long i = 0; i < n; i++) {
    const double x = (i % 2) ? (u01() * 2. - 1.) * 1.6 : (u01() * 2. - 1.) * 700.;
    const double a = eco::exp_lib(x), r = std::exp(x);
    if (std::memcmp(&a, &r, 8) != 0) {
      if (bad < 3) std::fprintf(stderr, "x=%a lib=%a exp_lib=%a\n", x, r, a);
      bad++;
    }
  }
  const double sp[3] = {0., 1e-300, -1e-20};
  for (double x : sp) if (eco::exp_lib(x) != std::exp(x)) bad++;
  return bad;
}

// Disagreements of log_lib with the C library's log over the arguments of
// the saturation laws, drawn uniformly (not as exp of a uniform, whose logs
// sit next to a double, far from any rounding boundary): temperatures in
// [200, 2000] K and pressures plus the stiffened-gas offset in [1e3, 1e10]
// Pa.  Prints the count, then up to 20 disagreeing arguments with both
// results (hexadecimal) for the caller to check against the exact log.
static long check_log(long n) {
  long bad = 0;
  for (long i = 0; i < n; i++) {
    const double x = (i % 2) ? 200. + 1800. * u01() : 1e3 * std::pow(1e7, u01());
    const double a = eco::log_lib(x), r = std::log(x);
    if (std::memcmp(&a, &r, 8) != 0) {
      if (bad < 20) std::fprintf(stderr, "%a %a %a\n", x, r, a);
      bad++;
    }
  }
  const double sp[3] = {0., -1., INFINITY};
  for (double x : sp) {
    const double a = eco::log_lib(x), r = std::log(x);
    if (!((a == r) || (std::isnan(a) && std::isnan(r)))) bad += 1000000;
  }
  return bad;
}

int main(int argc, char **argv) {
  if (argc < 2) return 1;
  if (!std::strcmp(argv[1], "qdiv") || !std::strcmp(argv[1], "plain")) {
    const long n = argc > 2 ? std::atol(argv[2]) : 1000000;
    st = argc > 3 ? std::strtoull(argv[3], nullptr, 0) : 1;
    std::printf("%ld\n", check_div(n, argv[1][0] == 'p'));
    return 0;
  }
  if (!std::strcmp(argv[1], "log")) {
    const long n = argc > 2 ? std::atol(argv[2]) : 1000000;
    st = argc > 3 ? std::strtoull(argv[3], nullptr, 0) : 1;
    std::printf("%ld\n", check_log(n));
    return 0;
  }
  if (!std::strcmp(argv[1], "exp")) {
    const long n = argc > 2 ? std::atol(argv[2]) : 1000000;
    st = argc > 3 ? std::strtoull(argv[3], nullptr, 0) : 1;
    std::printf("%ld\n", check_exp(n));
    return 0;
  }
  if (!std::strcmp(argv[1], "pow")) {
    const long n = argc > 2 ? std::atol(argv[2]) : 1000000;
    const double g = argc > 3 ? std::atof(argv[3]) : 1.4;
    st = argc > 4 ? std::strtoull(argv[4], nullptr, 0) : 1;
    std::printf("%ld\n", check_pow(n, g));
    // special values follow the library pow
    const double sp[4] = {0., -1., INFINITY, NAN};
    for (double x : sp) {
      const double a = eco::pow_pos(x, g), r = std::pow(x, g);
      if (!((a == r) || (std::isnan(a) && std::isnan(r)))) { std::printf("special %g\n", x); return 2; }
    }
    return 0;
  }
  return 1;
}
