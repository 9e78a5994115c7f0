/* Checks that the shared-reciprocal quotient of the device kernels,
 * q = RN(a*y), r = fma(-b, q, a), RN(q + r*y) with y = RN(1/b),
 * equals the IEEE quotient a/b bit for bit.  Build: cc -O2 -ffp-contract=off.
 * Usage: qdiv_check <pairs> <seed>; prints the mismatch count. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t st;
static uint64_t next64(void) {
  st += 0x9E3779B97F4A7C15ull;
  uint64_t z = st;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double draw(int emin, int emax) {
  uint64_t m = next64() & 0xFFFFFFFFFFFFFull;
  int e = emin + (int)(next64() % (uint64_t)(emax - emin + 1));
  uint64_t b = ((uint64_t)(e + 1023) << 52) | m;
  if (next64() & 1) b |= 1ull << 63;
  double d;
  memcpy(&d, &b, 8);
  return d;
}
static double set_low_bits(double x, uint64_t mask) {
  uint64_t b;
  memcpy(&b, &x, 8);
  b |= mask;
  memcpy(&x, &b, 8);
  return x;
}
static double qdiv(double a, double b, double y) {
  const double q = a * y;
  const double r = fma(-b, q, a);
  return fma(r, y, q);
}
int main(int argc, char **argv) {
  long n = argc > 1 ? atol(argv[1]) : 1000000;
  st = argc > 2 ? strtoull(argv[2], 0, 0) : 1;
  long bad = 0;
  for (long i = 0; i < n; i++) {
    double b = draw(-40, 40), a = draw(-60, 60);
    if (i % 4 == 1) b = set_low_bits(b, 0xFFFFFFFFFFF00ull); /* near all-ones significands */
    if (i % 4 == 2) a = set_low_bits(a, 0xFFFFFFFFFFFF0ull);
    if (i % 8 == 3) b = ldexp(1.0, (int)(next64() % 40) - 20) * (1.0 + ldexp((double)(next64() % 16), -52));
    volatile double y = 1.0 / b;
    double t = a / b, q = qdiv(a, b, y);
    if (memcmp(&t, &q, 8) != 0) bad++;
  }
  printf("%ld\n", bad);
  return 0;
}
