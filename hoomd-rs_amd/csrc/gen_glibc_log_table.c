/* Regenerates the 128-entry (invc, logc) table of glibc 2.35's log
 * (sysdeps/ieee754/dbl-64/e_log_data.c, from ARM optimized-routines; MIT /
 * LGPL-2.1+) by the search its header describes, and pins every entry
 * against the host C library.
 *
 * Entry i serves the inputs x = 2^k z with z in [0x1.6p-1, 0x1.6p0) whose
 * top seven mantissa bits (counted from z = 0x1.6p-1) are i.  Its invc is a
 * double near 1/center, c = 1/invc exactly, and logc = (double)log(c) with
 * the constraints that 0x1.8p9 + logc is exact (logc a multiple of 2^-43)
 * and |log(c) - logc| < 2^-66.  This program enumerates every such invc
 * within 2^31 ulps of rn(1/center) -- about a thousand per entry, -log(invc)
 * tracked in 2^-110 fixed point from quad-precision block starts -- and
 * keeps the candidate whose log(x) (the FMA build's evaluation order, as in
 * glibc_math.hpp) equals the host library's on 2*10^4 inputs of the entry
 * (k in [-28, 1]) and then on 10^6 more (k in {-2, -1, 1, 2}).  Exactly one
 * candidate per entry survives; every other one differs from the host on
 * about 600 of the 10^6.
 *
 * Build and run (about 10 minutes on 8 cores):
 *   gcc -O2 -ffp-contract=off gen_glibc_log_table.c -lquadmath -lm -o gen_log
 *   for p in 0 1 2 3 4 5 6 7; do ./gen_log $((16*p)) $((16*p+16)) > log_$p.txt & done; wait
 *   cat log_*.txt | python3 gen_glibc_tables.py --log-table -
 * Output lines: "i invc logc" (hex floats).
 */
#include <math.h>
#include <quadmath.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef __int128 i128;
#define OFF 0x3fe6000000000000ULL

static inline uint64_t asu(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static inline double asd(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }

/* e_log_data.c: poly (A) and ln2hi / ln2lo */
static const double A0 = -0x1.0000000000001p-1, A1 = 0x1.555555551305bp-2, A2 = -0x1.fffffffeb459p-3,
                    A3 = 0x1.999b324f10111p-3, A4 = -0x1.55575e506c89fp-3;
static const double Ln2hi = 0x1.62e42fefa3800p-1, Ln2lo = 0x1.ef35793c76730p-45;

/* the table path of log (x outside [1 - 2^-4, 1 + 0x1.09p-4)), FMA build */
static double log_table_path(double x, double invc, double logc) {
  uint64_t ix = asu(x), tmp = ix - OFF;
  int k = (int)((int64_t)tmp >> 52);
  uint64_t iz = ix - (tmp & 0xfffULL << 52);
  double z = asd(iz), r = fma(z, invc, -1.0), kd = (double)k;
  double w = fma(kd, Ln2hi, logc), hi = w + r, lo = fma(kd, Ln2lo, w - hi + r);
  double r2 = r * r;
  return fma(r * r2, fma(r2, fma(r, A4, A3), fma(r, A2, A1)), fma(r2, A0, lo)) + hi;
}

static double (*volatile host_log)(double) = log;
static uint64_t xs;
static inline uint64_t xorshift(void) { xs ^= xs << 13; xs ^= xs >> 7; xs ^= xs << 17; return xs; }
static int near_one(double x) { return asu(x) - asu(1.0 - 0x1p-4) < asu(1.0 + 0x1.09p-4) - asu(1.0 - 0x1p-4); }

int main(int argc, char** argv) {
  if (argc != 3) { fprintf(stderr, "usage: %s first_entry end_entry\n", argv[0]); return 2; }
  const int i0 = atoi(argv[1]), i1 = atoi(argv[2]);
  enum { NT = 20000, NT2 = 1000000 };
  static double tx[NT], ty[NT];
  for (int i = i0; i < i1; i++) {
    const double zlo = asd(OFF + ((uint64_t)i << 45)), zhi = asd(OFF + ((uint64_t)(i + 1) << 45));
    const double invc0 = 1.0 / ((zlo + zhi) / 2);
    const double u = invc0 >= 1.0 ? 0x1p-52 : 0x1p-53; /* no binade is crossed within 2^31 ulps */
    xs = 88172645463325252ULL ^ (uint64_t)i * 0x9E3779B97F4A7C15ULL;
    for (int t = 0; t < NT; t++) {
      const uint64_t frac = xorshift() & ((1ULL << 45) - 1);
      const int k = 1 - (int)(xorshift() % 30);
      double x = ldexp(asd(OFF + ((uint64_t)i << 45) + frac), k);
      if (near_one(x)) x = ldexp(x, -3);
      tx[t] = x;
      ty[t] = host_log(x);
    }
    int found = 0;
    const long JB = 1L << 16;
    const i128 MASK = (((i128)1) << 67) - 1, HALF = ((i128)1) << 66, TH = ((i128)1) << 44;
    for (long j0 = -(1L << 31); j0 < (1L << 31); j0 += JB) {
      /* f(j) = -log(invc0 + j u) in units of 2^-110, by second differences */
      const __float128 inv0 = (__float128)invc0 + (__float128)j0 * u, sc = 0x1p110Q;
      const __float128 d1 = -log1pq((__float128)u / inv0), d1b = -log1pq((__float128)u / (inv0 + u));
      i128 F = (i128)(-logq(inv0) * sc), D = (i128)(d1 * sc);
      const i128 D2 = (i128)((d1b - d1) * sc * 0x1p20Q); /* with 20 fraction bits */
      i128 acc2 = 0;
      for (long kk = 0; kk < JB; kk++) {
        i128 rr = F & MASK;
        if (rr >= HALF) rr -= ((i128)1) << 67;
        if (rr < TH && rr > -TH) { /* -log(invc) within 2^-66 of m 2^-43 */
          const double invc = invc0 + (double)(j0 + kk) * u;
          const double logc = ldexp((double)(int64_t)((F - rr) >> 67), -43);
          int ok = 1;
          for (int t = 0; t < NT && ok; t++) ok = asu(log_table_path(tx[t], invc, logc)) == asu(ty[t]);
          if (ok) {
            uint64_t s2 = 12345 + (uint64_t)i;
            for (long t = 0; t < NT2 && ok; t++) {
              s2 ^= s2 << 13; s2 ^= s2 >> 7; s2 ^= s2 << 17;
              const int q = (int)((s2 >> 50) % 4), kx = q < 2 ? q - 2 : q - 1;
              const double x = ldexp(asd(OFF + ((uint64_t)i << 45) + (s2 & ((1ULL << 45) - 1))), kx);
              ok = asu(log_table_path(x, invc, logc)) == asu(host_log(x));
            }
            if (ok) {
              found++;
              printf("%d %a %a\n", i, invc, logc);
              fflush(stdout);
            }
          }
        }
        F += D;
        acc2 += D2;
        D += acc2 >> 20;
        acc2 &= (((i128)1) << 20) - 1;
      }
    }
    if (found != 1) fprintf(stderr, "entry %d: %d candidates reproduce the host log\n", i, found);
  }
  return 0;
}
