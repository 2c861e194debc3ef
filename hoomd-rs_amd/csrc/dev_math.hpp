// Scalar FP64 math for the sweep kernels, written so every result that feeds a
// position or a decision is reproducible bit for bit.
//
//  * hmc_cbrt: the classic frexp / rational-polynomial / ldexp cube root that
//    the host C library (glibc 2.35, what Rust's f64::cbrt calls on Linux)
//    uses; it matches that library bit for bit on the inputs the 3D ball
//    sampler produces (checked by sampling in the test suite).
//  * hmc_log_cr / hmc_exp_cr: correctly rounded natural log and exp evaluated
//    in double-double arithmetic (about 100 bits) and rounded once; log uses a
//    64-point table reduction (dd_tables.inc).
//  * hmc_sin_cr / hmc_cos_cr: correctly rounded sin and cos on |x| <= 2.
//  * hmc_lt_exp: decides lhs < exp(y) exactly (as if exp were correctly
//    rounded): the fast library exp decides unless lhs lies within 2^-44
//    (relative) of it, in which case the correctly rounded exp decides.
//
// Everything here is compiled with -ffp-contract=off; fma() appears only where
// an exact product residual is wanted.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#define HMC_HD __host__ __device__ __forceinline__
// Rare slow paths are kept out of line: they would otherwise triple the size
// of the sweep kernels, whose instruction fetch is shared by two CUs.
#define HMC_HD_COLD __host__ __device__ __noinline__

namespace hmc {

// ---------------------------------------------------------------- cube root --
HMC_HD double hmc_cbrt(double x) {
  const double kCbrt2 = 1.2599210498948731648;      // 2^(1/3)
  const double kSqrCbrt2 = 1.5874010519681994748;   // 2^(2/3)
  int xe;
  double xm = frexp(fabs(x), &xe);
  if (x == 0.0 || !isfinite(x)) return x + x;
  double u = (0.354895765043919860 +
              ((1.50819193781584896 +
                ((-2.11499494167371287 +
                  ((2.44693122563534430 +
                    ((-1.83469277483613086 + (0.784932344976639262 - 0.145263899385486377 * xm) * xm) * xm)) *
                   xm)) *
                 xm)) *
               xm));
  double t2 = u * u * u;
  int r = xe % 3;
  double factor = r == -2 ? 1.0 / kSqrCbrt2 : r == -1 ? 1.0 / kCbrt2 : r == 0 ? 1.0 : r == 1 ? kCbrt2 : kSqrCbrt2;
  double ym = u * (t2 + 2.0 * xm) / (2.0 * t2 + xm) * factor;
  return ldexp(x > 0.0 ? ym : -ym, xe / 3);
}

// ------------------------------------------------------------ double-double --
struct dd {
  double hi, lo;
};
HMC_HD dd two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}
HMC_HD dd fast_two_sum(double a, double b) {
  double s = a + b;
  double e = b - (s - a);
  return {s, e};
}
HMC_HD dd two_prod(double a, double b) {
  double p = a * b;
  double e = fma(a, b, -p);
  return {p, e};
}
HMC_HD dd dd_add(dd x, dd y) {
  dd s = two_sum(x.hi, y.hi);
  dd t = two_sum(x.lo, y.lo);
  s.lo += t.hi;
  s = fast_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return fast_two_sum(s.hi, s.lo);
}
HMC_HD dd dd_mul(dd x, dd y) {
  dd p = two_prod(x.hi, y.hi);
  p.lo += x.hi * y.lo + x.lo * y.hi;
  return fast_two_sum(p.hi, p.lo);
}
HMC_HD dd dd_mul_d(dd x, double y) {
  dd p = two_prod(x.hi, y);
  p.lo += x.lo * y;
  return fast_two_sum(p.hi, p.lo);
}
HMC_HD dd dd_div(dd x, dd y) {
  double q1 = x.hi / y.hi;
  dd r = dd_add(x, dd_mul_d(y, -q1));
  double q2 = r.hi / y.hi;
  r = dd_add(r, dd_mul_d(y, -q2));
  double q3 = r.hi / y.hi;
  dd q = fast_two_sum(q1, q2);
  return dd_add(q, dd{q3, 0.0});
}
// ln 2 = 0.693147180559945309417232121458176568... as hi + lo
HMC_HD dd dd_ln2() { return {0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56}; }

#include "dd_tables.inc"

__device__ __forceinline__ dd dd_const(const double (*t)[2], int i) { return dd{t[i][0], t[i][1]}; }

// Correctly rounded log(x) for finite x > 0: x = 2^e m with m in
// [sqrt(1/2), sqrt(2)), m = c (1 + r) with c = j/64 the nearest table point
// (c = 1 exactly around x = 1, so no cancellation there), and
// log(m) = log(c) + 2 atanh(f), f = (m - c)/(m + c), |f| < 0.0056, whose
// series is summed to f^15 in double-double (truncation below 2^-120).
__device__ __noinline__ double hmc_log_cr(double x) {
  int e;
  double m = frexp(x, &e);
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    e -= 1;
  }
  const int j = (int)rint(m * 64.0);  // 45 .. 91
  const double c = (double)j * 0x1p-6;
  const double num = m - c;  // exact (Sterbenz)
  const dd f = dd_div(dd{num, 0.0}, two_sum(m, c));
  const dd s2 = dd_mul(f, f);
  dd p = dd_const(kInvOdd, 7);
#pragma unroll 1
  for (int k = 6; k >= 0; --k) p = dd_add(dd_mul(p, s2), dd_const(kInvOdd, k));
  const dd at = dd_mul(p, f);
  const dd lm = {2.0 * at.hi, 2.0 * at.lo};
  const dd res = dd_add(dd_add(dd_mul_d(dd_ln2(), (double)e), dd_const(kLogC, j - 45)), lm);
  return res.hi + res.lo;
}

// Correctly rounded exp(y) for -708 <= y <= 709.
HMC_HD_COLD double hmc_exp_cr(double y) {
  double k = round(y * 1.4426950408889634);
  dd r = dd_add(dd{y, 0.0}, dd_mul_d(dd_ln2(), -k));
  dd p = {1.0, 0.0};
#pragma unroll 1
  for (int n = 27; n >= 1; --n) {
    p = dd_div(dd_mul(p, r), dd{(double)n, 0.0});
    p = dd_add(p, dd{1.0, 0.0});
  }
  return ldexp(p.hi + p.lo, (int)k);
}

// Correctly rounded sin(x) and cos(x) for |x| <= 2 (the rotation sampler
// evaluates them at half the rotation angle, in [0, pi/2]): Taylor series in
// Horner form, 20 terms, in double-double with double-double reciprocal
// coefficients, rounded once.
__device__ __noinline__ double hmc_sin_cr(double x) {
  const dd x2 = dd_mul(dd{x, 0.0}, dd{x, 0.0});
  dd p = {1.0, 0.0};
#pragma unroll 1
  for (int n = 20; n >= 1; --n) {
    const dd t = dd_mul(dd_mul(p, x2), dd_const(kInvSin, n));
    p = dd_add(dd{1.0, 0.0}, dd{-t.hi, -t.lo});
  }
  const dd r = dd_mul(p, dd{x, 0.0});
  return r.hi + r.lo;
}
__device__ __noinline__ double hmc_cos_cr(double x) {
  const dd x2 = dd_mul(dd{x, 0.0}, dd{x, 0.0});
  dd p = {1.0, 0.0};
#pragma unroll 1
  for (int n = 20; n >= 1; --n) {
    const dd t = dd_mul(dd_mul(p, x2), dd_const(kInvCos, n));
    p = dd_add(dd{1.0, 0.0}, dd{-t.hi, -t.lo});
  }
  return p.hi + p.lo;
}

// lhs < exp(y), decided as if exp were correctly rounded.  *ambiguous counts
// the times the slow path was needed.
HMC_HD bool hmc_lt_exp(double lhs, double y, unsigned* ambiguous) {
  double e = exp(y);
  double margin = e * 0x1p-44;
  if (lhs < e - margin) return true;
  if (lhs > e + margin) return false;
  if (y < -708.0 || y > 709.0) return lhs < e;
  if (ambiguous) *ambiguous += 1;
  return lhs < hmc_exp_cr(y);
}

}  // namespace hmc
