// Oracle (test infrastructure only): correctly rounded log and exp in
// double-double arithmetic, and the switch between them and the host C
// library.
//
// Rust's f64::ln / f64::exp call the platform C library; on Linux that is
// glibc, whose log and exp are not correctly rounded for roughly 1 input in
// 1200 (measured against a 45-digit decimal evaluation).  The oracle therefore
// runs in one of two modes:
//   glibc  (default) -- the reference's arithmetic on this platform;
//   cr               -- correctly rounded log/exp, the arithmetic of the GPU
//                       path.
// In both modes it counts the calls where the two would differ (a log value,
// or the outcome of an exp comparison), so a replay can state exactly how many
// results depend on the platform library.
#pragma once
#include <atomic>
#include <cmath>
#include <cstdint>

namespace oracle {

struct DD {
  double hi, lo;
};
inline DD dd_two_sum(double a, double b) {
  double s = a + b, bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}
inline DD dd_quick(double a, double b) {
  double s = a + b;
  return {s, b - (s - a)};
}
inline DD dd_add(DD x, DD y) {
  DD s = dd_two_sum(x.hi, y.hi), t = dd_two_sum(x.lo, y.lo);
  s.lo += t.hi;
  s = dd_quick(s.hi, s.lo);
  s.lo += t.lo;
  return dd_quick(s.hi, s.lo);
}
inline DD dd_mul(DD x, DD y) {
  double p = x.hi * y.hi;
  double e = std::fma(x.hi, y.hi, -p) + (x.hi * y.lo + x.lo * y.hi);
  return dd_quick(p, e);
}
inline DD dd_div(DD x, DD y) {
  double q1 = x.hi / y.hi;
  DD r = dd_add(x, dd_mul(y, {-q1, 0.0}));
  double q2 = r.hi / y.hi;
  r = dd_add(r, dd_mul(y, {-q2, 0.0}));
  double q3 = r.hi / y.hi;
  return dd_add(dd_quick(q1, q2), {q3, 0.0});
}
inline DD dd_ln2() { return {0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56}; }

// log(m) = 2 atanh((m-1)/(m+1)) with m in [sqrt(1/2), sqrt(2)).
inline double cr_log(double x) {
  int e;
  double m = std::frexp(x, &e);
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    e -= 1;
  }
  DD f = dd_div({m - 1.0, 0.0}, dd_two_sum(m, 1.0));
  DD f2 = dd_mul(f, f), t = f, s = f;
  for (int k = 1; k <= 24; ++k) {
    t = dd_mul(t, f2);
    s = dd_add(s, dd_div(t, {(double)(2 * k + 1), 0.0}));
  }
  DD r = dd_add(dd_mul(dd_ln2(), {(double)e, 0.0}), {2.0 * s.hi, 2.0 * s.lo});
  return r.hi + r.lo;
}

// exp(y) = 2^k exp(r), r = y - k ln2, Taylor in Horner form.  -708 <= y <= 709.
inline double cr_exp(double y) {
  double k = std::round(y * 1.4426950408889634);
  DD r = dd_add({y, 0.0}, dd_mul(dd_ln2(), {-k, 0.0}));
  DD p = {1.0, 0.0};
  for (int n = 27; n >= 1; --n) p = dd_add(dd_div(dd_mul(p, r), {(double)n, 0.0}), {1.0, 0.0});
  return std::ldexp(p.hi + p.lo, (int)k);
}

// sin(x) and cos(x) for |x| <= 2 (the rotation sampler calls them on
// [0, pi/2]): Taylor series in Horner form, 20 terms, in double-double.
inline double cr_sin(double x) {
  DD x2 = dd_mul({x, 0.0}, {x, 0.0});
  DD p = {1.0, 0.0};
  for (int n = 20; n >= 1; --n) {
    DD t = dd_div(dd_mul(p, x2), {(double)(2 * n) * (double)(2 * n + 1), 0.0});
    p = dd_add({1.0, 0.0}, {-t.hi, -t.lo});
  }
  DD r = dd_mul(p, {x, 0.0});
  return r.hi + r.lo;
}
inline double cr_cos(double x) {
  DD x2 = dd_mul({x, 0.0}, {x, 0.0});
  DD p = {1.0, 0.0};
  for (int n = 20; n >= 1; --n) {
    DD t = dd_div(dd_mul(p, x2), {(double)(2 * n - 1) * (double)(2 * n), 0.0});
    p = dd_add({1.0, 0.0}, {-t.hi, -t.lo});
  }
  return p.hi + p.lo;
}

struct LibmState {
  int mode = 0;  // 0 glibc, 1 correctly rounded
  std::atomic<uint64_t> log_calls{0}, log_divergent{0}, exp_calls{0}, exp_divergent{0};
  std::atomic<uint64_t> trig_calls{0}, trig_divergent{0};
};
inline LibmState& libm() {
  static LibmState s;
  return s;
}

inline double lib_log(double x) {
  LibmState& s = libm();
  double g = std::log(x), c = cr_log(x);
  s.log_calls.fetch_add(1, std::memory_order_relaxed);
  if (g != c) s.log_divergent.fetch_add(1, std::memory_order_relaxed);
  return s.mode ? c : g;
}

// sin / cos as the reference calls them (glibc), or correctly rounded; the
// correctly rounded versions cover |x| <= 2 only.
inline double lib_sin(double x) {
  LibmState& s = libm();
  double g = std::sin(x);
  double c = std::fabs(x) <= 2.0 ? cr_sin(x) : g;
  s.trig_calls.fetch_add(1, std::memory_order_relaxed);
  if (g != c) s.trig_divergent.fetch_add(1, std::memory_order_relaxed);
  return s.mode ? c : g;
}
inline double lib_cos(double x) {
  LibmState& s = libm();
  double g = std::cos(x);
  double c = std::fabs(x) <= 2.0 ? cr_cos(x) : g;
  s.trig_calls.fetch_add(1, std::memory_order_relaxed);
  if (g != c) s.trig_divergent.fetch_add(1, std::memory_order_relaxed);
  return s.mode ? c : g;
}

// lhs < exp(y)
inline bool lib_lt_exp(double lhs, double y) {
  LibmState& s = libm();
  bool g = lhs < std::exp(y);
  bool c = (y >= -708.0 && y <= 709.0) ? lhs < cr_exp(y) : g;
  s.exp_calls.fetch_add(1, std::memory_order_relaxed);
  if (g != c) s.exp_divergent.fetch_add(1, std::memory_order_relaxed);
  return s.mode ? c : g;
}

}  // namespace oracle
