// Oracle (test infrastructure only): the C library calls of the reference's
// arithmetic.
//
// Rust's f64::ln, f64::exp and f64::cbrt call the platform C library (glibc
// on Linux), and LLVM lowers f64::sin and f64::cos of one value in one basic
// block -- the rotation sampler's half.cos() and half.sin()
// (hoomd-mc/src/rotate/versor.rs:58-61) -- to one call of its sincos.  The
// oracle makes the same calls to the host library, so it computes what the
// reference computes on the machine it runs on.  (glibc's x86-64 exp and log
// are selected at run time: the FMA builds on CPUs with FMA and AVX2; its
// sincos has one build.)
#pragma once
#include <cmath>

namespace oracle {

inline double lib_log(double x) { return std::log(x); }

inline double lib_exp(double x) { return std::exp(x); }

// lhs < exp(y)
inline bool lib_lt_exp(double lhs, double y) { return lhs < std::exp(y); }

inline void lib_sincos(double x, double* s, double* c) { ::sincos(x, s, c); }

}  // namespace oracle
