// Scalar FP64 math for the sweep kernels, written so every result that feeds a
// position or a decision is the one the reference computes on Linux.
//
//  * hmc_cbrt: Rust's f64::cbrt calls the C library's cbrt; this restates
//    glibc 2.35's sysdeps/ieee754/dbl-64/s_cbrt.c (IBM Accurate Mathematical
//    Library lineage; LGPL-2.1+): frexp, the rational polynomial, one
//    Newton-like step, then scaling by 2^(e/3).  It matches the host library
//    bit for bit on the inputs the 3D ball sampler produces.
//  * glibc_exp, glibc_log, glibc_sincos (glibc_math.hpp): the library's exp,
//    log and sincos, for the normal sampler's tail and wedge test, Metropolis,
//    and the rotation sampler.
//
// Everything here is compiled with -ffp-contract=off; fma() appears only where
// the restated library fuses a product.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#define HMC_HD __host__ __device__ __forceinline__

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

}  // namespace hmc

#include "glibc_math.hpp"
