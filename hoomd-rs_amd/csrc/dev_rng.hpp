// Counter-based random streams for the sweep kernels.
//
// One stream per (step, substep, seed, index): the SFC64 generator of
// hoomd-rand (hoomd-rand/src/sfc.rs:49-119) seeded with the Counter layout of
// hoomd-rand/src/counter.rs:160-167 and warmed up 12 rounds.  The sampling
// adapters reproduce what the reference's call sites get from rand 0.10 and
// rand_distr 0.6 (uniform f64, Open01, u32-range index, ziggurat normal).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dev_math.hpp"

namespace hmc {

#include "zig_tables.inc"

struct Sfc64 {
  uint64_t a, b, c, ctr;

  __device__ __forceinline__ uint64_t next() {
    uint64_t out = a + b + ctr;
    a = b ^ (b >> 11);
    // c + 8c as a shift and an add: folded into a multiply it would read an
    // unused high half from whatever register holds it, and a register with a
    // load in flight makes the wave wait for that load
    uint64_t c8 = c << 3;
    asm("" : "+v"(c8));
    b = c + c8;
    c = ((c << 24) | (c >> 40)) + out;
    ctr += 1;
    return out;
  }
  __device__ __forceinline__ uint32_t next_u32() { return (uint32_t)next(); }
};

// Counter(step, substep, seed).index(index).make_rng()
__device__ __forceinline__ Sfc64 counter_rng(uint64_t step, uint32_t substep, uint32_t seed, uint64_t index) {
  Sfc64 r{step, ((uint64_t)substep << 32) | (uint64_t)seed, index, 0};
#pragma unroll 1
  for (int i = 0; i < 12; ++i) r.next();
  return r;
}

// The samplers below take any generator with next() / next_u32(): Sfc64,
// or a pre-drawn stream (the tuner reads one stream from every offset).

// StandardUniform f64: 53 high bits times 2^-53, in [0, 1).
template <class R>
__device__ __forceinline__ double uniform01(R& r) { return (double)(r.next() >> 11) * 0x1p-53; }

// Open01 f64: 52 high bits as the mantissa of [1, 2) minus (1 - 2^-53).
template <class R>
__device__ __forceinline__ double open01(R& r) {
  uint64_t bits = (r.next() >> 12) | 0x3ff0000000000000ull;
  return __longlong_as_double((long long)bits) - (1.0 - 0x1p-53);
}

// Index in [0, len), len <= 2^32 - 1: u32 sample with Canon's widening multiply
// and at most one extra draw (choose on a slice).
template <class R>
__device__ __forceinline__ uint32_t choose_index(R& r, uint32_t len) {
  uint64_t m = (uint64_t)r.next_u32() * (uint64_t)len;
  uint32_t result = (uint32_t)(m >> 32);
  uint32_t lo = (uint32_t)m;
  if (lo > (uint32_t)(0u - len)) {
    uint32_t hi2 = (uint32_t)(((uint64_t)r.next_u32() * (uint64_t)len) >> 32);
    if ((uint64_t)lo + (uint64_t)hi2 > 0xffffffffull) result += 1;
  }
  return result;
}

// StandardNormal by the symmetric 256-layer ziggurat with the Marsaglia tail
// (rand_distr 0.6).  The tail's log and the wedge test's exp are glibc's,
// as Rust's f64::ln / f64::exp call them on Linux.  zx / zf: the layer
// tables (kZigX / kZigF or a copy of them in LDS).
template <class G>
__device__ __forceinline__ double standard_normal(G& r, const double* zx, const double* zf) {
  for (;;) {
    uint64_t bits = r.next();
    int i = (int)(bits & 0xff);
    double u = __longlong_as_double((long long)((bits >> 12) | 0x4000000000000000ull)) - 3.0;
    double xi = zx[i];
    double x = u * xi;
    double xn = zx[i + 1];
    if (fabs(x) < xn) return x;
    if (i == 0) {
      const double R = zx[1];
      double tx = 1.0, ty = 0.0;
      while (-2.0 * ty < tx * tx) {
        double x_ = open01(r);
        double y_ = open01(r);
        tx = glibc_log(x_) / R;
        ty = glibc_log(y_);
      }
      return (u < 0.0) ? tx - R : R - tx;
    }
    double f0 = zf[i], f1 = zf[i + 1];
    double lhs = f1 + (f0 - f1) * uniform01(r);
    if (lhs < glibc_exp(-x * x / 2.0)) return x;
  }
}

}  // namespace hmc
