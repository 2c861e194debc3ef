// Initial site positions and orientations, boundary maps and test hooks.
#include "sweep_device.hpp"

namespace hmc {

// Site positions (given, or wrap(body + 0)) and their ghost flags.
template <int D>
__global__ void k_init_sites(Geometry g, Bodies B, int have_sites) {
  uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.n_bodies) return;
  double4 b = B.bpos[e];
  double x[3] = {b.x, b.y, b.z};
  double s[3] = {0.0, 0.0, 0.0};
  if (have_sites) {
    double4 sp = B.spos[e];
    s[0] = sp.x;
    s[1] = sp.y;
    s[2] = sp.z;
  } else {
    for (int a = 0; a < D; ++a) s[a] = wrap1(x[a] + 0.0, g.L[a]);
  }
  B.spos[e] = pack_site(s, ghost_flags<D>(g, s));
}

// Microstate::clone_with_boundary with every body mapped (microstate.rs:1529-1565):
// Hypercuboid::map_point (cuboid.rs:553-567), p' = clamp(p (L'/L), -L'/2,
// next_down(L'/2)) per axis, then update_body_properties (microstate.rs:985-1034):
// body = wrap(p'), site = wrap(body + 0) with its ghost flags.  g is the new
// geometry; a body outside the old box (map_point's error, on which the
// reference panics) sets *outside.
template <int D>
__global__ void k_map_bodies(Geometry g, double3 lold, const double4* __restrict__ src, Bodies B,
                             uint32_t* outside) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.n_bodies) return;
  const double4 b = src[e];
  const double x[3] = {b.x, b.y, b.z};
  const double lo[3] = {lold.x, lold.y, lold.z};
  double p[3] = {0.0, 0.0, 0.0};
  bool inside = true;
  for (int a = 0; a < D; ++a) {
    inside = inside && -lo[a] / 2.0 <= x[a] && x[a] < lo[a] / 2.0;  // cuboid.rs:449-454
    const double scale = g.L[a] / lo[a];
    const double low = -g.L[a] / 2.0;
    const double high = __longlong_as_double(__double_as_longlong(g.L[a] / 2.0) - 1);  // next_down of a positive
    const double v = scale * x[a];
    p[a] = v < low ? low : (v > high ? high : v);  // f64::clamp
  }
  if (!inside) atomicOr(outside, 1u);
  double s[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < D; ++a) p[a] = wrap1(p[a], g.L[a]);
  for (int a = 0; a < D; ++a) s[a] = wrap1(p[a] + 0.0, g.L[a]);
  B.bpos[e] = make_double4(p[0], p[1], p[2], 0.0);
  B.spos[e] = pack_site(s, ghost_flags<D>(g, s));
}

// ------------------------------------------------------------ moved bodies --
// Bit i % 64 of mask[i / 64]: body i's stamp (the epoch of the apply that last
// moved it) is above `since`; *count receives the number of such bodies.
__global__ void __launch_bounds__(256) k_moved_mask(const double4* __restrict__ bpos, uint32_t n, uint32_t since,
                                                    unsigned long long* mask, unsigned long long* count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool moved = i < n && stamp_of(bpos[i].w) > since;
  const unsigned long long b = __ballot(moved);
  if (__lane_id() == 0 && i < n) {
    mask[i >> 6] = b;
    if (b) atomicAdd(count, (unsigned long long)__popcll(b));
  }
}

// Body and site positions (and orientations of oriented bodies) of the
// listed bodies: out row k = {body, site, body quat, site quat} as double4.
__global__ void k_gather_bodies(Bodies B, const uint32_t* __restrict__ idx, uint32_t n, double4* __restrict__ out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t i = idx[k];
  out[4 * (size_t)k] = B.bpos[i];
  out[4 * (size_t)k + 1] = B.spos[i];
  if (B.bquat) {
    out[4 * (size_t)k + 2] = B.bquat[i];
    out[4 * (size_t)k + 3] = B.squat[i];
  }
}

// A HOOMD GSD frame's particle arrays from the sites (tag order == body
// order): positions as FP32 triples (z = 0 in 2D; Rust's `as f32` rounds to
// nearest like the conversion here) and, for oriented bodies, orientations as
// FP32 (s, x, y, z) (hoomd-gsd/src/hoomd.rs particles_position and
// particles_orientation).
__global__ void k_pack_frame(Bodies B, uint32_t n, int dim, float* __restrict__ pos, float* __restrict__ quat) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 s = B.spos[i];
  pos[3 * (size_t)i] = (float)s.x;
  pos[3 * (size_t)i + 1] = (float)s.y;
  pos[3 * (size_t)i + 2] = dim == 3 ? (float)s.z : 0.0f;
  if (quat) {
    const double4 q = B.squat[i];
    quat[4 * (size_t)i] = (float)q.x;
    quat[4 * (size_t)i + 1] = (float)q.y;
    quat[4 * (size_t)i + 2] = (float)q.z;
    quat[4 * (size_t)i + 3] = (float)q.w;
  }
}

// ------------------------------------------------------------- debug hooks --
__global__ void k_debug_math(int kind, const double* in, double* out, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r = 0.0;
  switch (kind) {
    case 0: r = hmc_cbrt(in[i]); break;
    case 1: r = glibc_log(in[i]); break;
    case 2: r = glibc_exp(in[i]); break;
    case 3: r = sqrt(in[i]); break;
    case 5: {
      double c;
      glibc_sincos(in[i], &r, &c);
      break;
    }
    case 6: {
      double sn;
      glibc_sincos(in[i], &sn, &r);
      break;
    }
    default: r = in[2 * i] / in[2 * i + 1]; break;
  }
  out[i] = r;
}

__global__ void k_debug_normals(uint64_t step, uint32_t substep, uint32_t seed, uint64_t index, double* out,
                                uint64_t n) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  Sfc64 r = counter_rng(step, substep, seed, index);
  for (uint64_t i = 0; i < n; ++i) out[i] = standard_normal(r, kZigX, kZigF);
}

// explicit instantiations (launched from the host translation units)
template __global__ void k_init_sites<2>(Geometry g, Bodies B, int have_sites);
template __global__ void k_init_sites<3>(Geometry g, Bodies B, int have_sites);
template __global__ void k_map_bodies<2>(Geometry g, double3 lold, const double4* __restrict__ src, Bodies B, uint32_t* outside);
template __global__ void k_map_bodies<3>(Geometry g, double3 lold, const double4* __restrict__ src, Bodies B, uint32_t* outside);

}  // namespace hmc
