// Hard convex polyhedra on the device: quaternion algebra, the polytope
// support function, XenoCollide (Minkowski portal refinement) and the
// Rotate<Versor> proposal, each in the evaluation order of the reference so
// that every overlap decision and every orientation is reproduced bit for
// bit.  fma() appears exactly where the reference calls mul_add (the
// Cartesian dot product, hoomd-vector/src/cartesian.rs:293-296).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dev_math.hpp"
#include "dev_rng.hpp"

namespace hmc {

struct v3 {
  double x, y, z;
};
struct quat {
  double s;
  v3 v;
};

__device__ __forceinline__ double dot3(const v3& a, const v3& b) {
  return fma(a.z, b.z, fma(a.y, b.y, fma(a.x, b.x, 0.0)));
}
// cartesian.rs:604-610 (unfused)
__device__ __forceinline__ v3 cross3(const v3& a, const v3& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ v3 sub3(const v3& a, const v3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ v3 add3(const v3& a, const v3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ v3 neg3(const v3& a) { return {-a.x, -a.y, -a.z}; }

// quaternion.rs:405-418
__device__ __forceinline__ quat qmul(const quat& a, const quat& b) {
  const v3 c = cross3(a.v, b.v);
  return {a.s * b.s - dot3(a.v, b.v),
          {(b.v.x * a.s + a.v.x * b.s) + c.x, (b.v.y * a.s + a.v.y * b.s) + c.y, (b.v.z * a.s + a.v.z * b.s) + c.z}};
}
// quaternion.rs:740-746
__device__ __forceinline__ v3 qrotate(const quat& q, const v3& x) {
  const double A = q.s * q.s - dot3(q.v, q.v);
  const v3 c = cross3(q.v, x);
  const double Bf = 2.0 * q.s;
  const double Cf = 2.0 * dot3(q.v, x);
  return {(x.x * A + c.x * Bf) + q.v.x * Cf, (x.y * A + c.y * Bf) + q.v.y * Cf, (x.z * A + c.z * Bf) + q.v.z * Cf};
}

// RotationMatrix<3>::from(Versor) (quaternion.rs:665-690)
struct mat3 {
  v3 r0, r1, r2;
};
__device__ __forceinline__ mat3 rotmat(const quat& q) {
  const double a = q.s, b = q.v.x, c = q.v.y, d = q.v.z;
  mat3 m;
  m.r0 = {a * a + b * b - c * c - d * d, 2.0 * b * c - 2.0 * a * d, 2.0 * b * d + 2.0 * a * c};
  m.r1 = {2.0 * b * c + 2.0 * a * d, a * a - b * b + c * c - d * d, 2.0 * c * d - 2.0 * a * b};
  m.r2 = {2.0 * b * d - 2.0 * a * c, 2.0 * c * d + 2.0 * a * b, a * a - b * b - c * c + d * d};
  return m;
}
__device__ __forceinline__ v3 mrot(const mat3& m, const v3& x) { return {dot3(m.r0, x), dot3(m.r1, x), dot3(m.r2, x)}; }
// the inverse (transpose) applied without forming it
__device__ __forceinline__ v3 mrot_t(const mat3& m, const v3& x) {
  return {dot3(v3{m.r0.x, m.r1.x, m.r2.x}, x), dot3(v3{m.r0.y, m.r1.y, m.r2.y}, x),
          dot3(v3{m.r0.z, m.r1.z, m.r2.z}, x)};
}

// ConvexPolytope::support_mapping (shape/convex_polytope.rs:220-244): the
// vertex maximising (x n.x + y n.y) + z n.z, first maximum wins.  vx holds
// nv vertices as x, y, z triples; NV > 0 fixes nv at compile time (the scan
// is unrolled, with the vertex reads at constant offsets).
template <int NV>
__device__ __forceinline__ v3 support(const double* vx, uint32_t nv, const v3& n) {
  const uint32_t cnt = NV > 0 ? (uint32_t)NV : nv;
  uint32_t best = 0;
  double maxv = -__builtin_inf();
#pragma unroll
  for (uint32_t i = 0; i < cnt; ++i) {
    const double x = (vx[3 * i] * n.x + vx[3 * i + 1] * n.y) + vx[3 * i + 2] * n.z;
    if (x > maxv) {
      best = i;
      maxv = x;
    }
  }
  return {vx[3 * best], vx[3 * best + 1], vx[3 * best + 2]};
}

// xenocollide.rs:320-329: S(n) = R sup(R^-1 n) + v_ij - sup(-n)
template <int NV>
__device__ __forceinline__ v3 minkowski(const double* vx, uint32_t nv, const mat3& R, const v3& vij, const v3& n) {
  const v3 sb = add3(mrot(R, support<NV>(vx, nv, mrot_t(R, n))), vij);
  return sub3(sb, support<NV>(vx, nv, neg3(n)));
}

__device__ __forceinline__ bool tiny3(const v3& a, double tol) {
  return fabs(a.x) < tol && fabs(a.y) < tol && fabs(a.z) < tol;
}

// collide3d for a polytope with itself (xenocollide.rs:157-292, 378-420);
// NV as for support().
template <int NV>
__device__ __forceinline__ bool collide3d_n(const double* vx, uint32_t nv, double vx0, double vy0, double vz0,
                                            double qs, double qx, double qy, double qz) {
  const v3 vij = {vx0, vy0, vz0};
  const quat qij = {qs, {qx, qy, qz}};
  const double TOL = 2e-12;
  const int MAX_ITER = 1024;
  const mat3 R = rotmat(qij);
  const v3 v0 = vij;
  if (tiny3(v0, TOL)) return true;
  v3 v1 = minkowski<NV>(vx, nv, R, vij, neg3(v0));
  if (dot3(v1, v0) > 0.0) return false;
  v3 n = cross3(v1, v0);
  if (tiny3(n, TOL)) return true;
  v3 v2 = minkowski<NV>(vx, nv, R, vij, n);
  if (dot3(v2, n) < 0.0) return false;
  n = cross3(sub3(v1, v0), sub3(v2, v0));
  if (dot3(n, v0) > 0.0) {
    const v3 t = v1;
    v1 = v2;
    v2 = t;
    n = neg3(n);
  }
  v3 v3p;
  int count = 0;
  for (;;) {
    count += 1;
    if (count >= MAX_ITER) return true;
    v3p = minkowski<NV>(vx, nv, R, vij, n);
    if (dot3(v3p, n) <= 0.0) return false;
    if (dot3(cross3(v1, v3p), v0) < 0.0) {
      v2 = v3p;
      n = cross3(sub3(v1, v0), sub3(v2, v0));
      continue;
    }
    if (dot3(cross3(v3p, v2), v0) < 0.0) {
      v1 = v3p;
      n = cross3(sub3(v1, v0), sub3(v2, v0));
      continue;
    }
    break;
  }
  v3 p0 = v1, p1 = v2, p2 = v3p;
  count = 0;
  for (;;) {
    count += 1;
    v3 nrm = cross3(sub3(p1, p0), sub3(p2, p0));
    if (dot3(sub3(p0, v0), nrm) < 0.0) nrm = neg3(nrm);
    if (dot3(p0, nrm) >= 0.0) return true;
    const v3 vn = minkowski<NV>(vx, nv, R, vij, nrm);
    if (dot3(vn, nrm) < 0.0) return false;
    const double tol = TOL * sqrt(dot3(nrm, nrm));
    if (fabs(dot3(sub3(vn, p0), nrm)) < tol) return false;
    if (fabs(dot3(p0, nrm)) < tol) return true;
    const v3 vp = cross3(vn, v0);
    const bool t1 = dot3(vp, p0) > 0.0, t2 = dot3(vp, p1) > 0.0, t3 = dot3(vp, p2) > 0.0;
    if (t1 && t2)
      p0 = vn;
    else if (t1)
      p2 = vn;
    else if (t3)
      p1 = vn;
    else
      p0 = vn;
    if (count >= MAX_ITER) return true;
  }
}

__device__ __noinline__ bool collide3d(const double* vx, uint32_t nv, double vx0, double vy0, double vz0, double qs,
                                       double qx, double qy, double qz) {
  return collide3d_n<0>(vx, nv, vx0, vy0, vz0, qs, qx, qy, qz);
}

// Convex::intersects_at_global (hoomd-geometry/src/convex.rs:154-170): the
// circumsphere test on the squared distance d2 (computed by the caller from
// the same image), then pair_system_to_local (hoomd-vector/src/lib.rs:878-888)
// and collide3d.
// NV > 0: collide3d inlined with that vertex count; otherwise the out-of-line
// one for any vertex count.
template <int NV = 0>
__device__ __forceinline__ bool convex_overlap(const double* vx, uint32_t nv, double rcut2, double d2, const v3& ri,
                                               const quat& qi, const v3& rj, const quat& qj) {
  if (d2 >= rcut2) return false;
  const v3 rab = sub3(rj, ri);
  const quat inv = {qi.s, neg3(qi.v)};
  const v3 v = qrotate(inv, rab);
  const quat o = qmul(inv, qj);
  if constexpr (NV > 0)
    return collide3d_n<NV>(vx, nv, v.x, v.y, v.z, o.s, o.v.x, o.v.y, o.v.z);
  else
    return collide3d(vx, nv, v.x, v.y, v.z, o.s, o.v.x, o.v.y, o.v.z);
}

// VersorDisplacement::sample (hoomd-mc/src/rotate/versor.rs:39-72).
template <class G>
__device__ __forceinline__ quat versor_displacement(G& rng, double sigma, const double* zx,
                                                    const double* zf) {
  const double PI = 3.141592653589793;
  for (;;) {
    // three draws through one call site (a shift register keeps them in registers)
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll 1
    for (int i = 0; i < 3; ++i) {
      const double z = 0.0 + sigma * standard_normal(rng, zx, zf);
      s0 = s1;
      s1 = s2;
      s2 = z;
    }
    const v3 sv = {s0, s1, s2};
    const double theta = sqrt(dot3(sv, sv));
    if (theta > PI) continue;
    const double half = 0.5 * theta;
    // half.cos() and half.sin() (versor.rs:59-61): one sincos call, as LLVM
    // lowers the pair
    double sn, w;
    glibc_sincos(half, &sn, &w);
    double f = sn / theta;
    if (!isfinite(f)) f = 0.5;
    const v3 v = {sv.x * f, sv.y * f, sv.z * f};
    const double mag = sqrt(w * w + dot3(v, v));
    return {w / mag, {v.x / mag, v.y / mag, v.z / mag}};
  }
}

}  // namespace hmc
