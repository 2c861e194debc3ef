// Oracle (test infrastructure only): restatement of hoomd-rs's hard convex
// polyhedra in 3D -- OrientedPoint<Cartesian<3>, Versor> bodies with one site
// at the body origin, PairwiseCutoff<HardShape<Convex<ConvexPolyhedron>>>,
// XenoCollide (Minkowski portal refinement) and the Translate and
// Rotate<Versor> local moves under ParallelSweep.
//
// Every expression keeps the reference's evaluation order; std::fma appears
// exactly where the Rust source calls mul_add (the Cartesian dot product).
//
// Parity status:
//   quaternion algebra, rotation matrix, support mapping, collide3d,
//   intersects_at_global -- pinned by the reference's boolean test tables
//       (xenocollide.rs, convex_polytope.rs tests transcribed in the tests)
//       and by an independent separating-axis test away from contact;
//   VersorDisplacement -- follows rotate/versor.rs:39-72; its Normal comes
//       from rand_distr 0.6 (PARITY UNPINNED, as the other rand adapters) and
//       its sin/cos from the host library's sincos (libm.hpp).
#pragma once
#include <cmath>
#include <cstdint>
#include <vector>

#include "mc.hpp"

namespace oracle {

using V3 = Vec<3>;

// hoomd-vector/src/cartesian.rs:604-610 (unfused)
inline V3 cross(const V3& a, const V3& b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
inline V3 vsub(const V3& a, const V3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline V3 vadd(const V3& a, const V3& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline V3 vneg(const V3& a) { return {-a[0], -a[1], -a[2]}; }
inline V3 vscale(const V3& a, double s) { return {a[0] * s, a[1] * s, a[2] * s}; }
// lib.rs:574-606: norm = sqrt(dot(v, v))
inline double vnorm(const V3& a) { return std::sqrt(dot<3>(a, a)); }

struct Quat {
  double s = 1.0;
  V3 v{0.0, 0.0, 0.0};
};

// quaternion.rs:405-418
inline Quat qmul(const Quat& a, const Quat& b) {
  Quat r;
  r.s = a.s * b.s - dot<3>(a.v, b.v);
  V3 c = cross(a.v, b.v);
  for (int i = 0; i < 3; ++i) r.v[i] = (b.v[i] * a.s + a.v[i] * b.s) + c[i];
  return r;
}
// quaternion.rs:245-250, 810-814
inline Quat qconj(const Quat& q) { return {q.s, vneg(q.v)}; }
// quaternion.rs:740-746
inline V3 qrotate(const Quat& q, const V3& x) {
  const double A = q.s * q.s - dot<3>(q.v, q.v);
  const V3 c = cross(q.v, x);
  const double B = 2.0 * q.s;
  const double C = 2.0 * dot<3>(q.v, x);
  V3 r;
  for (int i = 0; i < 3; ++i) r[i] = (x[i] * A + c[i] * B) + q.v[i] * C;
  return r;
}

// RotationMatrix<3>::from(Versor) (quaternion.rs:665-690), rotate
// (cartesian.rs:981-989: each row dotted with the vector), inverted
// (cartesian.rs:897-901: transpose).
struct Mat3 {
  V3 r[3];
};
inline Mat3 rotmat(const Quat& q) {
  const double a = q.s, b = q.v[0], c = q.v[1], d = q.v[2];
  Mat3 m;
  m.r[0] = {a * a + b * b - c * c - d * d, 2.0 * b * c - 2.0 * a * d, 2.0 * b * d + 2.0 * a * c};
  m.r[1] = {2.0 * b * c + 2.0 * a * d, a * a - b * b + c * c - d * d, 2.0 * c * d - 2.0 * a * b};
  m.r[2] = {2.0 * b * d - 2.0 * a * c, 2.0 * c * d + 2.0 * a * b, a * a - b * b - c * c + d * d};
  return m;
}
inline Mat3 transpose(const Mat3& m) {
  Mat3 t;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t.r[i][j] = m.r[j][i];
  return t;
}
inline V3 mrotate(const Mat3& m, const V3& x) { return {dot<3>(m.r[0], x), dot<3>(m.r[1], x), dot<3>(m.r[2], x)}; }

// ConvexPolytope<3, 32> (shape/convex_polytope.rs:200-250)
struct Polytope {
  std::vector<V3> v;
  double radius = 0.0;

  void set(const std::vector<V3>& verts) {
    v = verts;
    double m = 0.0;  // fold(0.0, f64::max) over squared norms
    for (const V3& x : v) m = std::fmax(m, dot<3>(x, x));
    radius = std::sqrt(m);
  }
  // vertex maximising (x0 n0 + x1 n1) + x2 n2; the first maximum wins
  V3 support(const V3& n) const {
    size_t best = 0;
    double maxv = -INFINITY;
    for (size_t i = 0; i < v.size(); ++i) {
      double x = (v[i][0] * n[0] + v[i][1] * n[1]) + v[i][2] * n[2];
      if (x > maxv) {
        best = i;
        maxv = x;
      }
    }
    return v[best];
  }
};

// xenocollide.rs:300-355: support function of B (rotated by q_ij, offset by
// v_ij) minus A.
struct Minkowski {
  const Polytope& a;
  const Polytope& b;
  V3 vij;
  Mat3 q, qinv;
  V3 operator()(const V3& n) const {
    V3 sb = vadd(mrotate(q, b.support(mrotate(qinv, n))), vij);
    return vsub(sb, a.support(vneg(n)));
  }
};

// collide3d (xenocollide.rs:157-292 discovery/refinement, 378-420 driver)
inline bool collide3d(const Polytope& a, const Polytope& b, const V3& vij, const Quat& qij) {
  const double TOL = 2e-12;
  const int MAX_ITER = 1024;
  Minkowski S{a, b, vij, rotmat(qij), {}};
  S.qinv = transpose(S.q);
  const V3 v0 = vij;
  // discovery
  if (std::fabs(v0[0]) < TOL && std::fabs(v0[1]) < TOL && std::fabs(v0[2]) < TOL) return true;
  V3 v1 = S(vneg(v0));
  if (dot<3>(v1, v0) > 0.0) return false;
  V3 n = cross(v1, v0);
  if (std::fabs(n[0]) < TOL && std::fabs(n[1]) < TOL && std::fabs(n[2]) < TOL) return true;
  V3 v2 = S(n);
  if (dot<3>(v2, n) < 0.0) return false;
  n = cross(vsub(v1, v0), vsub(v2, v0));
  if (dot<3>(n, v0) > 0.0) {
    std::swap(v1, v2);
    n = vneg(n);
  }
  V3 v3;
  int count = 0;
  for (;;) {
    count += 1;
    if (count >= MAX_ITER) return true;
    v3 = S(n);
    if (dot<3>(v3, n) <= 0.0) return false;
    if (dot<3>(cross(v1, v3), v0) < 0.0) {
      v2 = v3;
      n = cross(vsub(v1, v0), vsub(v2, v0));
      continue;
    }
    if (dot<3>(cross(v3, v2), v0) < 0.0) {
      v1 = v3;
      n = cross(vsub(v1, v0), vsub(v2, v0));
      continue;
    }
    break;
  }
  V3 p[3] = {v1, v2, v3};
  // refinement
  count = 0;
  for (;;) {
    count += 1;
    V3 nrm = cross(vsub(p[1], p[0]), vsub(p[2], p[0]));
    if (dot<3>(vsub(p[0], v0), nrm) < 0.0) nrm = vneg(nrm);
    if (dot<3>(p[0], nrm) >= 0.0) return true;
    V3 vn = S(nrm);
    if (dot<3>(vn, nrm) < 0.0) return false;
    const double tol = TOL * vnorm(nrm);
    if (std::fabs(dot<3>(vsub(vn, p[0]), nrm)) < tol) return false;
    if (std::fabs(dot<3>(p[0], nrm)) < tol) return true;
    V3 vp = cross(vn, v0);
    const bool t1 = dot<3>(vp, p[0]) > 0.0, t2 = dot<3>(vp, p[1]) > 0.0, t3 = dot<3>(vp, p[2]) > 0.0;
    if (t1 && t2)
      p[0] = vn;
    else if (t1)
      p[2] = vn;
    else if (t3)
      p[1] = vn;
    else
      p[0] = vn;
    if (count >= MAX_ITER) return true;
  }
}

// Convex::intersects_at_global (hoomd-geometry/src/convex.rs:154-170) with
// pair_system_to_local (hoomd-vector/src/lib.rs:878-888).
inline bool intersects_at_global(const Polytope& s, const V3& ri, const Quat& qi, const V3& rj, const Quat& qj) {
  const double rcut = s.radius + s.radius;
  if (distance_squared<3>(ri, rj) >= rcut * rcut) return false;
  const V3 rab = vsub(rj, ri);
  const Quat inv = qconj(qi);
  const V3 vij = qrotate(inv, rab);
  const Quat oij = qmul(inv, qj);
  return collide3d(s, s, vij, oij);
}

// VersorDisplacement::sample (hoomd-mc/src/rotate/versor.rs:39-72): Normal
// (0, sigma) = 0 + sigma * StandardNormal per component (rand_distr 0.6),
// theta = |s|, rejected above pi, exponential map, then normalised by
// to_versor_unchecked (quaternion.rs:308-310: q / |q|,
// |q| = sqrt(s*s + dot(v, v))).
inline Quat versor_displacement(Sfc64& rng, double sigma) {
  const double PI = 3.141592653589793;
  for (;;) {
    V3 s;
    for (int i = 0; i < 3; ++i) s[i] = 0.0 + sigma * standard_normal(rng);
    const double theta = vnorm(s);
    if (theta > PI) continue;
    const double half = 0.5 * theta;
    double sn, w;
    lib_sincos(half, &sn, &w);  // half.cos(), half.sin(): one sincos call
    double f = sn / theta;
    if (!std::isfinite(f)) f = 0.5;
    const V3 v = vscale(s, f);
    const double mag = std::sqrt(w * w + dot<3>(v, v));
    Quat q;
    q.s = w / mag;
    for (int i = 0; i < 3; ++i) q.v[i] = v[i] / mag;
    return q;
  }
}

struct CSite {
  uint64_t site_tag = 0, body_tag = 0;
  V3 pos{};
  Quat q{};
};

// Microstate<OrientedPoint, OrientedPoint, VecCell, Periodic<Hypercuboid<3>>>
// with one site per body at the body origin, and ParallelSweep.
struct ConvexSystem {
  uint64_t step = 0;
  uint32_t substep = 0, seed = 0;
  std::vector<V3> bpos;
  std::vector<Quat> bq;
  VecWithTags<CSite> sites;
  std::vector<std::vector<uint64_t>> sites_ghosts;
  VecWithTags<CSite> ghosts;
  Periodic<3> boundary;
  VecCell<3> spatial;
  Polytope shape;

  double body_range = 1.0;
  Checkerboard<3> cb;
  std::vector<std::vector<uint64_t>> spaces;
  struct BodyTrial {
    V3 pos{};
    Quat q{};
    uint64_t body_index = 0;
    int8_t status = EmptySpace;
  };
  std::vector<BodyTrial> trials;
  bool tracing = false;
  std::vector<TraceRecord> trace;

  static Polytope make_shape(const std::vector<V3>& verts) {
    Polytope p;
    p.set(verts);
    return p;
  }

  // VecCell::with_search_radius(maximum interaction range = 2 R) (hard_shape.rs:111-118)
  ConvexSystem(const std::array<double, 3>& L, double boundary_range, const std::vector<V3>& verts)
      : spatial(make_shape(verts).radius * 2.0), shape(make_shape(verts)) {
    boundary.L = L;
    boundary.range = boundary_range;
  }

  double range() const { return shape.radius * 2.0; }

  // OrientedPoint::transform (property/oriented_point.rs:107-113) of the
  // site at the origin with the identity orientation, then wrapped.
  void site_of(const V3& b, const Quat& q, V3& sp, Quat& sq) const {
    const V3 r = qrotate(q, V3{0.0, 0.0, 0.0});
    sp = boundary.wrap(vadd(b, r));
    sq = qmul(q, Quat{});
  }

  void update_site_ghosts(int64_t site_index) {
    V3 ng[8];
    const CSite site = sites.items[site_index];
    int n = boundary.ghosts(site.pos, ng);
    auto& gt = sites_ghosts[site_index];
    if ((int)gt.size() < n) {
      int add = n - (int)gt.size();
      for (int k = 0; k < add; ++k) {
        CSite g;
        g.site_tag = site.site_tag;
        g.body_tag = site.body_tag;
        gt.push_back(ghosts.push(g));
      }
    } else if ((int)gt.size() > n) {
      int rem = (int)gt.size() - n;
      for (int k = 0; k < rem; ++k) {
        uint64_t tag = gt[gt.size() - 1 - k];
        ghosts.remove(ghosts.indices[tag]);
        spatial.remove(key_ghost(tag));
      }
      gt.resize(n);
    }
    for (int k = 0; k < n; ++k) {
      uint64_t tag = gt[k];
      spatial.insert(key_ghost(tag), ng[k]);
      CSite& g = ghosts.items[ghosts.indices[tag]];
      g.pos = ng[k];
      g.q = site.q;
    }
  }

  void add_body(const V3& position, const Quat& q) {
    uint64_t body_tag = bpos.size();
    V3 b = boundary.wrap(position);
    CSite s;
    s.site_tag = sites.indices.size();
    if (!sites.free_tags.empty()) s.site_tag = sites.free_tags.top();
    s.body_tag = body_tag;
    site_of(b, q, s.pos, s.q);
    spatial.insert(key_primary(s.site_tag), s.pos);
    sites.push(s);
    sites_ghosts.emplace_back();
    bpos.push_back(b);
    bq.push_back(q);
    update_site_ghosts(sites.indices[s.site_tag]);
  }

  void update_body_properties(uint64_t i, const V3& pos, const Quat& q) {
    V3 nb = boundary.wrap(pos);
    bpos[i] = nb;
    bq[i] = q;
    int64_t si = sites.indices[i];
    CSite& s = sites.items[si];
    site_of(nb, q, s.pos, s.q);
    spatial.insert(key_primary(i), s.pos);
    update_site_ghosts(si);
  }

  const CSite& site_by_key(uint64_t key) const {
    uint64_t tag = key >> 1;
    if (key & 1) return ghosts.items[ghosts.indices[tag]];
    return sites.items[sites.indices[tag]];
  }

  // pairwise_cutoff.rs:355-387 + hard_shape.rs:79-98: infinite on the first overlap
  bool overlaps(const V3& p, const Quat& q, uint64_t own) const {
    bool hit = false;
    spatial.points_near_ball(p, range(), [&](uint64_t key) {
      const CSite& sj = site_by_key(key);
      if (sj.body_tag != own && intersects_at_global(shape, p, q, sj.pos, sj.q)) {
        hit = true;
        return false;
      }
      return true;
    });
    return hit;
  }

  // any overlapping pair in the current state (site tags i < j)
  bool any_overlap() const {
    for (const CSite& si : sites.items) {
      bool hit = false;
      spatial.points_near_ball(si.pos, range(), [&](uint64_t key) {
        const CSite& sj = site_by_key(key);
        if (si.site_tag < sj.site_tag && si.body_tag != sj.body_tag &&
            intersects_at_global(shape, si.pos, si.q, sj.pos, sj.q)) {
          hit = true;
          return false;
        }
        return true;
      });
      if (hit) return true;
    }
    return false;
  }

  bool update_checkerboard() {
    Sfc64 rng = counter_rng(step, substep, seed, UINT64_MAX);
    if (!cb.update(rng, body_range, boundary.L)) return false;
    spaces.resize(cb.num_spaces());
    for (auto& s : spaces) s.clear();
    for (uint64_t b = 0; b < bpos.size(); ++b) {
      int64_t s = cb.point_to_space_index(bpos[b]);
      if (s < 0) return false;
      spaces[s].push_back(b);
    }
    return true;
  }

  // parallel_sweep.rs:209-279 for Translate<Cartesian<3>> (rotate = false)
  // or Rotate<Versor> (rotate = true; rotate/versor.rs:111-122: new
  // orientation = delta * old, position unchanged but wrapped again)
  void generate_trial_moves(const std::vector<uint64_t>& space_indices, bool rotate, double max_move) {
    trials.resize(space_indices.size());
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t k = 0; k < (int64_t)space_indices.size(); ++k) {
      BodyTrial& bt = trials[k];
      bt.status = EmptySpace;
      uint64_t s = space_indices[k];
      Sfc64 rng = counter_rng(step, substep, seed, s);
      const auto& list = spaces[s];
      if (list.empty()) continue;
      uint64_t bi = list[rand_index(rng, (uint32_t)list.size())];
      bt.body_index = bi;
      V3 np = bpos[bi];
      Quat nq = bq[bi];
      if (rotate)
        nq = qmul(versor_displacement(rng, max_move), nq);
      else
        np = System<3>::propose_translate(rng, np, max_move);
      np = boundary.wrap(np);
      if (cb.point_to_space_index(np) != (int64_t)s) {
        bt.status = Rejected;
        continue;
      }
      bt.pos = np;
      bt.q = nq;
      V3 sp;
      Quat sq;
      site_of(np, nq, sp, sq);
      // hard shapes: dH is 0 or +inf, u is never drawn
      bt.status = overlaps(sp, sq, bi) ? Rejected : Accepted;
    }
  }

  // tune_local_trial (hoomd-mc/src/tune_local.rs:21-89) with Translate or
  // Rotate<Versor>, all bodies eligible, then tune_by_scaling (lib.rs:514-524)
  // and the trial's adjust (translate.rs:87-92; rotate/versor.rs:125-137: a
  // rotation size at most pi/2).  Hard shapes: no acceptance draw.
  void tune(bool rotate, double target, uint64_t samples, uint32_t steps, double& size, uint64_t* counts) const {
    Sfc64 rng = counter_rng(step, substep, seed, 0);
    const uint64_t n = bpos.size();
    for (uint32_t t = 0; t < steps; ++t) {
      uint64_t acc = 0, rej = 0;
      while (acc + rej < samples) {
        const uint64_t bi = rand_index(rng, (uint32_t)n);
        V3 np = bpos[bi];
        Quat nq = bq[bi];
        if (rotate)
          nq = qmul(versor_displacement(rng, size), nq);
        else
          np = System<3>::propose_translate(rng, np, size);
        np = boundary.wrap(np);
        V3 sp;
        Quat sq;
        site_of(np, nq, sp, sq);
        if (overlaps(sp, sq, bi))
          ++rej;
        else
          ++acc;
      }
      counts[2 * t] = acc;
      counts[2 * t + 1] = rej;
      if (acc + rej > 0) {
        const double ratio = (double)acc / (double)(acc + rej);
        size *= (ratio + 1.5) / (target + 1.5);
        if (rotate && size > M_PI / 2.0) size = M_PI / 2.0;
      }
    }
  }

  void update_bodies(uint64_t& acc, uint64_t& rej) {
    for (const BodyTrial& bt : trials) {
      if (bt.status == Rejected) {
        ++rej;
      } else if (bt.status == Accepted) {
        update_body_properties(bt.body_index, bt.pos, bt.q);
        ++acc;
      }
    }
  }

  bool parallel_apply(bool rotate, double max_move, uint64_t& acc, uint64_t& rej) {
    acc = rej = 0;
    if (!update_checkerboard()) return false;
    uint64_t n = bpos.size();
    while (acc + rej < n) {
      for (uint32_t c = 0; c < cb.by_color.size(); ++c) {
        const auto& si = cb.by_color[c];
        generate_trial_moves(si, rotate, max_move);
        if (tracing) {
          for (size_t k = 0; k < si.size(); ++k) {
            const BodyTrial& bt = trials[k];
            TraceRecord tr{substep, c, si[k], bt.status == EmptySpace ? -1 : (int64_t)bt.body_index, bt.status,
                           0};
            trace.push_back(tr);
          }
        }
        update_bodies(acc, rej);
      }
      substep += 1;
    }
    return true;
  }

  void increment_step() {
    step += 1;
    substep = 0;
  }
};

}  // namespace oracle
