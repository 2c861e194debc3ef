// Oracle (test infrastructure only): restatement of hoomd-microstate's
// Microstate for single-site point bodies, hoomd-interaction's PairwiseCutoff
// energies (HardSphere, Isotropic<LennardJones>, Isotropic<Shifted<LJ>>), and
// hoomd-mc's Translate, ParallelSweep and Sweep.
//
// Data structures follow the reference (array-of-structures bodies and sites,
// VecCell, ghost sites, serial application of accepted moves) so that the same
// object serves as the parity oracle and as the timed CPU baseline.
#pragma once
#include <omp.h>

#include <cmath>
#include <cstdint>
#include <functional>
#include <limits>
#include <queue>
#include <vector>

#include "geometry.hpp"
#include "rng.hpp"
#include "vec_cell.hpp"

namespace oracle {

constexpr double INF = std::numeric_limits<double>::infinity();

// hoomd-microstate/src/microstate.rs:54-138 (VecWithTags): items in index
// order, tag -> index map, smallest free tag reused first.
template <class T>
struct VecWithTags {
  std::vector<T> items;
  std::vector<uint64_t> tags;
  std::vector<int64_t> indices;
  std::priority_queue<uint64_t, std::vector<uint64_t>, std::greater<uint64_t>> free_tags;

  uint64_t push(const T& item) {
    uint64_t tag;
    if (!free_tags.empty()) {
      tag = free_tags.top();
      free_tags.pop();
    } else {
      tag = indices.size();
    }
    int64_t index = (int64_t)items.size();
    items.push_back(item);
    tags.push_back(tag);
    if (tag == indices.size())
      indices.push_back(index);
    else
      indices[tag] = index;
    return tag;
  }
  void remove(int64_t index) {
    uint64_t removed = tags[index];
    items[index] = items.back();
    items.pop_back();
    tags[index] = tags.back();
    tags.pop_back();
    if ((size_t)index < items.size()) indices[tags[index]] = index;
    indices[removed] = -1;
    free_tags.push(removed);
  }
};

template <int D>
struct Site {
  uint64_t site_tag = 0, body_tag = 0;
  Vec<D> pos{};
};

// Pair energies.  kind: 0 HardSphere, 1 Isotropic<LennardJones>, 2 Isotropic<Shifted<LennardJones>>,
// 4 Isotropic<Expanded<OverlapPenalty>> with r_cut = delta (the placement
// Hamiltonian of benchmarks/src/mc/hard_sphere.rs:192-198).
struct PairModel {
  int kind = 0;
  double diameter = 1.0;          // HardSphere (hard_shape.rs:127-174)
  double epsilon = 1.0, sigma = 1.0, r_cut = 2.5, r_shift = 2.5;  // LJ
  double k = 10000.0, max_overlap = 0.2, eps_shoulder = 100.0, delta = 1.0;  // OverlapPenalty

  double range() const { return kind == 0 ? diameter : r_cut; }
  bool only_infinite_or_zero() const { return kind == 0; }

  // univariate/lennard_jones.rs:107-111 (powi expands to the square chain)
  double lj(double r) const {
    double s = sigma / r;
    double s2 = s * s, s4 = s2 * s2, s8 = s4 * s4;
    double s12 = s4 * s8, s6 = s2 * s4;
    return 4.0 * epsilon * (s12 - s6);
  }
  // univariate/overlap_penalty.rs:109-115 at x = r - delta (expanded.rs:62-66);
  // powi(2) is x * x
  double overlap_penalty(double r) const {
    double x = r - delta;
    if (x < -max_overlap) return INF;
    if (x < 0.0) return 0.5 * k * (x * x) + eps_shoulder;
    return 0.0;
  }
  template <int D>
  double pair(const Vec<D>& a, const Vec<D>& b) const {
    double r2 = distance_squared<D>(a, b);
    if (kind == 0) return r2 < diameter * diameter ? INF : 0.0;  // hard_shape.rs:143-153
    if (r2 >= r_cut * r_cut) return 0.0;                          // isotropic.rs:147-156
    double r = std::sqrt(r2);
    if (kind == 4) return overlap_penalty(r);
    if (kind == 1) return lj(r);
    return lj(r) - lj(r_shift);                                    // shifted.rs:92-94
  }
};

enum TrialStatus : int8_t { Rejected = 0, Accepted = 1, EmptySpace = 2 };

struct TraceRecord {
  uint32_t substep;
  uint32_t color;
  uint64_t space;
  int64_t body;  // -1 for EmptySpace
  int32_t status;
  int32_t pad;
};

template <int D>
struct System {
  // Microstate (hoomd-microstate/src/microstate.rs:192-226)
  uint64_t step = 0;
  uint32_t substep = 0, seed = 0;
  std::vector<Vec<D>> bodies;  // body properties (Point), index == tag
  VecWithTags<Site<D>> sites;
  std::vector<std::vector<uint64_t>> sites_ghosts;  // per site index
  VecWithTags<Site<D>> ghosts;
  Periodic<D> boundary;
  VecCell<D> spatial;
  PairModel model;

  // ParallelSweep caches (hoomd-mc/src/parallel_sweep.rs:83-100)
  double body_range = 1.0;
  Checkerboard<D> cb;
  std::vector<std::vector<uint64_t>> spaces;
  struct BodyTrial {
    Vec<D> pos{};
    uint64_t body_index = 0;
    int8_t status = EmptySpace;
  };
  std::vector<BodyTrial> trials;

  bool tracing = false;
  std::vector<TraceRecord> trace;

  // External(Harmonic) (hoomd-interaction/src/external_type.rs:341-375 with
  // the Harmonic site energy of hoomd-mc/src/sweep.rs:399-405), added to the
  // pair energies when ext_k > 0: 1.0 / 2.0 * K * (x - c).norm_squared(),
  // norm_squared = dot = fold(0, fma) (hoomd-vector/src/cartesian.rs:293-296).
  double ext_k = 0.0;
  Vec<D> ext_c{};
  double external(const Vec<D>& x) const {
    double ns = 0.0;
    for (int i = 0; i < D; ++i) {
      const double d = x[i] - ext_c[i];
      ns = std::fma(d, d, ns);
    }
    return 1.0 / 2.0 * ext_k * ns;
  }

  System(const std::array<double, D>& L, double boundary_range, double spatial_radius)
      : spatial(spatial_radius) {
    boundary.L = L;
    boundary.range = boundary_range;
  }

  Vec<D> add_site_local(const Vec<D>& body) const {  // property/point.rs:67-71 (+ site at origin)
    Vec<D> s;
    for (int i = 0; i < D; ++i) s[i] = body[i] + 0.0;
    return s;
  }

  // microstate.rs:661-706
  void update_site_ghosts(int64_t site_index) {
    Vec<D> ng[1 << D];
    const Site<D> site = sites.items[site_index];
    int n = boundary.ghosts(site.pos, ng);
    auto& gt = sites_ghosts[site_index];
    if ((int)gt.size() < n) {
      int add = n - (int)gt.size();
      for (int k = 0; k < add; ++k) {
        Site<D> g;
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
      ghosts.items[ghosts.indices[tag]].pos = ng[k];
    }
  }

  // microstate.rs:777-835 (periodic wrap never fails)
  void add_body(const Vec<D>& position) {
    uint64_t body_tag = bodies.size();
    Vec<D> b = boundary.wrap(position);
    Site<D> s;
    s.site_tag = sites.indices.size();
    if (!sites.free_tags.empty()) s.site_tag = sites.free_tags.top();
    s.body_tag = body_tag;
    s.pos = boundary.wrap(add_site_local(b));
    spatial.insert(key_primary(s.site_tag), s.pos);
    sites.push(s);
    sites_ghosts.emplace_back();
    bodies.push_back(b);
    update_site_ghosts(sites.indices[s.site_tag]);
  }

  // microstate.rs:985-1034 (body index == site tag for single-site bodies)
  void update_body_properties(uint64_t body_index, const Vec<D>& props) {
    Vec<D> nb = boundary.wrap(props);
    bodies[body_index] = nb;
    uint64_t site_tag = body_index;
    int64_t si = sites.indices[site_tag];
    Vec<D> sp = boundary.wrap(add_site_local(nb));
    spatial.insert(key_primary(site_tag), sp);
    sites.items[si].pos = sp;
    update_site_ghosts(si);
  }

  const Site<D>& site_by_key(uint64_t key) const {
    uint64_t tag = key >> 1;
    if (key & 1) return ghosts.items[ghosts.indices[tag]];
    return sites.items[sites.indices[tag]];
  }

  // pairwise_cutoff.rs:355-387 for one site, filter body_tag != own.
  double site_energy(const Vec<D>& site_i, uint64_t own_body, bool initial) const {
    double energy = 0.0;
    double r = model.range();
    if (!(r <= boundary.range)) std::abort();  // microstate.rs:1436-1439
    bool inf = false;
    spatial.points_near_ball(site_i, r, [&](uint64_t key) {
      const Site<D>& sj = site_by_key(key);
      if (sj.body_tag != own_body) {
        double one = (initial && model.only_infinite_or_zero()) ? 0.0 : model.pair<D>(site_i, sj.pos);
        if (one == INF) {
          inf = true;
          return false;
        }
        energy += one;
      }
      return true;
    });
    return inf ? INF : energy;
  }

  // pairwise_cutoff.rs:919-941 with :420-488
  double delta_energy_one(uint64_t body_index, const Vec<D>& trial_body) const {
    double ef = 0.0;
    {
      Vec<D> si = boundary.wrap(add_site_local(trial_body));
      double one = site_energy(si, body_index, false);
      if (one == INF) return INF;
      ef += one;
    }
    double ei = 0.0;
    if (!model.only_infinite_or_zero()) {
      const Site<D>& s = sites.items[sites.indices[body_index]];
      double one = site_energy(s.pos, body_index, true);
      if (one == INF)
        ei = INF;
      else
        ei += one;
    }
    double dh = ef - ei;
    if (ext_k > 0.0) {  // the sum of the two Hamiltonians' delta_energy_one
      const Site<D>& s = sites.items[sites.indices[body_index]];
      dh += external(boundary.wrap(add_site_local(trial_body))) - (0.0 + external(s.pos));
    }
    return dh;
  }

  // pairwise_cutoff.rs:1017-1025: the new body has no tag, nothing is filtered
  double delta_energy_insert(const Vec<D>& body) const {
    Vec<D> si = boundary.wrap(add_site_local(body));
    return site_energy(si, UINT64_MAX, false);
  }

  // pairwise_cutoff.rs:720-745
  double total_energy() const {
    double total = 0.0;
    double r = model.range();
    for (const Site<D>& si : sites.items) {
      double e = 0.0;
      bool inf = false;
      spatial.points_near_ball(si.pos, r, [&](uint64_t key) {
        const Site<D>& sj = site_by_key(key);
        if (si.site_tag < sj.site_tag && si.body_tag != sj.body_tag) {
          double one = model.pair<D>(si.pos, sj.pos);
          if (one == INF) {
            inf = true;
            return false;
          }
          e += one;
        }
        return true;
      });
      if (inf) return INF;
      total += e;
    }
    return total;
  }

  // Translate<Cartesian<D>>::propose (translate/cartesian.rs:39-52) with
  // Ball::sample (hoomd-vector/src/distribution.rs:39-70).
  static Vec<D> propose_translate(Sfc64& rng, Vec<D> pos, double r) {
    Vec<D> v;
    if (D == 2) {
      for (;;) {
        double c[D];
        for (int i = 0; i < D; ++i) c[i] = rand_f64(rng);
        for (int i = 0; i < D; ++i) v[i] = (c[i] * 2.0 - 1.0) * r;
        if (dot<D>(v, v) < r * r) break;
      }
    } else {
      Vec<D> p;
      for (int i = 0; i < D; ++i) p[i] = standard_normal(rng);
      double norm = std::sqrt(dot<D>(p, p));
      double q = r / norm;
      double s = q * std::cbrt(rand_f64(rng));
      for (int i = 0; i < D; ++i) v[i] = p[i] * s;
    }
    for (int i = 0; i < D; ++i) pos[i] += v[i];
    return pos;
  }

  // parallel_sweep.rs:176-201.  Returns false where the reference panics.
  bool update_checkerboard() {
    Sfc64 rng = counter_rng(step, substep, seed, UINT64_MAX);
    if (!cb.update(rng, body_range, boundary.L)) return false;
    spaces.resize(cb.num_spaces());
    for (auto& s : spaces) s.clear();
    for (uint64_t b = 0; b < bodies.size(); ++b) {
      int64_t s = cb.point_to_space_index(bodies[b]);
      if (s < 0) return false;
      spaces[s].push_back(b);
    }
    return true;
  }

  // parallel_sweep.rs:209-279
  void generate_trial_moves(const std::vector<uint64_t>& space_indices, double kt, double max_move) {
    trials.resize(space_indices.size());
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t k = 0; k < (int64_t)space_indices.size(); ++k) {
      BodyTrial& bt = trials[k];
      bt.status = EmptySpace;
      uint64_t s = space_indices[k];
      Sfc64 rng = counter_rng(step, substep, seed, s);
      const auto& list = spaces[s];
      if (list.empty()) continue;
      uint64_t bi = list[rand_index(rng, (uint32_t)list.size())];
      bt.body_index = bi;
      Vec<D> np = boundary.wrap(propose_translate(rng, bodies[bi], max_move));
      if (cb.point_to_space_index(np) != (int64_t)s) {
        bt.status = Rejected;
        continue;
      }
      bt.pos = np;
      double dh = delta_energy_one(bi, np);
      if (dh != INF && (dh <= 0.0 || lib_lt_exp(rand_f64(rng), -dh / kt)))
        bt.status = Accepted;
      else
        bt.status = Rejected;
    }
  }

  // parallel_sweep.rs:283-314
  void update_bodies(uint64_t& acc, uint64_t& rej) {
    for (const BodyTrial& bt : trials) {
      if (bt.status == Rejected) {
        ++rej;
      } else if (bt.status == Accepted) {
        update_body_properties(bt.body_index, bt.pos);
        ++acc;
      }
    }
  }

  // parallel_sweep.rs:390-421.  Returns false where the reference panics.
  bool parallel_apply(double kt, double max_move, uint64_t& acc, uint64_t& rej) {
    acc = rej = 0;
    if (!update_checkerboard()) return false;
    uint64_t n = bodies.size();
    while (acc + rej < n) {
      for (uint32_t c = 0; c < cb.by_color.size(); ++c) {
        const auto& si = cb.by_color[c];
        generate_trial_moves(si, kt, max_move);
        if (tracing) {
          for (size_t k = 0; k < si.size(); ++k) {
            const BodyTrial& bt = trials[k];
            TraceRecord tr{substep, c, si[k], bt.status == EmptySpace ? -1 : (int64_t)bt.body_index,
                           bt.status, 0};
            trace.push_back(tr);
          }
        }
        update_bodies(acc, rej);
      }
      substep += 1;
    }
    return true;
  }

  // hoomd-mc/src/sweep.rs:159-224 (serial, one RNG stream, body order)
  void sweep_apply(double kt, double max_move, uint64_t& acc, uint64_t& rej) {
    acc = rej = 0;
    Sfc64 rng = counter_rng(step, substep, seed, 0);
    for (uint64_t bi = 0; bi < bodies.size(); ++bi) {
      Vec<D> np = boundary.wrap(propose_translate(rng, bodies[bi], max_move));
      double dh = delta_energy_one(bi, np);
      if (dh != INF && (dh <= 0.0 || lib_lt_exp(rand_f64(rng), -dh / kt))) {
        update_body_properties(bi, np);
        ++acc;
      } else {
        ++rej;
      }
    }
    substep += 1;
  }

  // QuickInsert::apply (hoomd-mc/src/quick_insert.rs:286-334) with a
  // UniformIn distribution over the periodic box (uniform_in.rs:99-110:
  // one Uniform(-L/2, L/2) draw per axis, cuboid.rs:590-598), single-site
  // point bodies.  qi = {target, allowed_overlaps, inserted, complete}.
  void quick_insert_apply(uint64_t* qi, uint64_t& acc, uint64_t& rej) {
    acc = rej = 0;
    if (qi[3]) return;
    const double energy = total_energy();
    if (energy <= 0.0 && qi[2] >= qi[0]) {
      qi[3] = 1;
      return;
    }
    if (energy <= 0.0) {
      Sfc64 rng = counter_rng(step, substep, seed, 0);
      uint64_t overlapping = 0;
      for (uint64_t k = 0; k < qi[0]; ++k) {
        Vec<D> p;
        for (int a = 0; a < D; ++a) p[a] = rand_uniform(rng, -boundary.L[a] / 2.0, boundary.L[a] / 2.0);
        const double de = delta_energy_insert(p);
        if (std::isfinite(de)) {  // Microstate::add_body on a periodic boundary cannot fail
          add_body(p);
          acc += 1;
          qi[2] += 1;
          if (de > 0.0) overlapping += 1;
          if (qi[2] == qi[0] || overlapping >= qi[1]) break;
        } else {
          rej += 1;
        }
      }
    }
    substep += 1;
  }

  // tune_local_trial (hoomd-mc/src/tune_local.rs:21-89) with Translate, all
  // bodies eligible (parallel_sweep.rs:486-506), then tune_by_scaling
  // (lib.rs:514-524: GAMMA = 1.5) and Translate::adjust (translate.rs:87-92).
  // One stream for every step; the microstate is not modified.  body =
  // random_range(0..n): rand samples a usize range that fits in u32 as a
  // u32 (rand_index).  counts: accepted, rejected per step.
  void tune_translate(double kt, double target, uint64_t samples, uint32_t steps, double& max_move,
                      uint64_t* counts) const {
    Sfc64 rng = counter_rng(step, substep, seed, 0);
    const uint64_t n = bodies.size();
    for (uint32_t t = 0; t < steps; ++t) {
      uint64_t acc = 0, rej = 0;
      while (acc + rej < samples) {
        const uint64_t bi = rand_index(rng, (uint32_t)n);
        const Vec<D> np = boundary.wrap(propose_translate(rng, bodies[bi], max_move));
        const double dh = delta_energy_one(bi, np);
        if (dh != INF && (dh <= 0.0 || lib_lt_exp(rand_f64(rng), -dh / kt)))
          ++acc;
        else
          ++rej;
      }
      counts[2 * t] = acc;
      counts[2 * t + 1] = rej;
      if (acc + rej > 0) {
        const double ratio = (double)acc / (double)(acc + rej);
        max_move *= (ratio + 1.5) / (target + 1.5);
      }
    }
  }

  // pairwise_cutoff.rs:1097-1109 with :458-488: minus the initial energy of
  // the body's site with the other bodies' sites (0 for hard shapes)
  double delta_energy_remove(uint64_t body_index) const {
    double e = 0.0;
    if (!model.only_infinite_or_zero()) {
      const Site<D>& s = sites.items[sites.indices[body_index]];
      const double one = site_energy(s.pos, body_index, true);
      if (one == INF) return -INF;
      e += one;
    }
    return -e;
  }

  // The microstate of another body list, kept as stored (no wrap: wrap is
  // not idempotent in floating point).  Tags are renumbered in index order;
  // the reference reuses freed tags, which enter no decision here.
  void reset_bodies(const std::vector<Vec<D>>& nb, const std::vector<Vec<D>>& ns) {
    bodies.clear();
    sites = VecWithTags<Site<D>>{};
    sites_ghosts.clear();
    ghosts = VecWithTags<Site<D>>{};
    spatial = VecCell<D>(spatial.w);
    for (size_t k = 0; k < nb.size(); ++k) {
      Site<D> s;
      s.site_tag = k;
      s.body_tag = k;
      s.pos = ns[k];
      spatial.insert(key_primary(k), s.pos);
      sites.push(s);
      sites_ghosts.emplace_back();
      bodies.push_back(nb[k]);
      update_site_ghosts((int64_t)k);
    }
  }

  // GrandCanonical(UniformIn over the periodic box)::apply
  // (hoomd-mc/src/grand_canonical.rs:145-205).  The bool draw is rand's
  // StandardUniform bool, the sign bit of a u32 (rand 0.10.4, PARITY
  // UNPINNED).  cnt = {insert accepted, insert rejected, remove accepted,
  // remove rejected}, incremented.
  void grand_canonical_apply(double kt, double fugacity, uint64_t* cnt) {
    const uint64_t n = bodies.size();
    double volume = 1.0;
    for (int i = 0; i < D; ++i) volume *= boundary.L[i];
    Sfc64 rng = counter_rng(step, substep, seed, 0);
    if (n == 0 && fugacity <= 0.0) {
      substep += 1;
      return;
    }
    const bool insert = (int32_t)rng.next_u32() < 0;
    if (insert) {
      Vec<D> p;
      for (int a = 0; a < D; ++a) p[a] = rand_uniform(rng, -boundary.L[a] / 2.0, boundary.L[a] / 2.0);
      const double de = delta_energy_insert(p);
      if (std::isfinite(de)) {
        const double p_insert = (volume * fugacity * lib_exp(-de / kt)) / ((double)n + 1.0);
        if (p_insert > rand_f64(rng)) {
          add_body(p);
          cnt[0] += 1;
        } else {
          cnt[1] += 1;
        }
      } else {
        cnt[1] += 1;
      }
    } else if (n > 0) {
      const uint64_t index = rand_index(rng, (uint32_t)n);
      const double de = delta_energy_remove(index);
      const double p_remove = fugacity <= 0.0 ? 1.0 : ((double)n * lib_exp(-de / kt)) / (volume * fugacity);
      if (p_remove > rand_f64(rng)) {
        // Microstate::remove_body (microstate.rs:917-943): swap_remove
        std::vector<Vec<D>> nb = bodies, ns(n);
        for (uint64_t k = 0; k < n; ++k) ns[k] = sites.items[sites.indices[k]].pos;
        nb[index] = nb[n - 1];
        ns[index] = ns[n - 1];
        nb.pop_back();
        ns.pop_back();
        reset_bodies(nb, ns);
        cnt[2] += 1;
      } else {
        cnt[3] += 1;
      }
    }
    substep += 1;
  }

  // microstate.rs:412-415
  void increment_step() {
    step += 1;
    substep = 0;
  }
};

}  // namespace oracle
