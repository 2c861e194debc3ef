// Initial-state preparation through the C ABI: QuickCompress
// (hoomd-mc/src/quick_compress.rs) over device energies and boundary clones,
// QuickInsert (hoomd-mc/src/quick_insert.rs) over device insertion energies,
// the move-size tuner (hoomd-mc/src/tune_local.rs) over device trial
// evaluations, and GrandCanonical (hoomd-mc/src/grand_canonical.rs) over
// device insertion and removal energies.
#include "host.hpp"

#include <unordered_map>

extern "C" {

// ------------------------------------------------------------ QuickCompress --
// One draw of Counter(step, substep, seed).make_rng() (hoomd-rand/src/counter.rs:160-167,
// index 0): SFC64 (hoomd-rand/src/sfc.rs:49-119), 12 warm-up rounds.
static uint64_t host_counter_u64(uint64_t step, uint32_t substep, uint32_t seed) {
  uint64_t a = step, b = ((uint64_t)substep << 32) | (uint64_t)seed, c = 0, ctr = 0;
  uint64_t out = 0;
  for (int i = 0; i < 13; ++i) {
    out = a + b + ctr;
    a = b ^ (b >> 11);
    b = c + (c << 3);
    c = ((c << 24) | (c >> 40)) + out;
    ctr += 1;
  }
  return out;
}

// rand's Uniform::new(low, high) and sample for f64 (rand 0.10): scale =
// high - low, lowered by one ulp at a time while low + scale (1 - 2^-52) >=
// high; a sample is ((bits >> 12) as the mantissa of [1, 2) - 1) scale + low.
static double host_uniform(uint64_t bits, double low, double high) {
  double scale = high - low;
  const double max_rand = 1.0 - 0x1p-52;
  while (scale * max_rand + low >= high) scale = std::nextafter(scale, 0.0);
  uint64_t m = (bits >> 12) | 0x3FF0000000000000ull;
  double v12;
  std::memcpy(&v12, &m, sizeof v12);
  return (v12 - 1.0) * scale + low;
}

// Counter(step, substep, seed).make_rng() as a stream (index 0).
struct HostSfc64 {
  uint64_t a, b, c, ctr;
  uint64_t next() {
    const uint64_t out = a + b + ctr;
    a = b ^ (b >> 11);
    b = c + (c << 3);
    c = ((c << 24) | (c >> 40)) + out;
    ctr += 1;
    return out;
  }
};
static HostSfc64 host_counter_rng(uint64_t step, uint32_t substep, uint32_t seed) {
  HostSfc64 r{step, ((uint64_t)substep << 32) | (uint64_t)seed, 0, 0};
  for (int i = 0; i < 12; ++i) (void)r.next();
  return r;
}

static double box_volume(const Geometry& g) {  // Hypercuboid::volume: the product of the edges
  double v = 1.0;
  for (int a = 0; a < g.dim; ++a) v *= g.L[a];
  return v;
}

void hmc_quick_compress_init(hmc_quick_compress* qc, double target_volume) {
  if (!qc) return;
  std::memset(qc, 0, sizeof *qc);
  qc->target_volume = target_volume;
  qc->maximum_energy_per_site = 1000.0;
  qc->maximum_delta = 0.05;
  qc->state = HMC_QC_STARTUP;
}

int hmc_quick_compress_apply(hmc_quick_compress* qc, hmc_ctx** pc, uint64_t step, uint32_t substep,
                             uint32_t seed) {
  if (!qc || !pc || !*pc) return HMC_ERR_INVALID;
  hmc_ctx* c = *pc;
  if (!(qc->target_volume > 0.0) || !(qc->maximum_delta > 0.0 && qc->maximum_delta < 1.0))
    return fail(c, HMC_ERR_INVALID, "QuickCompress: target volume > 0 and 0 < maximum_delta < 1 are required");
  if (!c->have_state) return fail(c, HMC_ERR_STATE, "no state");
  const Geometry& g = c->g;
  const int D = g.dim;
  auto same_boundary = [&](const double* e) {
    for (int a = 0; a < D; ++a)
      if (e[a] != g.L[a]) return false;
    return true;
  };
  auto keep_boundary = [&](double* e) {
    for (int a = 0; a < 3; ++a) e[a] = a < D ? g.L[a] : 0.0;
  };
  // an energy comparison the summation-order bound could flip is counted
  auto note = [&](double x, double thr, double bound) {
    if (std::isfinite(x) && std::fabs(x - thr) <= bound && x != thr) qc->ambiguous += 1;
  };
  const uint64_t bits = host_counter_u64(step, substep, seed);
  switch (qc->state) {
    case HMC_QC_STARTUP:
      qc->state = HMC_QC_WAITING;
      qc->is_final = 0;
      keep_boundary(qc->boundary);
      return HMC_OK;
    case HMC_QC_COMPLETE:
      if (!same_boundary(qc->boundary)) {
        qc->state = HMC_QC_WAITING;
        qc->is_final = 0;
        keep_boundary(qc->boundary);
      }
      return HMC_OK;
    case HMC_QC_WAITING: {
      double e = 0.0, bound = 0.0;
      const int rc = hmc_total_energy(c, &e, &bound);
      if (rc) return rc;
      note(e, 0.0, bound);
      if (e <= 0.0) {
        if (qc->is_final && same_boundary(qc->boundary)) {
          qc->state = HMC_QC_COMPLETE;
          keep_boundary(qc->boundary);
        } else {
          qc->state = HMC_QC_COMPRESSING;
        }
      }
      return HMC_OK;
    }
    case HMC_QC_COMPRESSING: {
      const double delta = host_uniform(bits, 0.0, qc->maximum_delta);
      const double current = box_volume(g);
      const double target = qc->target_volume;
      const double trial = target > current ? std::min(current * (1.0 + delta), target)
                                            : std::max(current * (1.0 - delta), target);
      // Periodic::scale_volume -> Hypercuboid::scale_volume (cuboid.rs:514-518)
      const double f = std::pow(trial / current, 1.0 / (double)D);
      double edges[3] = {0.0, 0.0, 0.0};
      for (int a = 0; a < D; ++a) edges[a] = g.L[a] * f;
      hmc_ctx* d = nullptr;
      int rc = hmc_clone_with_boundary(c, edges, &d);
      if (rc) return rc;
      // PairwiseCutoff::delta_energy_total (pairwise_cutoff.rs:793-840)
      double ef = 0.0, bf = 0.0, ei = 0.0, bi = 0.0;
      rc = hmc_total_energy(d, &ef, &bf);
      if (rc == HMC_OK && std::isfinite(ef) && g.model != HMC_HARD_SPHERE && g.model != HMC_CONVEX_POLYHEDRON)
        rc = hmc_total_energy(c, &ei, &bi);
      if (rc) {
        hmc_destroy(d);
        return fail(c, rc, "QuickCompress: total energy failed");
      }
      const double dE = std::isinf(ef) ? ef : (std::isinf(ei) ? -ei : ef - ei);
      const double thr = qc->maximum_energy_per_site * (double)g.n_bodies;
      note(dE, thr, bf + bi);
      if (dE > thr) {  // too strained: stay compressing
        hmc_destroy(d);
        return HMC_OK;
      }
      qc->state = HMC_QC_WAITING;
      qc->is_final = trial == target ? 1 : 0;
      for (int a = 0; a < 3; ++a) qc->boundary[a] = a < D ? edges[a] : 0.0;
      hmc_destroy(c);
      *pc = d;
      return HMC_OK;
    }
    default:
      return fail(c, HMC_ERR_INVALID, "QuickCompress: unknown state");
  }
}

// -------------------------------------------------------------- QuickInsert --
// Host restatements, in the device's arithmetic, for the candidates of one
// call against each other: Periodic::wrap (cuboid.rs:90-103), the ghost
// flags of a site (cuboid.rs:108-161), the squared distance to a site over
// its images (the per-axis nearest image), and the pair energy.
static double host_wrap(double x, double L) {
  double lambda = x / L;
  lambda = lambda - std::round(lambda);
  if (lambda == 0.5) lambda = -0.5;
  return lambda * L;
}
static uint32_t host_ghost_flags(const Geometry& g, const double* x) {
  uint32_t f = 0;
  for (int a = 0; a < g.dim; ++a) {
    if (x[a] < -g.half[a] || x[a] >= g.half[a]) return kNoImages;
    if (x[a] > g.ghost_hi[a])
      f |= 2u << (2 * a);
    else if (x[a] < g.ghost_lo[a])
      f |= 1u << (2 * a);
  }
  return f;
}
static double host_r2(const Geometry& g, const double* t, const double* xj, uint32_t fl) {
  double d2 = 0.0;
  for (int a = 0; a < g.dim; ++a) {
    double d = t[a] - xj[a];
    const uint32_t f = (fl >> (2 * a)) & 3u;
    if (f && !(fl & kNoImages)) {
      const double dg = t[a] - (xj[a] + (f == 1u ? g.L[a] : -g.L[a]));
      if (std::fabs(dg) < std::fabs(d)) d = dg;
    }
    d2 = d2 + d * d;
  }
  return d2;
}
// Pair energy of two sites at squared distance d2 (hard_shape.rs:143-153,
// isotropic.rs:147-156, lennard_jones.rs:107-111, shifted.rs:92-94,
// overlap_penalty.rs:109-115); *edge when d2 lies within 2^-30 of the
// cutoff, where the reference's VecCell may not visit the image.
static double host_pair(const Geometry& g, double d2, bool* edge) {
  if (g.model == HMC_HARD_SPHERE) {
    if (d2 >= g.sigma2) return 0.0;
    if (d2 >= g.edge2) *edge = true;
    return INFINITY;
  }
  if (d2 >= g.lj_rc2) return 0.0;
  if (d2 >= g.lj_rc2_edge) *edge = true;
  const double r = std::sqrt(d2);
  if (g.model == HMC_OVERLAP_PENALTY) {
    const double x = r - g.op_delta;
    if (x < -g.op_dmax) return INFINITY;
    return x < 0.0 ? g.op_half_k * (x * x) + g.op_eps : 0.0;
  }
  const double sr = g.lj_sigma / r;
  const double s2 = sr * sr, s4 = s2 * s2, s8 = s4 * s4;
  const double s12 = s4 * s8, s6 = s2 * s4;
  return g.lj_eps4 * (s12 - s6) - g.lj_shift;
}

void hmc_quick_insert_init(hmc_quick_insert* qi, uint64_t target) {
  if (!qi) return;
  std::memset(qi, 0, sizeof *qi);
  qi->target = target;
  qi->allowed_overlaps = std::max<uint64_t>(target / 8, 1);
  qi->state = HMC_QI_RUNNING;
}

// The context of a changed body list (bodies and sites in index order, D
// per row): a new context with the same configuration replaces *pc.
static int rebuild_with(hmc_ctx** pc, const std::vector<double>& body, const std::vector<double>& site, uint64_t n,
                        const char* who) {
  hmc_ctx* c = *pc;
  hmc_config cfg = c->cfg_user;
  cfg.n_bodies = n;
  hmc_ctx* d = nullptr;
  int rc = hmc_create(&cfg, &d);
  if (rc) return fail(c, rc, std::string(who) + ": creating the context of the new microstate failed");
  rc = hmc_set_state(d, body.data(), site.data());
  if (rc) {
    hmc_destroy(d);
    return fail(c, rc, std::string(who) + ": setting the state of the new microstate failed");
  }
  hmc_destroy(c);
  *pc = d;
  return HMC_OK;
}

// random_range(0..len) for len < 2^32 (rand 0.10 samples a usize range that
// fits in u32 as a u32): Canon's widening multiply with at most one extra
// draw, as choose_index on the device.
static uint32_t host_range_u32(HostSfc64& r, uint32_t len) {
  const uint64_t m = (uint64_t)(uint32_t)r.next() * (uint64_t)len;
  uint32_t result = (uint32_t)(m >> 32);
  const uint32_t lo = (uint32_t)m;
  if (lo > (uint32_t)(0u - len)) {
    const uint32_t hi2 = (uint32_t)(((uint64_t)(uint32_t)r.next() * (uint64_t)len) >> 32);
    if ((uint64_t)lo + (uint64_t)hi2 > 0xffffffffull) result += 1;
  }
  return result;
}

int hmc_quick_insert_apply(hmc_quick_insert* qi, hmc_ctx** pc, uint64_t step, uint32_t substep, uint32_t seed,
                           hmc_count* count) {
  if (!qi || !pc || !*pc || !count) return HMC_ERR_INVALID;
  hmc_ctx* c = *pc;
  count->accepted = count->rejected = 0;
  if (!c->have_state) return fail(c, HMC_ERR_STATE, "no state");
  const Geometry& g = c->g;
  const int D = g.dim;
  if (g.model == HMC_CONVEX_POLYHEDRON)
    return fail(c, HMC_ERR_UNSUPPORTED, "QuickInsert of oriented bodies is not offloaded");
  if (qi->state == HMC_QI_COMPLETE) return HMC_OK;
  double e = 0.0, bound = 0.0;
  int rc = hmc_total_energy(c, &e, &bound);
  if (rc) return rc;
  if (std::isfinite(e) && e != 0.0 && std::fabs(e) <= bound) qi->ambiguous += 1;
  if (e <= 0.0 && qi->inserted >= qi->target) {
    qi->state = HMC_QI_COMPLETE;
    return HMC_OK;
  }
  if (!(e <= 0.0) || qi->target == 0) return HMC_OK;
  const uint64_t T = qi->target;
  if ((uint64_t)g.n_bodies + T >= 0xffffffffull) return fail(c, HMC_ERR_INVALID, "QuickInsert: too many bodies");
  // every candidate the call can draw, in the stream's order
  // (UniformIn::sample -> Hypercuboid's Distribution: one Uniform draw per axis)
  std::vector<double> cand((size_t)T * D);
  {
    HostSfc64 rng = host_counter_rng(step, substep, seed);
    for (uint64_t k = 0; k < T; ++k)
      for (int a = 0; a < D; ++a) cand[(size_t)k * D + a] = host_uniform(rng.next(), -g.L[a] / 2.0, g.L[a] / 2.0);
  }
  // insertion energies against the bodies already present, all on the device
  std::vector<double> es(T), eb(T);
  rc = hmc_delta_energy_insert(c, cand.data(), T, es.data(), eb.data());
  if (rc) return rc;
  // the candidates in order, each against the ones inserted before it in this
  // call (a hash grid of cells no narrower than the interaction range)
  const double range = g.model == HMC_HARD_SPHERE ? g.sigma : std::sqrt(g.lj_rc2);
  int32_t gs[3] = {1, 1, 1};
  for (int a = 0; a < D; ++a) gs[a] = std::max<int32_t>(1, (int32_t)std::floor(g.L[a] / range));
  auto cell_of = [&](const double* x, int32_t* m) {
    for (int a = 0; a < D; ++a) {
      int32_t k = (int32_t)std::floor((x[a] + g.half[a]) / g.L[a] * gs[a]);
      m[a] = std::min(std::max(k, 0), gs[a] - 1);
    }
  };
  std::unordered_map<uint64_t, std::vector<uint32_t>> grid;
  auto key = [&](const int32_t* m) {  // one key per cell
    return (uint64_t)m[0] + (uint64_t)gs[0] * ((uint64_t)m[1] + (uint64_t)gs[1] * (uint64_t)m[2]);
  };
  struct New {
    double body[3], site[3];
    uint32_t flags;
  };
  std::vector<New> added;
  uint64_t overlapping = 0;
  for (uint64_t k = 0; k < T; ++k) {
    const double* p = &cand[(size_t)k * D];
    double t[3] = {0.0, 0.0, 0.0};  // the candidate's site, as delta_energy_insert forms it
    for (int a = 0; a < D; ++a) t[a] = host_wrap(p[a] + 0.0, g.L[a]);
    double de = es[k], dbound = eb[k], mag = 0.0;
    if (std::isfinite(de)) {
      int32_t m[3] = {0, 0, 0};
      cell_of(t, m);
      int32_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
      for (int a = 0; a < D; ++a) {
        lo[a] = gs[a] < 3 ? 0 : -1;
        hi[a] = gs[a] < 3 ? gs[a] - 1 : 1;
      }
      for (int32_t o0 = lo[0]; o0 <= hi[0] && std::isfinite(de); ++o0)
        for (int32_t o1 = lo[1]; o1 <= hi[1] && std::isfinite(de); ++o1)
          for (int32_t o2 = (D == 3 ? lo[2] : 0); o2 <= (D == 3 ? hi[2] : 0) && std::isfinite(de); ++o2) {
            int32_t n[3] = {0, 0, 0};
            const int32_t off[3] = {o0, o1, o2};
            for (int a = 0; a < D; ++a) {
              n[a] = gs[a] < 3 ? off[a] : ((m[a] + off[a]) % gs[a] + gs[a]) % gs[a];
            }
            const auto it = grid.find(key(n));
            if (it == grid.end()) continue;
            for (uint32_t j : it->second) {
              bool edge = false;
              const double one = host_pair(g, host_r2(g, t, added[j].site, added[j].flags), &edge);
              if (edge) qi->ambiguous += 1;
              if (std::isinf(one)) {
                de = INFINITY;
                break;
              }
              de += one;
              mag += std::fabs(one);
            }
          }
    }
    if (std::isfinite(de)) {
      // Microstate::add_body (microstate.rs:777-835): body = wrap(p),
      // site = wrap(body + 0)
      New nb{};
      for (int a = 0; a < D; ++a) {
        nb.body[a] = host_wrap(p[a], g.L[a]);
        nb.site[a] = host_wrap(nb.body[a] + 0.0, g.L[a]);
      }
      nb.flags = host_ghost_flags(g, nb.site);
      int32_t m[3] = {0, 0, 0};
      cell_of(nb.site, m);
      grid[key(m)].push_back((uint32_t)added.size());
      added.push_back(nb);
      count->accepted += 1;
      qi->inserted += 1;
      // the sign of dE within the bound of the summation order (the
      // reference sums in VecCell order)
      const double b2 = dbound + mag * 0x1p-50;
      if (de != 0.0 && std::fabs(de) <= b2) qi->ambiguous += 1;
      if (de > 0.0) overlapping += 1;
      if (qi->inserted == qi->target || overlapping >= qi->allowed_overlaps) break;
    } else {
      count->rejected += 1;
    }
  }
  if (added.empty()) return HMC_OK;
  // the new context: the old bodies, then the inserted ones in order
  const uint64_t N = g.n_bodies, A = added.size();
  std::vector<double> body((size_t)(N + A) * D), site((size_t)(N + A) * D);
  rc = hmc_get_state(c, body.data(), site.data());
  if (rc) return rc;
  for (uint64_t k = 0; k < A; ++k)
    for (int a = 0; a < D; ++a) {
      body[(size_t)(N + k) * D + a] = added[k].body[a];
      site[(size_t)(N + k) * D + a] = added[k].site[a];
    }
  return rebuild_with(pc, body, site, N + A, "QuickInsert");
}

// ------------------------------------------------------------ tune_local --
int hmc_tune_local(hmc_ctx* c, uint32_t move, uint64_t step, uint32_t substep, uint32_t seed, double kT,
                   double target, uint64_t samples, uint32_t steps, double* maximum_size, hmc_count* counts,
                   uint64_t* ambiguous) {
  if (!c || !maximum_size) return HMC_ERR_INVALID;
  if (!c->have_state) return fail(c, HMC_ERR_STATE, "no state");
  const Geometry& g = c->g;
  const int D = g.dim;
  const bool convex = g.model == HMC_CONVEX_POLYHEDRON;
  if (move != HMC_MOVE_TRANSLATE && move != HMC_MOVE_ROTATE) return fail(c, HMC_ERR_INVALID, "tune: unknown move");
  if (move == HMC_MOVE_ROTATE && !convex) return fail(c, HMC_ERR_INVALID, "tune: point bodies have no orientation");
  if (!(target > 0.0 && target < 1.0) || !(*maximum_size > 0.0) || !std::isfinite(*maximum_size) || !(kT > 0.0))
    return fail(c, HMC_ERR_INVALID, "tune: target in (0, 1), maximum size and kT > 0");
  double* const maximum_distance = maximum_size;
  if (g.n_bodies == 0) return fail(c, HMC_ERR_INVALID, "tune: no bodies to sample");
  if (samples > (1ull << 26)) return fail(c, HMC_ERR_INVALID, "tune: at most 2^26 samples per step");
  if (ambiguous) *ambiguous = 0;
  HIPCHK(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  launch_cell_lists(c);  // the state is not modified: one set of cells serves every step
  HostSfc64 rng = host_counter_rng(step, substep, seed);
  uint64_t M = samples * (D == 3 ? 7 : 6) + 4096, cap = 0;
  uint64_t* dx = nullptr;
  uint32_t* dn = nullptr;
  uint8_t* df = nullptr;
  std::vector<uint64_t> hx;
  std::vector<uint32_t> hn;
  std::vector<uint8_t> hf;
  auto release = [&]() {
    if (dx) (void)hipFreeAsync(dx, s);
    if (dn) (void)hipFreeAsync(dn, s);
    if (df) (void)hipFreeAsync(df, s);
    dx = nullptr;
    dn = nullptr;
    df = nullptr;
  };
  int rc = HMC_OK;
  uint64_t amb = 0;
  for (uint32_t t = 0; t < steps && !rc; ++t) {
    uint64_t acc = 0, rej = 0, a = 0, p = 0;
    for (;;) {
      if (M > 0x7fffffffull) {
        rc = fail(c, HMC_ERR_DEVICE, "tune: the sample chain outgrew its stream");
        break;
      }
      if (M > cap) {
        release();
        hipError_t e = hipMallocAsync((void**)&dx, M * sizeof(uint64_t), s);
        if (e == hipSuccess) e = hipMallocAsync((void**)&dn, M * sizeof(uint32_t), s);
        if (e == hipSuccess) e = hipMallocAsync((void**)&df, M, s);
        if (e != hipSuccess) {
          rc = fail(c, HMC_ERR_DEVICE, std::string("tune: ") + hipGetErrorString(e));
          break;
        }
        cap = M;
        hx.resize(M);
        hn.resize(M);
        hf.resize(M);
      }
      // the step's stream from the generator's current state
      HostSfc64 r = rng;
      for (uint64_t i = 0; i < M; ++i) hx[i] = r.next();
      hipError_t e = hipMemcpyAsync(dx, hx.data(), M * sizeof(uint64_t), hipMemcpyHostToDevice, s);
      if (e == hipSuccess) {
        if (convex)
          k_tune_eval_convex<<<grid_for(M, 256), 256, 0, s>>>(g, c->st, c->bodies, c->w, dx, (uint32_t)M, move,
                                                               *maximum_distance, dn, df);
        else if (D == 2)
          k_tune_eval<2><<<grid_for(M, kEnergyBlock), kEnergyBlock, 0, s>>>(g, c->st, c->bodies, c->w, dx,
                                                                             (uint32_t)M, *maximum_distance, kT,
                                                                             dn, df);
        else
          k_tune_eval<3><<<grid_for(M, kEnergyBlock), kEnergyBlock, 0, s>>>(g, c->st, c->bodies, c->w, dx,
                                                                             (uint32_t)M, *maximum_distance, kT,
                                                                             dn, df);
        e = hipGetLastError();
      }
      if (e == hipSuccess) e = hipMemcpyAsync(hn.data(), dn, M * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipMemcpyAsync(hf.data(), df, M, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e != hipSuccess) {
        rc = fail(c, HMC_ERR_DEVICE, std::string("tune: ") + hipGetErrorString(e));
        break;
      }
      // the samples of this step: the chain from offset 0
      acc = rej = a = p = 0;
      bool ok = true;
      for (uint64_t k = 0; k < samples; ++k) {
        const uint32_t q = p < M ? hn[p] : 0xFFFFFFFFu;
        if (q == 0xFFFFFFFFu) {
          ok = false;
          break;
        }
        if (hf[p] & 4u) {
          rc = fail(c, HMC_ERR_STATE, "tune: a site outside the checkerboard");
          break;
        }
        if (hf[p] & 1u)
          ++acc;
        else
          ++rej;
        a += (hf[p] >> 1) & 1u;
        p = q;
      }
      if (rc || ok) break;
      M *= 2;  // a longer stream for this step
    }
    if (rc) break;
    for (uint64_t i = 0; i < p; ++i) (void)rng.next();  // the generator after the step's samples
    amb += a;
    if (counts) {
      counts[t].accepted = acc;
      counts[t].rejected = rej;
    }
    // tune_by_scaling (hoomd-mc/src/lib.rs:514-524), Translate::adjust
    // (translate.rs:87-92), Rotate<Versor>::adjust (rotate/versor.rs:125-137:
    // at most pi/2)
    if (acc + rej > 0) {
      const double ratio = (double)acc / (double)(acc + rej);
      const double scale = (ratio + 1.5) / (target + 1.5);
      *maximum_distance *= scale;
      if (move == HMC_MOVE_ROTATE && *maximum_distance > 1.5707963267948966) *maximum_distance = 1.5707963267948966;
    }
  }
  release();
  if (ambiguous) *ambiguous = amb;
  return rc;
}

// --------------------------------------------------------- GrandCanonical --
int hmc_grand_canonical_apply(hmc_ctx** pc, uint64_t step, uint32_t substep, uint32_t seed, double kT,
                              double fugacity, hmc_gc_count* count, uint64_t* ambiguous) {
  if (!pc || !*pc || !count) return HMC_ERR_INVALID;
  hmc_ctx* c = *pc;
  std::memset(count, 0, sizeof *count);
  if (ambiguous) *ambiguous = 0;
  if (!c->have_state) return fail(c, HMC_ERR_STATE, "no state");
  const Geometry& g = c->g;
  const int D = g.dim;
  if (!is_isotropic(g.model) && g.model != HMC_HARD_SPHERE)
    return fail(c, HMC_ERR_UNSUPPORTED, "GrandCanonical of oriented bodies is not offloaded");
  if (!(kT > 0.0) || !std::isfinite(fugacity)) return fail(c, HMC_ERR_INVALID, "GrandCanonical: kT > 0, finite f");
  const uint64_t n = g.n_bodies;
  const double volume = box_volume(g);
  if (n == 0 && fugacity <= 0.0) return HMC_OK;  // no move can be accepted
  HostSfc64 rng = host_counter_rng(step, substep, seed);
  const bool insert = (int32_t)(uint32_t)rng.next() < 0;  // StandardUniform bool: the sign bit of a u32
  auto u01 = [&]() { return (double)(rng.next() >> 11) * 0x1p-53; };
  std::vector<double> body, site;
  auto state = [&](uint64_t cap) {
    body.assign((size_t)cap * D, 0.0);
    site.assign((size_t)cap * D, 0.0);
    return n ? hmc_get_state(c, body.data(), site.data()) : HMC_OK;
  };
  int rc = HMC_OK;
  if (insert) {
    // UniformIn over the box (uniform_in.rs:99-110): one Uniform draw per axis
    double p[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < D; ++a) p[a] = host_uniform(rng.next(), -g.L[a] / 2.0, g.L[a] / 2.0);
    double de = 0.0, db = 0.0;
    rc = hmc_delta_energy_insert(c, p, 1, &de, &db);
    if (rc) return rc;
    if (!std::isfinite(de)) {
      count->insert_rejected = 1;
      return HMC_OK;
    }
    const double p_insert = (volume * fugacity * std::exp(-de / kT)) / ((double)n + 1.0);
    const double u = u01();
    if (ambiguous && db > 0.0) {
      const double plo = (volume * fugacity * std::exp(-(de + db) / kT)) / ((double)n + 1.0);
      const double phi = (volume * fugacity * std::exp(-(de - db) / kT)) / ((double)n + 1.0);
      if ((plo > u) != (phi > u)) *ambiguous = 1;
    }
    if (!(p_insert > u)) {
      count->insert_rejected = 1;
      return HMC_OK;
    }
    // Microstate::add_body (microstate.rs:777-835): body = wrap(p), site = wrap(body + 0)
    rc = state(n + 1);
    if (rc) return rc;
    for (int a = 0; a < D; ++a) {
      body[(size_t)n * D + a] = host_wrap(p[a], g.L[a]);
      site[(size_t)n * D + a] = host_wrap(body[(size_t)n * D + a] + 0.0, g.L[a]);
    }
    rc = rebuild_with(pc, body, site, n + 1, "GrandCanonical");
    if (!rc) count->insert_accepted = 1;
    return rc;
  }
  if (n == 0) return HMC_OK;
  const uint32_t index = host_range_u32(rng, (uint32_t)n);
  double out[2] = {0.0, 0.0};
  {
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    double* d = nullptr;
    hipError_t e = hipMallocAsync((void**)&d, 2 * sizeof(double), s);
    if (e == hipSuccess) {
      launch_cell_lists(c);
      if (D == 2)
        k_remove_energy<2><<<1, 64, 0, s>>>(g, c->st, c->bodies, c->w, index, d);
      else
        k_remove_energy<3><<<1, 64, 0, s>>>(g, c->st, c->bodies, c->w, index, d);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipMemcpyAsync(out, d, sizeof out, hipMemcpyDeviceToHost, s);
      (void)hipFreeAsync(d, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    if (e != hipSuccess) return fail(c, HMC_ERR_DEVICE, std::string("GrandCanonical: ") + hipGetErrorString(e));
    if (out[1] < 0.0) return fail(c, HMC_ERR_STATE, "GrandCanonical: a site outside the checkerboard");
  }
  const double de = out[0], db = out[1];
  const double p_remove = fugacity <= 0.0 ? 1.0 : ((double)n * std::exp(-de / kT)) / (volume * fugacity);
  const double u = u01();
  if (ambiguous && db > 0.0 && fugacity > 0.0) {
    const double plo = ((double)n * std::exp(-(de + db) / kT)) / (volume * fugacity);
    const double phi = ((double)n * std::exp(-(de - db) / kT)) / (volume * fugacity);
    if ((plo > u) != (phi > u)) *ambiguous = 1;
  }
  if (!(p_remove > u)) {
    count->remove_rejected = 1;
    return HMC_OK;
  }
  // Microstate::remove_body (microstate.rs:917-943): the last body takes the
  // removed one's index (swap_remove)
  rc = state(n);
  if (rc) return rc;
  for (int a = 0; a < D; ++a) {
    body[(size_t)index * D + a] = body[(size_t)(n - 1) * D + a];
    site[(size_t)index * D + a] = site[(size_t)(n - 1) * D + a];
  }
  body.resize((size_t)(n - 1) * D);
  site.resize((size_t)(n - 1) * D);
  rc = rebuild_with(pc, body, site, n - 1, "GrandCanonical");
  if (!rc) count->remove_accepted = 1;
  return rc;
}

}  // extern "C"
