// QuickCompress (hoomd-mc/src/quick_compress.rs) over device energies and
// boundary clones.
#include "host.hpp"

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

}  // extern "C"
