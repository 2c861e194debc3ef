// The C ABI of one context (include/hmc.h): creation, state, applies and
// runs, traces, statistics and profiling.
#include "host.hpp"

namespace hmc {

hipEvent_t take_event(hmc_ctx* c) {
  if (!c->ev_pool.empty()) {
    hipEvent_t e = c->ev_pool.back();
    c->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

template <int D, int PREC>
int launch_apply(hmc_ctx* c, int from_args, int advance, uint64_t step, uint32_t substep, uint32_t seed, double kT,
                 double max_move, uint32_t move) {
  const Geometry& g = c->g;
  hipStream_t s = c->stream;
  const uint32_t N = g.n_bodies;
  const bool prof = c->profiling && !c->capturing && c->applies % c->profile_period == 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (prof) {
    e0 = take_event(c);
    (void)hipEventRecord(e0, s);
  }
  k_setup<<<1, 64, 0, s>>>(g, c->st, step, substep, seed, kT, max_move, from_args, advance, move, 1);
  // (the per-space counts are zero: cleared at creation and by k_scan_add)
  if (N) k_bin<D><<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->st, c->bodies, c->w);
  if (N < (1u << kLookbackBits) && g.n_spaces < (1u << kLookbackBits)) {
    k_scan_lookback<<<c->n_scan_blocks, 256, 0, s>>>(g, c->st, c->w.cnt, c->w.start, c->w.scan_status, g.n_spaces,
                                                     c->n_scan_blocks);
  } else {
    k_scan_blocks<<<c->n_scan_blocks, 256, 0, s>>>(c->w.cnt, c->w.start, c->w.block_sums, c->w.block_ne,
                                                   g.n_spaces);
    k_scan_top<<<1, 1024, 0, s>>>(g, c->st, c->w.block_sums, c->w.block_ne, c->n_scan_blocks, c->w.start, 1);
    k_scan_add<<<grid_for(g.n_spaces, 256), 256, 0, s>>>(c->w.start, c->w.block_sums, c->w.cnt, g.n_spaces);
  }
  if (N) {
    k_scatter<<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->w);
    // the persistent apply stages its sites from the FP64 copies
    if (c->persist)
      k_place<PREC, false><<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->bodies, c->cells, c->w);
    else
      k_place<PREC><<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->bodies, c->cells, c->w);
  }
  if (prof) {
    e1 = take_event(c);
    (void)hipEventRecord(e1, s);
    c->build_events.push_back({e0, e1});
  }
  if (c->persist) {
    // one persistent launch runs the apply: hard shapes (mixed precision) or
    // isotropic pair energies (FP32 energies with the FP64 layout)
    hipEvent_t p0 = nullptr, p1 = nullptr;
    if (prof) {
      p0 = take_event(c);
      (void)hipEventRecord(p0, s);
    }
    ApplyState* stp = c->st;
    void* args[] = {(void*)&g, (void*)&stp, (void*)&c->bodies, (void*)&c->cells, (void*)&c->w, (void*)&c->pc};
    const void* fn = is_isotropic(g.model) ? (const void*)k_apply_lj<D> : (const void*)k_apply_tile<D>;
    const hipError_t le =
        hipLaunchCooperativeKernel(fn, dim3(c->pgrid), dim3(c->pthreads), args, c->pc.lds_bytes, s);
    if (le != hipSuccess) return fail(c, HMC_ERR_DEVICE, std::string("cooperative launch: ") + hipGetErrorString(le));
    if (prof) {
      p1 = take_event(c);
      (void)hipEventRecord(p1, s);
      c->pass_events.push_back({p0, p1});
    }
    c->applies += 1;
    return HMC_OK;
  }
  const uint32_t blocks = grid_for(g.per_colour, kBlock);
  const uint32_t chunk = (blocks + 7) / 8;
  const uint32_t grid = chunk * 8;
  const uint32_t tile_threads = ((c->tc.trials + 63) / 64) * 64;
  const bool lj = is_isotropic(g.model);
  const bool convex = g.model == HMC_CONVEX_POLYHEDRON;
  // LJ needs many cycles per apply and its guard is slow: launch one spare
  // cycle (its passes exit at once when not needed)
  const uint32_t launched = c->cycles_guess + (lj ? 1u : 0u);
  const uint32_t lj_blocks = grid_for(g.per_colour, kBlock / 64);
  const uint32_t lj_chunk = (lj_blocks + 7) / 8;
  const uint32_t cv_chunk = (grid_for(g.per_colour, (kBlock / 64) * kCvTrials) + 7) / 8;
  for (uint32_t k = 0; k < launched; ++k) {
    for (int col = 0; col < (1 << D); ++col) {
      hipEvent_t p0 = nullptr, p1 = nullptr;
      if (prof) {
        p0 = take_event(c);
        (void)hipEventRecord(p0, s);
      }
      if (lj)
      {
        if (g.lj_mixed)
          k_pass_lj<D, HMC_MIXED><<<lj_chunk * 8, kBlock, 0, s>>>(g, c->st, c->bodies, c->cells, c->w, (uint32_t)col,
                                                                  k, lj_chunk);
        else
          k_pass_lj<D, HMC_FP64><<<lj_chunk * 8, kBlock, 0, s>>>(g, c->st, c->bodies, c->cells, c->w, (uint32_t)col,
                                                                 k, lj_chunk);
      }
      else if (convex && g.cv_mixed)
        k_pass_convex<true><<<cv_chunk * 8, kBlock, 0, s>>>(g, c->st, c->bodies, c->cells, c->w, (uint32_t)col, k,
                                                            cv_chunk, c->sat);
      else if (convex)
        k_pass_convex<false><<<cv_chunk * 8, kBlock, 0, s>>>(g, c->st, c->bodies, c->cells, c->w, (uint32_t)col, k,
                                                             cv_chunk, c->sat);
      else if (c->tiled && PREC == HMC_MIXED && g.lj_fp32_error)  // the FP32 error report
        k_pass_tile<D, PREC, true><<<c->tc.xcd_chunk * 8, tile_threads, c->tc.lds_bytes, s>>>(
            g, c->st, c->bodies, c->cells, c->w, (uint32_t)col, k, c->tc);
      else if (c->tiled)
        k_pass_tile<D, PREC><<<c->tc.xcd_chunk * 8, tile_threads, c->tc.lds_bytes, s>>>(
            g, c->st, c->bodies, c->cells, c->w, (uint32_t)col, k, c->tc);
      else if (c->fast)
        k_pass<D, PREC, true><<<grid, kBlock, 0, s>>>(g, c->st, c->bodies, c->cells, c->w, (uint32_t)col, k, chunk);
      else
        k_pass<D, PREC, false><<<grid, kBlock, 0, s>>>(g, c->st, c->bodies, c->cells, c->w, (uint32_t)col, k, chunk);
      if (prof) {
        p1 = take_event(c);
        (void)hipEventRecord(p1, s);
        c->pass_events.push_back({p0, p1});
      }
    }
  }
  if (lj)
    k_guard_lj<D><<<1, kGuardThreads, 0, s>>>(g, c->st, c->bodies, c->cells, c->w, launched);
  else if (convex)
    k_guard_convex<<<1, kGuardThreads, 0, s>>>(g, c->st, c->bodies, c->cells, c->w, launched);
  else
    k_guard<D, PREC><<<1, kGuardThreads, 0, s>>>(g, c->st, c->bodies, c->cells, c->w, launched);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(c, HMC_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
  c->applies += 1;
  return HMC_OK;
}

int dispatch_apply(hmc_ctx* c, int from_args, int advance, uint64_t step, uint32_t substep, uint32_t seed,
                   double kT, double max_move, uint32_t move) {
  int d = c->g.dim, p = c->cfg.precision;
  if (d == 2 && p == HMC_FP64)
    return launch_apply<2, HMC_FP64>(c, from_args, advance, step, substep, seed, kT, max_move, move);
  if (d == 2) return launch_apply<2, HMC_MIXED>(c, from_args, advance, step, substep, seed, kT, max_move, move);
  if (p == HMC_FP64) return launch_apply<3, HMC_FP64>(c, from_args, advance, step, substep, seed, kT, max_move, move);
  return launch_apply<3, HMC_MIXED>(c, from_args, advance, step, substep, seed, kT, max_move, move);
}

// Forget the captured step graph (its kernel arguments name buffers that are
// about to change).
void drop_graph(hmc_ctx* c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
}

void collect_events(hmc_ctx* c, std::vector<std::pair<hipEvent_t, hipEvent_t>>& evs, double& ms_total,
                    uint64_t& n) {
  for (auto& pr : evs) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
      ms_total += ms;
      n += 1;
    }
    c->ev_pool.push_back(pr.first);
    c->ev_pool.push_back(pr.second);
  }
  evs.clear();
}

int read_state(hmc_ctx* c, ApplyState* out, bool check_errors) {
  std::vector<unsigned long long> slots(kAccSlots);
  HIPCHK(c, hipMemcpyAsync(out, c->st, sizeof(ApplyState), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(slots.data(), c->w.acc_slots, kAccSlots * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  out->accepted_total = 0;
  for (unsigned long long v : slots) out->accepted_total += v;
  if (c->profiling) {
    collect_events(c, c->pass_events, c->pass_ms, c->pass_n);
    collect_events(c, c->build_events, c->build_ms, c->build_n);
  }
  if (!check_errors) return HMC_OK;
  if (out->error & kErrBarrier) return fail(c, HMC_ERR_DEVICE, "a device-side wait (persistent apply neighbour sync or look-back scan) timed out");
  if (out->error) return fail(c, HMC_ERR_COVER, "a body lies outside the checkerboard");
  return HMC_OK;
}

// FP32 thresholds around contact for a per-axis error bound E on the FP32
// displacement: d2 < lo overlaps (and lies below the edge region), d2 >= hi
// does not.  The bound on d^2 near sigma^2 is 2 E sqrt(D) d + D E^2 +
// 4 (2^-24 + 2^-53) d^2 at d = 1.01 sigma (E must be far below sigma), with
// 10% margin; the float thresholds are rounded outward.
bool contact_band(double E, double sigma, double sigma2, double edge2, int D, float* lo, float* hi) {
  const double dref = 1.01 * sigma;
  const double Bd = (2.0 * E * std::sqrt((double)D) * dref + D * E * E + 4.0 * (0x1p-24 + 0x1p-53) * dref * dref) * 1.1;
  if (!(Bd < 0.01 * sigma2)) return false;
  float l = (float)(edge2 - Bd);
  if ((double)l > edge2 - Bd) l = std::nextafter(l, 0.0f);
  float h = (float)(sigma2 + Bd);
  if ((double)h < sigma2 + Bd) h = std::nextafter(h, 1e30f);
  *lo = l;
  *hi = h;
  return true;
}

// Tile shape of the LDS-staged pass for a system with `occ` bodies per space
// on a device with `cus` CUs.  Candidates: at most kTileThreads trials, at
// most 64 kRowsPerLane rows and kValsPerThread row starts per thread, at most
// kTileLds bytes of dynamic LDS with a body capacity of 1.2x the box's
// expected count plus 64 (a tile whose box holds more runs its trials from
// global memory).  A
// workgroup's time is set by a fixed number of dependent memory phases, so
// the shape that needs the fewest rounds of resident workgroups wins; among
// those, the least launched lanes and staging.
bool choose_tiles(const Geometry& g, int D, double occ, size_t vt, int cus, TileCfg* out) {
  // the box (2T + 1 spaces) must span at most shape - 2 spaces per axis, so
  // that box-relative positions are minimum-image positions
  uint32_t lim[3] = {1, 1, 1};
  const uint32_t cap[3] = {16, 16, 32};
  for (int a = 0; a < D; ++a) {
    if (g.shape[a] < 5) return false;
    lim[a] = std::min<uint32_t>(std::min<uint32_t>(g.half_shape[a], cap[a]), (uint32_t)(g.shape[a] - 3) / 2);
  }
  const size_t static_lds = 2 * 257 * sizeof(double) + 64;
  const size_t cu_lds = 160 * 1024;
  const uint32_t cu_waves = 8;  // two waves per SIMD at the kernel's register use
  double best_rounds = 0.0, best_work = 0.0;
  bool found = false;
  for (uint32_t t0 = 1; t0 <= lim[0]; ++t0)
    for (uint32_t t1 = 1; t1 <= lim[1]; ++t1)
      for (uint32_t t2 = 1; t2 <= lim[2]; ++t2) {
        const uint32_t T[3] = {t0, t1, t2};
        const uint32_t trials = T[0] * T[1] * T[2];
        if (trials > (uint32_t)kTileThreads) continue;
        const uint32_t threads = ((trials + 63) / 64) * 64;
        uint32_t ncell = 1, ntiles = 1;
        for (int a = 0; a < D; ++a) {
          ncell *= 2 * T[a] + 1;
          ntiles *= (g.half_shape[a] + T[a] - 1) / T[a];
        }
        const uint32_t EL = 2 * T[D - 1] + 1, rows = ncell / EL;
        if (rows > 64u * kRowsPerLane || rows * (EL + 2) > threads * kValsPerThread) continue;
        const uint32_t maxb = (uint32_t)std::ceil(1.2 * occ * ncell) + 64;
        const size_t lds = ((16 * ((size_t)rows + 1) + 4 * (size_t)rows * (EL + 1) + 15) & ~(size_t)15) +
                           (size_t)maxb * vt;
        if (lds > kTileLds) continue;
        const uint32_t per_cu = std::max<uint32_t>(
            1, std::min<uint32_t>((uint32_t)(cu_lds / (lds + static_lds)), cu_waves / (threads / 64)));
        const double rounds = std::ceil((double)ntiles / ((double)per_cu * cus));
        const double work = (double)ntiles * threads * (1.0 + 0.02 * (double)ncell / trials);
        if (!found || rounds < best_rounds || (rounds == best_rounds && work < best_work)) {
          found = true;
          best_rounds = rounds;
          best_work = work;
          for (int a = 0; a < 3; ++a) {
            out->T[a] = T[a];
            out->nt[a] = a < D ? (g.half_shape[a] + T[a] - 1) / T[a] : 1;
            out->E[a] = a < D ? 2 * T[a] + 1 : 1;
          }
          out->trials = trials;
          out->ncell = ncell;
          out->rows = rows;
          out->maxb = maxb;
          out->ntiles = ntiles;
          out->xcd_chunk = (ntiles + 7) / 8;
          out->lds_bytes = (uint32_t)lds;
          uint32_t r = threads;
          for (int a = D - 1; a > 0; --a) {
            const uint32_t rad = a == D - 1 ? EL + 2 : out->E[a];
            out->step[a] = r % rad;
            r /= rad;
          }
          out->step[0] = r;  // the leading digit is not reduced
          for (int a = D; a < 3; ++a) out->step[a] = 0;
        }
      }
  return found;
}

// Region shape of the persistent apply (k_apply_tile).  A region is 2 T
// spaces per axis; its box (the region and a one-space halo, 2 T + 2 spaces)
// must span at most shape - 2 spaces.  Every region must be resident at once,
// so a shape qualifies only if all its workgroups fit on the device (LDS: the
// box's offsets and a site capacity of 1.1x its expected count plus 64).
// Among those the fewest lanes per workgroup win -- a colour pass lasts as
// long as one workgroup's trials -- then the most trials on those lanes, then
// the least staging.
bool choose_regions(const Geometry& g, int D, double occ, int cus, TileCfg* out) {
  uint32_t lim[3] = {1, 1, 1};
  const uint32_t cap[3] = {16, 16, 32};
  for (int a = 0; a < D; ++a) {
    if (g.shape[a] < 6) return false;
    lim[a] = std::min<uint32_t>(std::min<uint32_t>(g.half_shape[a], cap[a]), (uint32_t)(g.shape[a] - 4) / 2);
  }
  const size_t static_lds = 2 * 257 * sizeof(double) + 64;
  const size_t cu_lds = 160 * 1024;
  const uint32_t cu_waves = 8;  // two waves per SIMD at the kernel's register use
  uint32_t best_threads = 0, best_trials = 0;
  double best_work = 0.0;
  bool found = false;
  for (uint32_t t0 = 1; t0 <= lim[0]; ++t0)
    for (uint32_t t1 = 1; t1 <= lim[1]; ++t1)
      for (uint32_t t2 = 1; t2 <= lim[2]; ++t2) {
        const uint32_t T[3] = {t0, t1, t2};
        const uint32_t trials = T[0] * T[1] * T[2];
        if (trials > (uint32_t)kTileThreads) continue;
        const uint32_t threads = ((trials + 63) / 64) * 64;
        uint32_t ncell = 1, ntiles = 1;
        for (int a = 0; a < D; ++a) {
          ncell *= 2 * T[a] + 2;
          ntiles *= (g.half_shape[a] + T[a] - 1) / T[a];
        }
        const uint32_t EL = 2 * T[D - 1] + 2, rows = ncell / EL;
        if (rows > 64u * kRowsPerLane || rows * (EL + 2) > threads * kValsPerThread) continue;
        const uint32_t maxb = (uint32_t)std::ceil(1.1 * occ * ncell) + 64;
        const size_t lds = ((16 * ((size_t)rows + 1) + 2 * (size_t)rows * (EL + 1) + 15) & ~(size_t)15) +
                           (size_t)maxb * sizeof(float4);
        if (lds > kTileLds) continue;
        const uint32_t per_cu = std::min<uint32_t>((uint32_t)(cu_lds / (lds + static_lds)), cu_waves / (threads / 64));
        const uint32_t grid = ((ntiles + 7) / 8) * 8;
        if (per_cu == 0 || grid > per_cu * (uint32_t)cus) continue;
        const double work = (double)ntiles * ncell;
        if (!found || threads < best_threads || (threads == best_threads && trials > best_trials) ||
            (threads == best_threads && trials == best_trials && work < best_work)) {
          found = true;
          best_threads = threads;
          best_trials = trials;
          best_work = work;
          for (int a = 0; a < 3; ++a) {
            out->T[a] = a < D ? T[a] : 1;
            out->nt[a] = a < D ? (g.half_shape[a] + T[a] - 1) / T[a] : 1;
            out->E[a] = a < D ? 2 * T[a] + 2 : 1;
          }
          out->trials = trials;
          out->ncell = ncell;
          out->rows = rows;
          out->maxb = maxb;
          out->ntiles = ntiles;
          out->xcd_chunk = (ntiles + 7) / 8;
          out->lds_bytes = (uint32_t)lds;
          uint32_t r = threads;
          for (int a = D - 1; a > 0; --a) {
            const uint32_t rad = a == D - 1 ? EL + 2 : out->E[a];
            out->step[a] = r % rad;
            r /= rad;
          }
          out->step[0] = r;
          for (int a = D; a < 3; ++a) out->step[a] = 0;
        }
      }
  return found;
}

// Whether every region of the context's persistent apply is resident at
// once, checked on the device: the persistent kernel itself in probe mode
// (region_probe), so the check sees the kernel's real register, LDS and
// private-memory use, which the occupancy API may not all account for.  The
// mailboxes and the state's error word are cleared afterwards.
bool persist_probe(hmc_ctx* c) {
  TileCfg pc = c->pc;
  pc.probe = 1;
  Geometry g = c->g;
  ApplyState* stp = c->st;
  void* args[] = {(void*)&g, (void*)&stp, (void*)&c->bodies, (void*)&c->cells, (void*)&c->w, (void*)&pc};
  const int D = g.dim;
  const void* fn = is_isotropic(g.model) ? (D == 2 ? (const void*)k_apply_lj<2> : (const void*)k_apply_lj<3>)
                                         : (D == 2 ? (const void*)k_apply_tile<2> : (const void*)k_apply_tile<3>);
  const size_t fbytes = (size_t)c->pc.ntiles * kMailboxStride * sizeof(uint32_t);
  uint32_t err = 0;
  bool ok = hipLaunchCooperativeKernel(fn, dim3(c->pgrid), dim3(c->pthreads), args, c->pc.lds_bytes, c->stream) ==
                hipSuccess &&
            hipMemcpyAsync(&err, &c->st->error, sizeof err, hipMemcpyDeviceToHost, c->stream) == hipSuccess &&
            hipStreamSynchronize(c->stream) == hipSuccess;
  ok = ok && err == 0;
  (void)hipMemsetAsync(c->w.mbox, 0, fbytes, c->stream);
  (void)hipMemsetAsync(&c->st->error, 0, sizeof(uint32_t), c->stream);
  (void)hipStreamSynchronize(c->stream);
  return ok;
}

// Workgroup size of k_apply_lj for a region of `trials` trials: one
// half-wavefront per trial up to eight, in 1, 2 or 4 wavefronts (a count
// that divides the four SIMDs of a CU, so that the wavefronts of several
// workgroups on one CU spread evenly; with five, two workgroups did not
// always fit on a CU that the occupancy count said holds them).
static uint32_t lj_region_threads(uint32_t trials) { return trials <= 2 ? 64u : (trials <= 4 ? 128u : 256u); }

// Region shape of the persistent isotropic apply (k_apply_lj): as
// choose_regions, with at most 12 trials (half-wavefronts taking one or two
// each), at most 128 cells of a colour's halo, and a site capacity of 1.1x
// the box's expected count plus 128.  Among the shapes whose workgroups all
// fit on the device at once (two waves per SIMD, the kernel's register
// budget; the host checks the real occupancy): the least staging over all
// regions, then the fewest lanes.
bool choose_regions_lj(const Geometry& g, int D, double occ, int cus, TileCfg* out) {
  uint32_t lim[3] = {1, 1, 1};
  for (int a = 0; a < D; ++a) {
    if (g.shape[a] < 6 || g.reach[a] != 1) return false;
    lim[a] = std::min<uint32_t>(std::min<uint32_t>(g.half_shape[a], 8u), (uint32_t)(g.shape[a] - 4) / 2);
  }
  // ziggurat tables, halo lists, 12 proposals of 160 bytes, the shared
  // trials' partial sums (2 x 8 x 11 floats)
  const size_t static_lds = 2 * 257 * sizeof(double) + (2 * 128 + 1) * sizeof(uint32_t) + 12 * 160 + 2 * 8 * 11 * 4 + 64;
  const size_t cu_lds = 160 * 1024;
  const uint32_t cu_waves = 8;
  bool found = false;
  double best_work = 0.0;
  uint32_t best_threads = 0;
  for (uint32_t t0 = 1; t0 <= lim[0]; ++t0)
    for (uint32_t t1 = 1; t1 <= lim[1]; ++t1)
      for (uint32_t t2 = 1; t2 <= lim[2]; ++t2) {
        const uint32_t T[3] = {t0, t1, t2};
        const uint32_t trials = T[0] * T[1] * T[2];
        if (trials > 12) continue;
        const uint32_t threads = lj_region_threads(trials);
        uint32_t ncell = 1, ntiles = 1, halo = 1;
        for (int a = 0; a < D; ++a) {
          ncell *= 2 * T[a] + 2;
          ntiles *= (g.half_shape[a] + T[a] - 1) / T[a];
          halo *= T[a] + 1;
        }
        if (halo > 128) continue;
        const uint32_t EL = 2 * T[D - 1] + 2, rows = ncell / EL;
        if (rows > 64u * kRowsPerLane || rows * (EL + 2) > threads * kValsPerThread) continue;
        const uint32_t maxb = (uint32_t)std::ceil(1.1 * occ * ncell) + 128;  // >= 96: the overflow path's tables
        const size_t lds = ((16 * ((size_t)rows + 1) + 2 * (size_t)rows * (EL + 1) + 15) & ~(size_t)15) +
                           (size_t)maxb * sizeof(float4);
        if (lds + static_lds > cu_lds) continue;
        const uint32_t per_cu = std::min<uint32_t>((uint32_t)(cu_lds / (lds + static_lds)), cu_waves / (threads / 64));
        const uint32_t grid = ((ntiles + 7) / 8) * 8;
        if (per_cu == 0 || grid > per_cu * (uint32_t)cus) continue;
        const double work = (double)ntiles * ncell;
        if (!found || work < best_work || (work == best_work && threads < best_threads)) {
          found = true;
          best_work = work;
          best_threads = threads;
          for (int a = 0; a < 3; ++a) {
            out->T[a] = a < D ? T[a] : 1;
            out->nt[a] = a < D ? (g.half_shape[a] + T[a] - 1) / T[a] : 1;
            out->E[a] = a < D ? 2 * T[a] + 2 : 1;
          }
          out->trials = trials;
          out->ncell = ncell;
          out->rows = rows;
          out->maxb = maxb;
          out->ntiles = ntiles;
          out->xcd_chunk = (ntiles + 7) / 8;
          out->lds_bytes = (uint32_t)lds;
          uint32_t r = threads;
          for (int a = D - 1; a > 0; --a) {
            const uint32_t rad = a == D - 1 ? EL + 2 : out->E[a];
            out->step[a] = r % rad;
            r /= rad;
          }
          out->step[0] = r;
          for (int a = D; a < 3; ++a) out->step[a] = 0;
        }
      }
  return found;
}

// The FP32 separating-axis data of a polytope for the mixed-precision convex
// pass: the face normals and edge directions of its convex hull, found by
// brute force over vertex triples (at most 32 vertices) and kept once per
// +-pair, the slab [hmin, hmax] of every face axis, and the scanned vertices
// (one of each antipodal pair when the vertex set is closed under negation).
// A flat or degenerate vertex set, or one with more face axes than
// kSatFaces, gets use = 0: all its pairs go to XenoCollide in FP64.
static void sat_build(const double* vt, uint32_t nv, SatShape* S) {
  std::memset(S, 0, sizeof *S);
  using d3 = std::array<double, 3>;
  auto V = [&](uint32_t i) { return d3{vt[3 * i], vt[3 * i + 1], vt[3 * i + 2]}; };
  auto sub = [](const d3& a, const d3& b) { return d3{a[0] - b[0], a[1] - b[1], a[2] - b[2]}; };
  auto dot = [](const d3& a, const d3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
  auto cross = [](const d3& a, const d3& b) {
    return d3{a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
  };
  double R2 = 0.0;
  for (uint32_t i = 0; i < nv; ++i) R2 = std::max(R2, dot(V(i), V(i)));
  const double R = std::sqrt(R2);
  if (!(R > 0.0)) return;
  const double eps = 1e-12 * R;
  std::vector<d3> pn;       // oriented face planes: outward unit normal
  std::vector<double> ph;   // and offset
  for (uint32_t i = 0; i < nv; ++i)
    for (uint32_t j = i + 1; j < nv; ++j)
      for (uint32_t k = j + 1; k < nv; ++k) {
        d3 n = cross(sub(V(j), V(i)), sub(V(k), V(i)));
        const double nn = std::sqrt(dot(n, n));
        if (!(nn > 1e-12 * R2)) continue;  // (nearly) collinear
        for (double& x : n) x /= nn;
        double h = dot(n, V(i));
        bool pos = false, neg = false;
        for (uint32_t m = 0; m < nv; ++m) {
          const double d = dot(n, V(m)) - h;
          pos |= d > eps;
          neg |= d < -eps;
        }
        if (pos && neg) continue;
        if (pos) {
          for (double& x : n) x = -x;
          h = -h;
        }
        bool seen = false;
        for (size_t q = 0; q < pn.size() && !seen; ++q) seen = dot(pn[q], n) > 1.0 - 1e-12 && std::fabs(ph[q] - h) <= eps;
        if (!seen) {
          pn.push_back(n);
          ph.push_back(h);
        }
      }
  if (pn.size() < 4) return;  // flat or degenerate
  auto unique_axis = [&](std::vector<d3>& set, const d3& a) {
    for (const d3& b : set)
      if (std::fabs(dot(a, b)) > 1.0 - 1e-12) return;
    set.push_back(a);
  };
  std::vector<d3> faces, edges;
  for (const d3& n : pn) unique_axis(faces, n);
  for (size_t q = 0; q < pn.size(); ++q) {
    std::vector<uint32_t> on;
    for (uint32_t m = 0; m < nv; ++m)
      if (std::fabs(dot(pn[q], V(m)) - ph[q]) <= eps) on.push_back(m);
    for (size_t a = 0; a < on.size(); ++a)
      for (size_t b = a + 1; b < on.size(); ++b) {
        d3 e = sub(V(on[b]), V(on[a]));
        const double el = std::sqrt(dot(e, e));
        if (!(el > eps)) continue;
        const d3 m = cross(pn[q], e);
        bool pos = false, neg = false;
        for (size_t r = 0; r < on.size(); ++r) {
          const double d = dot(m, sub(V(on[r]), V(on[a])));
          pos |= d > eps * el;
          neg |= d < -eps * el;
        }
        if (pos && neg) continue;  // a diagonal of the face
        for (double& x : e) x /= el;
        unique_axis(edges, e);
      }
  }
  S->nf = (uint32_t)faces.size();
  S->ne = (uint32_t)edges.size();
  if (S->nf > kSatFaces) return;
  for (uint32_t f = 0; f < S->nf; ++f) {
    double lo = INFINITY, hi = -INFINITY;
    for (uint32_t m = 0; m < nv; ++m) {
      lo = std::min(lo, dot(faces[f], V(m)));
      hi = std::max(hi, dot(faces[f], V(m)));
    }
    for (int a = 0; a < 3; ++a) S->face[f][a] = (float)faces[f][a];
    S->face[f][3] = (float)hi;
    S->fmin[f] = (float)lo;
  }
  if (S->ne <= kSatEdges)
    for (uint32_t e = 0; e < S->ne; ++e)
      for (int a = 0; a < 3; ++a) S->edge[e][a] = (float)edges[e][a];
  // the scanned vertices
  bool sym = true;
  for (uint32_t i = 0; i < nv && sym; ++i) {
    bool found = false;
    for (uint32_t j = 0; j < nv && !found; ++j)
      found = V(j)[0] == -V(i)[0] && V(j)[1] == -V(i)[1] && V(j)[2] == -V(i)[2];
    sym = found;
  }
  std::vector<bool> covered(nv, false);
  uint32_t nh = 0;
  for (uint32_t i = 0; i < nv; ++i) {
    if (sym && covered[i]) continue;
    for (int a = 0; a < 3; ++a) S->hv[nh][a] = (float)vt[3 * i + a];
    ++nh;
    if (sym)
      for (uint32_t j = 0; j < nv; ++j)
        if (V(j)[0] == -V(i)[0] && V(j)[1] == -V(i)[1] && V(j)[2] == -V(i)[2]) covered[j] = true;
  }
  S->nh = nh;
  S->sym = sym ? 1u : 0u;
  S->tol = (float)std::ldexp(R, -12);
  S->use = 1;
  S->use_edges = S->ne >= 1 && S->ne <= kSatEdges ? 1u : 0u;
}

// Geometry of one system: validation, checkerboard dimensions, model
// constants and the FP32 contact band.  Shared by a context and by every
// replica of a batch.
int make_geometry(const hmc_config* cfg, Geometry& g, bool* fast, double* sigma_out, double* lmax_out) {
  if (cfg->dim != 2 && cfg->dim != 3) return HMC_ERR_UNSUPPORTED;
  const bool lj = is_isotropic(cfg->model);
  const bool convex = cfg->model == HMC_CONVEX_POLYHEDRON;
  const bool op = cfg->model == HMC_OVERLAP_PENALTY;
  if (cfg->model != HMC_HARD_SPHERE && !lj && !convex) return HMC_ERR_UNSUPPORTED;
  if (cfg->precision != HMC_FP64 && cfg->precision != HMC_MIXED) return HMC_ERR_INVALID;
  if (op && cfg->precision != HMC_FP64) return HMC_ERR_UNSUPPORTED;  // FP64 only
  if (convex && cfg->precision != HMC_FP64) return HMC_ERR_UNSUPPORTED;  // FP64 only
  if (convex && cfg->dim != 3) return HMC_ERR_UNSUPPORTED;
  if (convex && (!cfg->vertices || cfg->n_vertices < 1 || cfg->n_vertices > (uint32_t)kMaxVertices))
    return HMC_ERR_INVALID;
  if (cfg->n_bodies >= 0xffffffffull) return HMC_ERR_UNSUPPORTED;
  const int D = (int)cfg->dim;
  // circumsphere radius of a polytope (convex_polytope.rs:248-257): sqrt of
  // fold(0, max) over the vertices' squared norms
  double cv_radius = 0.0;
  if (convex) {
    double m2 = 0.0;
    for (uint32_t i = 0; i < cfg->n_vertices; ++i) {
      const double* v = cfg->vertices + 3 * i;
      if (!std::isfinite(v[0]) || !std::isfinite(v[1]) || !std::isfinite(v[2])) return HMC_ERR_INVALID;
      m2 = std::fmax(m2, std::fma(v[2], v[2], std::fma(v[1], v[1], std::fma(v[0], v[0], 0.0))));
    }
    cv_radius = std::sqrt(m2);
  }
  // interaction range: the diameter (hard_shape.rs:170-174), r_cut (isotropic.rs)
  // or twice the circumsphere radius (hard_shape.rs:111-118)
  const double sigma = op ? cfg->params[3] : lj ? cfg->params[2] : convex ? cv_radius * 2.0 : cfg->params[0];
  if (!(sigma > 0.0) || !(cfg->body_interaction_range > 0.0) || !(cfg->boundary_range >= 0.0) ||
      !(cfg->spatial_cell_width > 0.0))
    return HMC_ERR_INVALID;
  if (lj && !op && (!(cfg->params[0] >= 0.0) || !(cfg->params[1] > 0.0) || !std::isfinite(cfg->params[0]) ||
                   !std::isfinite(cfg->params[1])))
    return HMC_ERR_INVALID;
  if (op && (!(cfg->params[0] >= 0.0) || !(cfg->params[1] >= 0.0) || !(cfg->params[2] >= 0.0) ||
             !std::isfinite(cfg->params[0]) || !std::isfinite(cfg->params[1]) || !std::isfinite(cfg->params[2]) ||
             !std::isfinite(cfg->params[3])))
    return HMC_ERR_INVALID;
  if (cfg->model == HMC_LJ_SHIFTED && !(cfg->params[3] > 0.0)) return HMC_ERR_INVALID;
  double lmin = cfg->box_edges[0];
  for (int a = 0; a < D; ++a) {
    if (!(cfg->box_edges[a] > 0.0)) return HMC_ERR_INVALID;
    lmin = std::min(lmin, cfg->box_edges[a]);
  }
  // Periodic::new (boundary/periodic.rs:96-110)
  if (cfg->boundary_range > lmin / 2.0) return HMC_ERR_INVALID;
  // The pair query needs sigma <= boundary range (microstate.rs:1436-1439) and
  // the checkerboard needs sigma <= body interaction range (trials in one
  // colour must not interact).
  if (sigma > cfg->boundary_range || sigma > cfg->body_interaction_range) return HMC_ERR_INVALID;

  std::memset(&g, 0, sizeof g);
  g.dim = D;
  g.n_bodies = (uint32_t)cfg->n_bodies;
  uint64_t spaces = 1;
  for (int a = 0; a < D; ++a) {
    double L = cfg->box_edges[a];
    g.L[a] = L;
    g.half[a] = L / 2.0;
    g.ghost_hi[a] = g.half[a] - cfg->boundary_range;
    g.ghost_lo[a] = -g.half[a] + cfg->boundary_range;
    g.Lf[a] = (float)L;
    // HypercuboidCheckerboard::compute_dimensions (hypercuboid.rs:188-228)
    int64_t inside = (int64_t)std::floor(L / cfg->body_interaction_range);
    if (inside < 2) {
        return HMC_ERR_INVALID;
    }
    if (inside % 2 != 0) inside -= 1;
    g.shape[a] = (int32_t)inside;
    g.half_shape[a] = (uint32_t)(inside / 2);
    g.width[a] = L / (double)inside;
    // a pair closer than sigma lies in neighbouring spaces unless the width is
    // within rounding of sigma; then the scan reaches two spaces out
    g.reach[a] = g.width[a] > sigma * (1.0 + 1e-9) ? 1 : 2;
    spaces *= (uint64_t)inside;
  }
  if (spaces >= 0xffffffffull || spaces > (uint64_t)kScanBlock * 1024 * 16) {
    return HMC_ERR_UNSUPPORTED;
  }
  uint32_t str = 1;
  for (int a = D - 1; a >= 0; --a) {
    g.stride[a] = str;
    str *= (uint32_t)g.shape[a];
  }
  for (int a = D; a < 3; ++a) {
    g.reach[a] = 0;
    g.shape[a] = 1;
    g.half_shape[a] = 1;
    g.stride[a] = 0;
    g.L[a] = 1.0;
    g.Lf[a] = 1.0f;
    g.width[a] = 1.0;
  }
  g.n_spaces = (uint32_t)spaces;
  g.per_colour = g.n_spaces >> D;
  *fast = true;
  for (int a = 0; a < D; ++a) *fast = *fast && g.reach[a] == 1 && g.shape[a] >= 3;
  g.sigma2 = sigma * sigma;
  g.sigma = sigma;
  g.edge2 = g.sigma2 * (1.0 - 0x1p-30);
  g.model = cfg->model;
  if (convex) {
    g.nverts = cfg->n_vertices;
    g.cv_rcut2 = (cv_radius + cv_radius) * (cv_radius + cv_radius);
    g.cv_edge2 = g.cv_rcut2 * (1.0 - 0x1p-30);
  }
  if (op) {
    // overlap_penalty.rs:109-115 (0.5 k as the reference forms it), r_cut = delta
    g.op_half_k = 0.5 * cfg->params[0];
    g.op_dmax = cfg->params[1];
    g.op_eps = cfg->params[2];
    g.op_delta = cfg->params[3];
    g.lj_rc2 = sigma * sigma;
    g.lj_rc2_edge = g.lj_rc2 * (1.0 - 0x1p-30);
  } else if (lj) {
    // lennard_jones.rs:107-111, shifted.rs:92-94, isotropic.rs:147-156
    const double eps = cfg->params[0], lsig = cfg->params[1], rc = cfg->params[2];
    g.lj_eps4 = 4.0 * eps;
    g.lj_sigma = lsig;
    g.lj_rc2 = rc * rc;
    g.lj_rc2_edge = g.lj_rc2 * (1.0 - 0x1p-30);
    g.lj_shift = 0.0;
    if (cfg->model == HMC_LJ_SHIFTED) {
      const double sr = lsig / cfg->params[3];
      const double s2 = sr * sr, s4 = s2 * s2, s8 = s4 * s4;
      g.lj_shift = g.lj_eps4 * (s4 * s8 - s2 * s4);
    }
  }
  g.vc_width = cfg->spatial_cell_width;
  g.vc_k = std::ceil(sigma / cfg->spatial_cell_width);
  // Mixed mode band of the global pass.  Per axis |d32 - d64| <= E with
  // E = 2^-24 (3.5 Lmax + 2 sigma) + 2^-53 (1.5 Lmax + 2 sigma), from the
  // roundings of t, x, L, x + T and t - x.
  double lmax = 0.0;
  for (int a = 0; a < D; ++a) lmax = std::max(lmax, g.L[a]);
  const double E = (0x1p-24 * (3.5 * lmax + 2.0 * sigma) + 0x1p-53 * (1.5 * lmax + 2.0 * sigma)) * 1.01;
  if (!contact_band(E, sigma, g.sigma2, g.edge2, D, &g.band_lo, &g.band_hi) && cfg->precision == HMC_MIXED) {
    return HMC_ERR_UNSUPPORTED;
  }

  *sigma_out = sigma;
  *lmax_out = lmax;
  return HMC_OK;
}

}  // namespace hmc

extern "C" {

int hmc_create(const hmc_config* cfg_in, hmc_ctx** out) {
  if (!cfg_in || !out) return HMC_ERR_INVALID;
  *out = nullptr;
  // LJ in mixed precision keeps the FP64 data layout and passes; only its
  // pair energies run in FP32 (Geometry::lj_mixed)
  hmc_config cfg_local = *cfg_in;
  const bool lj_mixed = (cfg_in->model == HMC_LJ || cfg_in->model == HMC_LJ_SHIFTED) &&
                        cfg_in->precision == HMC_MIXED;
  if (lj_mixed) cfg_local.precision = HMC_FP64;
  // so do convex polyhedra: their listed pairs go through FP32 separating
  // axes first (Geometry::cv_mixed)
  const bool cv_mixed = cfg_in->model == HMC_CONVEX_POLYHEDRON && cfg_in->precision == HMC_MIXED;
  if (cv_mixed) cfg_local.precision = HMC_FP64;
  const hmc_config* cfg = &cfg_local;
  Geometry g0;
  bool fast0 = false;
  double sigma = 0.0, lmax = 0.0;
  {
    const int rc = make_geometry(cfg, g0, &fast0, &sigma, &lmax);
    if (rc) return rc;
  }
  const int D = (int)cfg->dim;
  const bool lj = is_isotropic(cfg->model);
  const bool convex = cfg->model == HMC_CONVEX_POLYHEDRON;
  const uint64_t spaces = g0.n_spaces;
  auto* c = new hmc_ctx;
  c->cfg = *cfg;
  c->cfg_user = *cfg_in;
  if (convex) c->verts_host.assign(cfg_in->vertices, cfg_in->vertices + 3 * (size_t)cfg_in->n_vertices);
  c->cfg_user.vertices = nullptr;
  c->device = cfg->device;
  c->g = g0;
  c->g.lj_mixed = lj_mixed ? 1 : 0;
  c->g.cv_mixed = cv_mixed ? 1 : 0;
  if (cv_mixed) sat_build(cfg_in->vertices, cfg_in->n_vertices, &c->sat);
  c->g.lj_fp32_error = (cfg_in->flags & HMC_FLAG_FP32_ERROR) && cfg_in->precision == HMC_MIXED ? 1 : 0;
  c->fast = fast0;
  Geometry& g = c->g;
  if (hipSetDevice(c->device) != hipSuccess) {
    delete c;
    return HMC_ERR_DEVICE;
  }
  if (noalloc_setup(c->device) != HMC_OK) {  // the static scratch of hmc_total_energy_noalloc
    delete c;
    return HMC_ERR_DEVICE;
  }
  if (c->fast && !lj && !convex && !(cfg->flags & HMC_FLAG_GLOBAL_PASS)) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || cus <= 0) {
      delete c;
      return HMC_ERR_DEVICE;
    }
    const size_t vt = cfg->precision == HMC_MIXED ? sizeof(float4) : sizeof(double4);
    c->tiled = choose_tiles(g, D, (double)cfg->n_bodies / (double)spaces, vt, cus, &c->tc);
    if (c->tiled && cfg->precision == HMC_MIXED) {
      // Band of the box-relative test.  Per axis: the FP32 roundings of x and
      // t (2^-25 L each), of x - c and t - c (2^-24 L each), of L in the image
      // shift (2^-24 L, twice), of the shifted values (2^-24 R each, R the
      // largest box-relative coordinate) and of their difference
      // (2^-24 |d|); the reference's own roundings add 2^-53 (1.5 L + |d|).
      double R = 0.0;
      for (int a = 0; a < D; ++a) R = std::max(R, 0.5 * c->tc.E[a] * g.width[a] + g.width[a]);
      const double Et = (0x1p-24 * (5.1 * lmax + 2.0 * R + 2.0 * sigma) + 0x1p-53 * (1.5 * lmax + 2.0 * sigma)) * 1.01;
      c->tiled = contact_band(Et, sigma, g.sigma2, g.edge2, D, &c->tc.band_lo, &c->tc.band_hi);
    }
    if (c->tiled) {
      const void* fn = nullptr;
      if (D == 2)
        fn = cfg->precision == HMC_MIXED ? (const void*)k_pass_tile<2, HMC_MIXED> : (const void*)k_pass_tile<2, HMC_FP64>;
      else
        fn = cfg->precision == HMC_MIXED ? (const void*)k_pass_tile<3, HMC_MIXED> : (const void*)k_pass_tile<3, HMC_FP64>;
      const void* fe = D == 2 ? (const void*)k_pass_tile<2, HMC_MIXED, true> : (const void*)k_pass_tile<3, HMC_MIXED, true>;
      if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)c->tc.lds_bytes) != hipSuccess ||
          hipFuncSetAttribute(fe, hipFuncAttributeMaxDynamicSharedMemorySize, (int)c->tc.lds_bytes) != hipSuccess) {
        delete c;
        return HMC_ERR_DEVICE;
      }
    }
    int coop = 0;
    // (the FP32 error report runs the per-colour passes)
    if (c->tiled && cfg->precision == HMC_MIXED && !(cfg->flags & (HMC_FLAG_PER_COLOUR | HMC_FLAG_FP32_ERROR)) &&
        hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, c->device) == hipSuccess && coop) {
      TileCfg pc{};
      if (choose_regions(g, D, (double)cfg->n_bodies / (double)spaces, cus, &pc)) {
        // Band of the persistent layout, whose LDS entries are the FP64
        // box-relative positions rounded once to FP32.  Per axis: that
        // rounding for the site and for the trial site (2^-24 R each, R the
        // largest box-relative coordinate), the rounding of their FP32
        // difference (2^-24 |d|), and the FP64 roundings of x - c and of the
        // image shift on both sides as well as of the reference's own
        // displacement (2^-53 (4 L + 2 |d|)).
        double R = 0.0;
        for (int a = 0; a < D; ++a) R = std::max(R, 0.5 * pc.E[a] * g.width[a] + g.width[a]);
        const double Et = (0x1p-24 * (2.0 * R + 2.0 * sigma) + 0x1p-53 * (4.0 * lmax + 2.0 * sigma)) * 1.01;
        const void* fn = D == 2 ? (const void*)k_apply_tile<2> : (const void*)k_apply_tile<3>;
        const uint32_t threads = ((pc.trials + 63) / 64) * 64;
        int per_cu = 0;
        if (contact_band(Et, sigma, g.sigma2, g.edge2, D, &pc.band_lo, &pc.band_hi) &&
            hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pc.lds_bytes) == hipSuccess &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, (int)threads, pc.lds_bytes) == hipSuccess &&
            (uint64_t)per_cu * (uint64_t)cus >= (uint64_t)pc.xcd_chunk * 8u) {
          c->persist = true;
          c->pc = pc;
          c->pgrid = pc.xcd_chunk * 8u;
          c->pthreads = threads;
          c->pper_cu = (uint32_t)per_cu;
          const size_t fbytes = (size_t)pc.ntiles * kMailboxStride * sizeof(uint32_t);
          if (hipMalloc((void**)&c->w.mbox, fbytes) != hipSuccess || hipMemset(c->w.mbox, 0, fbytes) != hipSuccess) {
            delete c;
            return HMC_ERR_DEVICE;
          }
        }
      }
    }
  }
  // Isotropic pair energies in mixed precision: one persistent launch per
  // apply (k_apply_lj) when every axis has at least six spaces and the
  // neighbour scan reaches one space out.
  if (lj_mixed && !(cfg_in->flags & (HMC_FLAG_PER_COLOUR | HMC_FLAG_FP32_ERROR | HMC_FLAG_GLOBAL_PASS))) {
    int cus = 0, coop = 0;
    TileCfg pc{};
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) == hipSuccess && cus > 0 &&
        hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, c->device) == hipSuccess && coop &&
        choose_regions_lj(g, D, (double)cfg->n_bodies / (double)spaces, cus, &pc)) {
      // Per-axis bound E on the FP32 displacement of two staged entries
      // against the reference's FP64 displacement: the rounding of each
      // box-relative coordinate to FP32 (2^-24 R, R the largest), the FP64
      // roundings of x - c and of the image shift on both sides
      // (2^-53 (L + R) each) and the reference's own displacement and image
      // shift (2^-53 (r_cut + 1.5 L)).  The membership band around r_cut^2
      // is the pair bound beta of k_apply_lj at |d| = r_cut, widened by 1%.
      double R = 0.0;
      for (int a = 0; a < D; ++a) R = std::max(R, 0.5 * pc.E[a] * g.width[a] + g.width[a]);
      const double rc2 = g.lj_rc2, rc = std::sqrt(rc2) * (1.0 + 0x1p-20);
      const double E = (0x1p-23 * R + 0x1p-53 * (3.5 * lmax + 2.0 * R + rc)) * 1.01;
      const double beta = (E * (2.0 * std::sqrt((double)D) * rc + D * E) + rc * rc * (5.0 * 0x1p-24 + 0x1p-50)) * 1.01;
      float lo = (float)(rc2 - beta), hi = (float)(rc2 + beta);
      while ((double)lo > rc2 - beta) lo = std::nextafter(lo, 0.0f);
      while ((double)hi < rc2 + beta) hi = std::nextafter(hi, INFINITY);
      float pe = (float)E;
      while ((double)pe < E) pe = std::nextafter(pe, INFINITY);
      pc.band_lo = lo;
      pc.band_hi = hi;
      pc.perr = pe;
      // |r2_32 - r2_ref| <= E (2 sum|d_a| + D E) + (5 u + 2^-50) r2 for every
      // pair counted (r2 < band_hi, so sum|d_a| <= sqrt(D band_hi)), and the
      // first-order bound on the pair energy holds within 2% where that is
      // at most r2 / 1024
      const double bmax = (E * (2.0 * std::sqrt((double)D * (double)hi) + D * E) +
                           (double)hi * (5.0 * 0x1p-24 + 0x1p-50)) * 1.01;
      float bm = (float)bmax;
      while ((double)bm < bmax) bm = std::nextafter(bm, INFINITY);
      float ro = (float)(1024.0 * (double)bm);
      while ((double)ro < 1024.0 * (double)bm) ro = std::nextafter(ro, INFINITY);
      pc.beta_max = bm;
      pc.r2_open = ro;
      const void* fn = D == 2 ? (const void*)k_apply_lj<2> : (const void*)k_apply_lj<3>;
      const uint32_t threads = lj_region_threads(pc.trials);
      int per_cu = 0;
      if (beta < 0.01 * rc2 &&
          hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pc.lds_bytes) == hipSuccess &&
          hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, (int)threads, pc.lds_bytes) == hipSuccess &&
          (uint64_t)per_cu * (uint64_t)cus >= (uint64_t)pc.xcd_chunk * 8u) {
        c->persist = true;
        c->pc = pc;
        c->pgrid = pc.xcd_chunk * 8u;
        c->pthreads = threads;
        c->pper_cu = (uint32_t)per_cu;
        const size_t fbytes = (size_t)pc.ntiles * kMailboxStride * sizeof(uint32_t);
        if (hipMalloc((void**)&c->w.mbox, fbytes) != hipSuccess || hipMemset(c->w.mbox, 0, fbytes) != hipSuccess) {
          delete c;
          return HMC_ERR_DEVICE;
        }
      }
    }
  }
  const uint32_t N = g.n_bodies;
  c->n_scan_blocks = grid_for(g.n_spaces, kScanBlock);
  bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess;
  ok = ok && dalloc(&c->bodies.bpos, N) == hipSuccess && dalloc(&c->bodies.spos, N) == hipSuccess &&
       dalloc(&c->cells.body, N) == hipSuccess;
  if (cfg->precision == HMC_MIXED)
    ok = ok && dalloc(&c->cells.pos32, N) == hipSuccess;
  else
    ok = ok && dalloc(&c->cells.pos64, N) == hipSuccess;
  ok = ok && dalloc(&c->w.cnt, g.n_spaces) == hipSuccess && dalloc(&c->w.start, g.n_spaces + 1) == hipSuccess &&
       dalloc(&c->w.elem_space, N) == hipSuccess && dalloc(&c->w.elem_rank, N) == hipSuccess &&
       dalloc(&c->w.tmp_body, N) == hipSuccess && dalloc(&c->w.block_sums, c->n_scan_blocks) == hipSuccess &&
       dalloc(&c->w.block_ne, c->n_scan_blocks) == hipSuccess && dalloc(&c->st, 1) == hipSuccess &&
       dalloc(&c->w.acc_slots, kAccSlots) == hipSuccess && dalloc(&c->w.scan_status, c->n_scan_blocks) == hipSuccess;
  if (ok && convex) {
    // orientations start at the identity; vertices live on the device
    ok = dalloc(&c->bodies.bquat, N) == hipSuccess && dalloc(&c->bodies.squat, N) == hipSuccess &&
         dalloc(&c->cells.quat64, N) == hipSuccess && dalloc(&c->verts, 3 * (size_t)g.nverts) == hipSuccess;
    std::vector<double4> ident(std::max<uint32_t>(N, 1), make_double4(1.0, 0.0, 0.0, 0.0));
    ok = ok && hipMemcpy(c->bodies.bquat, ident.data(), N * sizeof(double4), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(c->bodies.squat, ident.data(), N * sizeof(double4), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(c->verts, cfg->vertices, 3 * (size_t)g.nverts * sizeof(double), hipMemcpyHostToDevice) ==
             hipSuccess;
    g.verts = c->verts;
  }
  if (!ok) {
    hmc_destroy(c);
    return HMC_ERR_DEVICE;
  }
  if (hipMemset(c->st, 0, sizeof(ApplyState)) != hipSuccess ||
      hipMemset(c->w.scan_status, 0, c->n_scan_blocks * sizeof(unsigned long long)) != hipSuccess ||
      hipMemset(c->w.acc_slots, 0, kAccSlots * sizeof(unsigned long long)) != hipSuccess ||
      hipMemset(c->w.cnt, 0, (size_t)std::max<uint32_t>(g.n_spaces, 1) * sizeof(uint32_t)) != hipSuccess) {
    hmc_destroy(c);
    return HMC_ERR_DEVICE;
  }
  // first guess of the number of colour cycles per apply: every space occupied
  c->cycles_guess = (uint32_t)std::max<uint64_t>(1, (N + g.n_spaces - 1) / std::max<uint32_t>(g.n_spaces, 1));
  // the device copy of the geometry (complete from here on)
  if (dalloc(&c->gdev, 1) != hipSuccess) {
    hmc_destroy(c);
    return HMC_ERR_DEVICE;
  }
  g.self = c->gdev;
  if (hipMemcpy(c->gdev, &g, sizeof(Geometry), hipMemcpyHostToDevice) != hipSuccess) {
    hmc_destroy(c);
    return HMC_ERR_DEVICE;
  }
  if (c->persist && !persist_probe(c)) c->persist = false;  // the per-colour passes instead
  *out = c;
  return HMC_OK;
}

int hmc_set_state(hmc_ctx* c, const double* body_pos, const double* site_pos) {
  if (!c || (!body_pos && c->g.n_bodies)) return HMC_ERR_INVALID;
  const uint32_t N = c->g.n_bodies;
  const int D = c->g.dim;
  std::vector<double4> hb(N), hs(site_pos ? N : 0);
  for (uint32_t i = 0; i < N; ++i) {
    double v[3] = {0, 0, 0};
    for (int a = 0; a < D; ++a) v[a] = body_pos[(size_t)i * D + a];
    hb[i] = make_double4(v[0], v[1], v[2], 0.0);
    if (site_pos) {
      double u[3] = {0, 0, 0};
      for (int a = 0; a < D; ++a) u[a] = site_pos[(size_t)i * D + a];
      hs[i] = make_double4(u[0], u[1], u[2], 0.0);
    }
  }
  HIPCHK(c, hipSetDevice(c->device));
  if (N) {
    HIPCHK(c, hipMemcpyAsync(c->bodies.bpos, hb.data(), N * sizeof(double4), hipMemcpyHostToDevice, c->stream));
    if (site_pos)
      HIPCHK(c, hipMemcpyAsync(c->bodies.spos, hs.data(), N * sizeof(double4), hipMemcpyHostToDevice, c->stream));
    if (D == 2)
      k_init_sites<2><<<grid_for(N, 256), 256, 0, c->stream>>>(c->g, c->bodies, site_pos ? 1 : 0);
    else
      k_init_sites<3><<<grid_for(N, 256), 256, 0, c->stream>>>(c->g, c->bodies, site_pos ? 1 : 0);
    HIPCHK(c, hipGetLastError());
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->have_state = true;
  return HMC_OK;
}

int hmc_set_orientations(hmc_ctx* c, const double* body_quat, const double* site_quat) {
  if (!c || (!body_quat && c->g.n_bodies)) return HMC_ERR_INVALID;
  if (c->g.model != HMC_CONVEX_POLYHEDRON) return fail(c, HMC_ERR_UNSUPPORTED, "bodies have no orientation");
  const uint32_t N = c->g.n_bodies;
  std::vector<double4> hb(N), hs(site_quat ? N : 0);
  for (uint32_t i = 0; i < N; ++i) {
    const double* q = body_quat + 4 * (size_t)i;
    hb[i] = make_double4(q[0], q[1], q[2], q[3]);
    if (site_quat) {
      const double* u = site_quat + 4 * (size_t)i;
      hs[i] = make_double4(u[0], u[1], u[2], u[3]);
    }
  }
  HIPCHK(c, hipSetDevice(c->device));
  if (N) {
    HIPCHK(c, hipMemcpyAsync(c->bodies.bquat, hb.data(), N * sizeof(double4), hipMemcpyHostToDevice, c->stream));
    if (site_quat)
      HIPCHK(c, hipMemcpyAsync(c->bodies.squat, hs.data(), N * sizeof(double4), hipMemcpyHostToDevice, c->stream));
    k_init_site_quats<<<grid_for(N, 256), 256, 0, c->stream>>>(c->g, c->bodies, site_quat ? 1 : 0);
    HIPCHK(c, hipGetLastError());
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HMC_OK;
}

int hmc_get_orientations(hmc_ctx* c, double* body_quat, double* site_quat) {
  if (!c) return HMC_ERR_INVALID;
  if (c->g.model != HMC_CONVEX_POLYHEDRON) return fail(c, HMC_ERR_UNSUPPORTED, "bodies have no orientation");
  const uint32_t N = c->g.n_bodies;
  std::vector<double4> hb(N), hs(N);
  HIPCHK(c, hipSetDevice(c->device));
  if (N) {
    HIPCHK(c, hipMemcpyAsync(hb.data(), c->bodies.bquat, N * sizeof(double4), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(hs.data(), c->bodies.squat, N * sizeof(double4), hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (uint32_t i = 0; i < N; ++i) {
    const double4 b = hb[i], q = hs[i];
    if (body_quat) {
      body_quat[4 * (size_t)i] = b.x;
      body_quat[4 * (size_t)i + 1] = b.y;
      body_quat[4 * (size_t)i + 2] = b.z;
      body_quat[4 * (size_t)i + 3] = b.w;
    }
    if (site_quat) {
      site_quat[4 * (size_t)i] = q.x;
      site_quat[4 * (size_t)i + 1] = q.y;
      site_quat[4 * (size_t)i + 2] = q.z;
      site_quat[4 * (size_t)i + 3] = q.w;
    }
  }
  return HMC_OK;
}

int hmc_debug_convex(int device, int kind, const double* vertices, uint32_t n_vertices, const double* in,
                     double* out, uint64_t n) {
  if (!vertices || n_vertices < 1 || n_vertices > (uint32_t)kMaxVertices || (n && (!in || !out)) || kind < 0 ||
      kind > 3)
    return HMC_ERR_INVALID;
  if (hipSetDevice(device) != hipSuccess) return HMC_ERR_DEVICE;
  const size_t w_in = kind == 0 ? 7 : (kind == 1 || kind == 3) ? 14 : 5, w_out = kind == 2 ? 4 : 1;
  SatShape sat;
  sat_build(vertices, n_vertices, &sat);
  double m2 = 0.0;
  for (uint32_t i = 0; i < n_vertices; ++i) {
    const double* v = vertices + 3 * i;
    m2 = std::fmax(m2, std::fma(v[2], v[2], std::fma(v[1], v[1], std::fma(v[0], v[0], 0.0))));
  }
  const double R = std::sqrt(m2);
  double *dv = nullptr, *di = nullptr, *dout = nullptr;
  int rc = HMC_OK;
  if (hipMalloc((void**)&dv, 3 * n_vertices * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&di, std::max<uint64_t>(n, 1) * w_in * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&dout, std::max<uint64_t>(n, 1) * w_out * sizeof(double)) != hipSuccess) {
    rc = HMC_ERR_DEVICE;
  } else if (n) {
    if (hipMemcpy(dv, vertices, 3 * n_vertices * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(di, in, n * w_in * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      rc = HMC_ERR_DEVICE;
    if (!rc) {
      k_debug_convex<<<grid_for(n, 64), 64>>>(kind, dv, n_vertices, (R + R) * (R + R), di, dout, n, sat);
      if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
          hipMemcpy(out, dout, n * w_out * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        rc = HMC_ERR_DEVICE;
    }
  }
  if (dv) (void)hipFree(dv);
  if (di) (void)hipFree(di);
  if (dout) (void)hipFree(dout);
  return rc;
}

int hmc_debug_convex_axes(const double* vertices, uint32_t n_vertices, uint32_t* counts, double* faces,
                          double* edges) {
  if (!vertices || n_vertices < 1 || n_vertices > (uint32_t)kMaxVertices || !counts) return HMC_ERR_INVALID;
  SatShape S;
  sat_build(vertices, n_vertices, &S);
  const uint32_t v[7] = {S.nf, S.ne, S.nh, S.sym, S.use, S.use_edges, 0};
  for (int i = 0; i < 6; ++i) counts[i] = v[i];
  if (faces)
    for (uint32_t f = 0; f < std::min(S.nf, kSatFaces); ++f) {
      for (int a = 0; a < 3; ++a) faces[5 * f + a] = S.face[f][a];
      faces[5 * f + 3] = S.fmin[f];
      faces[5 * f + 4] = S.face[f][3];
    }
  if (edges)
    for (uint32_t e = 0; e < std::min(S.ne, kSatEdges); ++e)
      for (int a = 0; a < 3; ++a) edges[3 * e + a] = S.edge[e][a];
  return HMC_OK;
}

int hmc_clone_with_boundary(hmc_ctx* c, const double* new_edges, hmc_ctx** out) {
  if (!c || !new_edges || !out) return HMC_ERR_INVALID;
  *out = nullptr;
  if (!c->have_state) return fail(c, HMC_ERR_STATE, "no state");
  const int D = c->g.dim;
  const uint32_t N = c->g.n_bodies;
  hmc_config cfg = c->cfg_user;
  for (int a = 0; a < D; ++a) cfg.box_edges[a] = new_edges[a];
  if (!c->verts_host.empty()) cfg.vertices = c->verts_host.data();
  // the same ranges on the new edges: hmc_create applies Periodic::new's
  // criterion (the scale_volume assertion, periodic.rs:288-300) and the
  // checkerboard's
  hmc_ctx* d = nullptr;
  const int rc = hmc_create(&cfg, &d);
  if (rc) return fail(c, rc, "the new boundary does not fit the configuration's interaction ranges");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  uint32_t outside = 0;
  hipError_t e = hipSuccess;
  if (N) {
    uint32_t* flag = nullptr;
    e = hipMallocAsync((void**)&flag, sizeof(uint32_t), d->stream);
    if (e == hipSuccess) e = hipMemsetAsync(flag, 0, sizeof(uint32_t), d->stream);
    if (e == hipSuccess) {
      const double3 lold = make_double3(c->g.L[0], c->g.L[1], c->g.L[2]);
      if (D == 2)
        k_map_bodies<2><<<grid_for(N, 256), 256, 0, d->stream>>>(d->g, lold, c->bodies.bpos, d->bodies, flag);
      else
        k_map_bodies<3><<<grid_for(N, 256), 256, 0, d->stream>>>(d->g, lold, c->bodies.bpos, d->bodies, flag);
      e = hipGetLastError();
    }
    if (e == hipSuccess && c->g.model == HMC_CONVEX_POLYHEDRON) {
      // orientations are kept; the sites' are rebuilt from the bodies'
      e = hipMemcpyAsync(d->bodies.bquat, c->bodies.bquat, N * sizeof(double4), hipMemcpyDeviceToDevice, d->stream);
      if (e == hipSuccess) {
        k_init_site_quats<<<grid_for(N, 256), 256, 0, d->stream>>>(d->g, d->bodies, 0);
        e = hipGetLastError();
      }
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&outside, flag, sizeof(uint32_t), hipMemcpyDeviceToHost, d->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    if (flag) (void)hipFreeAsync(flag, d->stream);
  }
  if (e != hipSuccess) {
    hmc_destroy(d);
    return fail(c, HMC_ERR_DEVICE, std::string("clone_with_boundary: ") + hipGetErrorString(e));
  }
  if (outside) {
    hmc_destroy(d);
    return fail(c, HMC_ERR_INVALID, "a body lies outside the boundary (map_point: PointOutsideShape)");
  }
  d->have_state = true;
  *out = d;
  return HMC_OK;
}

int hmc_get_state(hmc_ctx* c, double* body_pos, double* site_pos) {
  if (!c) return HMC_ERR_INVALID;
  if (!c->have_state) return fail(c, HMC_ERR_STATE, "no state");
  const uint32_t N = c->g.n_bodies;
  const int D = c->g.dim;
  std::vector<double4> hb(N), hs(N);
  HIPCHK(c, hipSetDevice(c->device));
  if (N) {
    HIPCHK(c, hipMemcpyAsync(hb.data(), c->bodies.bpos, N * sizeof(double4), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(hs.data(), c->bodies.spos, N * sizeof(double4), hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (uint32_t i = 0; i < N; ++i) {
    double v[3] = {hb[i].x, hb[i].y, hb[i].z};
    double u[3] = {hs[i].x, hs[i].y, hs[i].z};
    for (int a = 0; a < D; ++a) {
      if (body_pos) body_pos[(size_t)i * D + a] = v[a];
      if (site_pos) site_pos[(size_t)i * D + a] = u[a];
    }
  }
  return HMC_OK;
}

static int apply_one(hmc_ctx* c, uint64_t step, uint32_t substep, uint32_t seed, double kT, double max_move,
                     uint32_t move, hmc_count* out, uint32_t* substeps_used) {
  if (!c) return HMC_ERR_INVALID;
  if (!c->have_state) return fail(c, HMC_ERR_STATE, "hmc_set_state was not called");
  if (!(max_move > 0.0) || !(kT > 0.0)) return fail(c, HMC_ERR_INVALID, "the move size and kT must be > 0");
  if (move == 1 && c->g.model != HMC_CONVEX_POLYHEDRON)
    return fail(c, HMC_ERR_UNSUPPORTED, "rotation moves need oriented bodies");
  HIPCHK(c, hipSetDevice(c->device));
  ApplyState before;
  int rc = read_state(c, &before);
  if (rc) return rc;
  rc = dispatch_apply(c, 1, 0, step, substep, seed, kT, max_move, move);
  if (rc) return rc;
  ApplyState after;
  rc = read_state(c, &after);
  if (rc) return rc;
  c->cycles_guess = std::max<uint32_t>(after.cycles, 1);
  uint64_t acc = after.accepted_total - before.accepted_total;
  uint64_t tri = after.trials_total - before.trials_total;
  if (out) {
    out->accepted = acc;
    out->rejected = tri - acc;
  }
  if (substeps_used) *substeps_used = after.cycles;
  return HMC_OK;
}

int hmc_apply_translate(hmc_ctx* c, uint64_t step, uint32_t substep, uint32_t seed, double kT,
                        double max_distance, hmc_count* out, uint32_t* substeps_used) {
  return apply_one(c, step, substep, seed, kT, max_distance, 0u, out, substeps_used);
}

int hmc_apply_rotate(hmc_ctx* c, uint64_t step, uint32_t substep, uint32_t seed, double kT, double max_rotation,
                     hmc_count* out, uint32_t* substeps_used) {
  return apply_one(c, step, substep, seed, kT, max_rotation, 1u, out, substeps_used);
}

static bool same_moves(const std::vector<hmc_move>& a, const hmc_move* b, uint32_t n) {
  if (a.size() != n) return false;
  for (uint32_t i = 0; i < n; ++i)
    if (a[i].move != b[i].move || a[i].size != b[i].size) return false;
  return true;
}

int hmc_run(hmc_ctx* c, uint64_t step0, uint32_t seed, double kT, const hmc_move* moves, uint32_t n_moves,
            uint32_t n_steps, hmc_count* out) {
  if (!c) return HMC_ERR_INVALID;
  if (!c->have_state) return fail(c, HMC_ERR_STATE, "hmc_set_state was not called");
  if (!moves || n_moves < 1 || n_moves > (uint32_t)kMaxMoves)
    return fail(c, HMC_ERR_INVALID, "hmc_run: 1 to 4 moves per step");
  if (!(kT > 0.0)) return fail(c, HMC_ERR_INVALID, "kT must be > 0");
  for (uint32_t m = 0; m < n_moves; ++m) {
    if (moves[m].move != HMC_MOVE_TRANSLATE && moves[m].move != HMC_MOVE_ROTATE)
      return fail(c, HMC_ERR_INVALID, "hmc_run: unknown move");
    if (!(moves[m].size > 0.0)) return fail(c, HMC_ERR_INVALID, "the move sizes must be > 0");
    if (moves[m].move == HMC_MOVE_ROTATE && c->g.model != HMC_CONVEX_POLYHEDRON)
      return fail(c, HMC_ERR_UNSUPPORTED, "rotation moves need oriented bodies");
  }
  HIPCHK(c, hipSetDevice(c->device));
  ApplyState before;
  int rc = read_state(c, &before);
  if (rc) return rc;
  const bool tally = n_moves > 1;
  if (tally) k_tally<<<1, kAccSlots, 0, c->stream>>>(c->st, c->w.acc_slots, (uint32_t)kMaxMoves);
  const bool lj = is_isotropic(c->g.model);
  // (a cooperative launch is not captured: the persistent apply launches directly)
  const bool graphs = !(c->cfg.flags & HMC_FLAG_NO_GRAPH) && !c->persist;
  // one step: apply each move in order (the step's substeps follow on), then
  // Microstate::increment_step before the next step
  auto step_launches = [&](bool first) -> int {
    for (uint32_t m = 0; m < n_moves; ++m) {
      const int from_args = first && m == 0 ? 1 : 0;
      const int advance = m > 0 ? 2 : (first ? 0 : 1);
      const int r = dispatch_apply(c, from_args, advance, step0, 0, seed, kT, moves[m].size, (uint32_t)moves[m].move);
      if (r) return r;
      if (tally) k_tally<<<1, kAccSlots, 0, c->stream>>>(c->st, c->w.acc_slots, m);
    }
    return HMC_OK;
  };
  for (uint32_t i = 0; i < n_steps; ++i) {
    const bool timed = c->profiling && c->applies % c->profile_period == 0;
    if (i == 0 || !graphs || timed) {
      rc = step_launches(i == 0);
      if (rc) return rc;
      continue;
    }
    // later steps: replay one captured step
    const uint32_t launched = c->cycles_guess + (lj ? 1u : 0u);
    if (!c->graph_exec || c->graph_kT != kT || !same_moves(c->graph_moves, moves, n_moves) ||
        c->graph_launched != launched) {
      drop_graph(c);
      HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      c->capturing = true;
      rc = step_launches(false);
      c->capturing = false;
      c->applies -= n_moves;  // counted when replayed
      hipGraph_t gr = nullptr;
      const hipError_t ce = hipStreamEndCapture(c->stream, &gr);
      if (rc) return rc;
      HIPCHK(c, ce);
      c->graph = gr;
      HIPCHK(c, hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
      c->graph_kT = kT;
      c->graph_moves.assign(moves, moves + n_moves);
      c->graph_launched = launched;
    }
    HIPCHK(c, hipGraphLaunch(c->graph_exec, c->stream));
    c->applies += n_moves;
  }
  ApplyState after;
  rc = read_state(c, &after);
  if (rc) return rc;
  if (n_steps) c->cycles_guess = std::max<uint32_t>(after.cycles, 1);
  if (out) {
    if (tally) {
      for (uint32_t m = 0; m < n_moves; ++m) {
        out[m].accepted = after.move_acc[m];
        out[m].rejected = after.move_trials[m] - after.move_acc[m];
      }
    } else {
      const uint64_t acc = after.accepted_total - before.accepted_total;
      const uint64_t tri = after.trials_total - before.trials_total;
      out[0].accepted = acc;
      out[0].rejected = tri - acc;
    }
  }
  return HMC_OK;
}

int hmc_run_translate(hmc_ctx* c, uint64_t step0, uint32_t seed, double kT, double max_distance,
                      uint32_t n_steps, hmc_count* out) {
  if (!c) return HMC_ERR_INVALID;
  if (!(max_distance > 0.0) || !(kT > 0.0)) return fail(c, HMC_ERR_INVALID, "max_distance and kT must be > 0");
  hmc_move m{};
  m.move = HMC_MOVE_TRANSLATE;
  m.size = max_distance;
  return hmc_run(c, step0, seed, kT, &m, 1, n_steps, out);
}

int hmc_moved_since(hmc_ctx* c, uint32_t since, uint32_t* now, uint64_t* mask, uint64_t* n_moved) {
  if (!c || !mask) return HMC_ERR_INVALID;
  if (!c->have_state) return fail(c, HMC_ERR_STATE, "hmc_set_state was not called");
  HIPCHK(c, hipSetDevice(c->device));
  const uint32_t N = c->g.n_bodies;
  const size_t words = ((size_t)N + 63) / 64;
  hipStream_t s = c->stream;
  unsigned long long* d = nullptr;  // mask words, then the count
  HIPCHK(c, hipMallocAsync((void**)&d, (words + 1) * sizeof(unsigned long long), s));
  hipError_t e = hipMemsetAsync(d + words, 0, sizeof(unsigned long long), s);
  if (e == hipSuccess && N) {
    k_moved_mask<<<grid_for(N, 256), 256, 0, s>>>(c->bodies.bpos, N, since, d, d + words);
    e = hipGetLastError();
  }
  unsigned long long cnt = 0;
  if (e == hipSuccess && words) e = hipMemcpyAsync(mask, d, words * sizeof(uint64_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(&cnt, d + words, sizeof cnt, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFreeAsync(d, s);
  if (e != hipSuccess) return fail(c, HMC_ERR_DEVICE, std::string("moved bodies: ") + hipGetErrorString(e));
  ApplyState st;
  const int rc = read_state(c, &st);
  if (rc) return rc;
  if (now) *now = st.epoch;
  if (n_moved) *n_moved = cnt;
  return HMC_OK;
}

int hmc_get_bodies(hmc_ctx* c, const uint32_t* index, uint64_t n, double* body_pos, double* site_pos,
                   double* body_quat, double* site_quat) {
  if (!c || (n && !index)) return HMC_ERR_INVALID;
  if (!c->have_state) return fail(c, HMC_ERR_STATE, "hmc_set_state was not called");
  const uint32_t N = c->g.n_bodies;
  const int D = c->g.dim;
  if ((body_quat || site_quat) && !c->bodies.bquat)
    return fail(c, HMC_ERR_UNSUPPORTED, "the bodies have no orientations");
  if (n >= 0xffffffffull) return fail(c, HMC_ERR_INVALID, "too many bodies");
  for (uint64_t k = 0; k < n; ++k)
    if (index[k] >= N) return fail(c, HMC_ERR_INVALID, "a body index is out of range");
  if (n == 0) return HMC_OK;
  HIPCHK(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  uint32_t* di = nullptr;
  double4* dout = nullptr;
  hipError_t e = hipMallocAsync((void**)&di, n * sizeof(uint32_t), s);
  if (e == hipSuccess) e = hipMallocAsync((void**)&dout, 4 * n * sizeof(double4), s);
  if (e == hipSuccess) e = hipMemcpyAsync(di, index, n * sizeof(uint32_t), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    k_gather_bodies<<<grid_for(n, 256), 256, 0, s>>>(c->bodies, di, (uint32_t)n, dout);
    e = hipGetLastError();
  }
  std::vector<double4> h(4 * n);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), dout, 4 * n * sizeof(double4), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (di) (void)hipFreeAsync(di, s);
  if (dout) (void)hipFreeAsync(dout, s);
  if (e != hipSuccess) return fail(c, HMC_ERR_DEVICE, std::string("gather bodies: ") + hipGetErrorString(e));
  for (uint64_t k = 0; k < n; ++k) {
    const double4 b = h[4 * k], st = h[4 * k + 1], bq = h[4 * k + 2], sq = h[4 * k + 3];
    const double bv[3] = {b.x, b.y, b.z}, sv[3] = {st.x, st.y, st.z};
    for (int a = 0; a < D; ++a) {
      if (body_pos) body_pos[k * D + a] = bv[a];
      if (site_pos) site_pos[k * D + a] = sv[a];
    }
    if (body_quat) {
      body_quat[4 * k] = bq.x;
      body_quat[4 * k + 1] = bq.y;
      body_quat[4 * k + 2] = bq.z;
      body_quat[4 * k + 3] = bq.w;
    }
    if (site_quat) {
      site_quat[4 * k] = sq.x;
      site_quat[4 * k + 1] = sq.y;
      site_quat[4 * k + 2] = sq.z;
      site_quat[4 * k + 3] = sq.w;
    }
  }
  return HMC_OK;
}

int hmc_trace_enable(hmc_ctx* c, int on, uint64_t capacity) {
  if (!c) return HMC_ERR_INVALID;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graph(c);
  if (c->w.trace) {
    HIPCHK(c, hipFree(c->w.trace));
    c->w.trace = nullptr;
  }
  c->w.tracing = 0;
  c->w.trace_cap = 0;
  if (on) {
    HIPCHK(c, dalloc(&c->w.trace, capacity));
    c->w.trace_cap = capacity;
    c->w.tracing = 1;
  }
  unsigned long long zero = 0;
  HIPCHK(c, hipMemcpy((char*)c->st + offsetof(ApplyState, trace_n), &zero, sizeof zero, hipMemcpyHostToDevice));
  return HMC_OK;
}

int hmc_trace_read(hmc_ctx* c, hmc_trace_record* buf, uint64_t cap, uint64_t* n) {
  if (!c || !n) return HMC_ERR_INVALID;
  ApplyState s;
  int rc = read_state(c, &s);
  if (rc) return rc;
  uint64_t have = s.trace_n;
  if (have > c->w.trace_cap) return fail(c, HMC_ERR_STATE, "trace buffer overflow");
  *n = have;
  if (have > cap) return fail(c, HMC_ERR_STATE, "caller buffer too small");
  if (have) HIPCHK(c, hipMemcpy(buf, c->w.trace, have * sizeof(hmc_trace_record), hipMemcpyDeviceToHost));
  unsigned long long zero = 0;
  HIPCHK(c, hipMemcpy((char*)c->st + offsetof(ApplyState, trace_n), &zero, sizeof zero, hipMemcpyHostToDevice));
  return HMC_OK;
}

int hmc_get_stats(hmc_ctx* c, hmc_stats* out) {
  if (!c || !out) return HMC_ERR_INVALID;
  ApplyState s;
  int rc = read_state(c, &s, false);  // (the stats also describe a failed apply)
  if (rc) return rc;
  out->applies = c->applies;
  out->cycles = s.cycles_total;
  out->pair_tests = 0;
  out->band_retests = s.band_retests;
  out->guard_cycles = s.guard_cycles;
  out->tile_overflows = s.tile_overflows;
  out->ambiguous = s.ambiguous;
  std::memcpy(&out->fp32_error_max, &s.fp32_err_max, sizeof(double));
  std::memcpy(&out->fp32_error_bound, &s.fp32_err_ratio, sizeof(double));
  out->tiled = c->persist ? 2u : (c->tiled ? 1u : 0u);
  for (int a = 0; a < 3; ++a) out->tile_shape[a] = c->persist ? c->pc.T[a] : (c->tiled ? c->tc.T[a] : 0u);
  out->regions = c->persist ? c->pc.ntiles : 0u;
  out->region_threads = c->persist ? c->pthreads : 0u;
  out->regions_per_cu = c->persist ? c->pper_cu : 0u;
  out->wait_timeout_at = s.wait_at;
  return HMC_OK;
}

int hmc_debug_regions(hmc_ctx* c, uint32_t* out, uint64_t cap, uint64_t* n) {
  if (!c || !out || !n) return HMC_ERR_INVALID;
  *n = 0;
  if (!c->persist || !c->w.mbox) return HMC_OK;
  const uint64_t words = (uint64_t)c->pc.ntiles * kMailboxStride;
  std::vector<uint32_t> h(words);
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(h.data(), c->w.mbox, words * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const uint64_t m = std::min<uint64_t>(cap / 6, c->pc.ntiles);
  for (uint64_t r = 0; r < m; ++r)
    for (int k = 0; k < 6; ++k) out[6 * r + k] = h[r * kMailboxStride + k];
  *n = m;
  return HMC_OK;
}

int hmc_profile_enable(hmc_ctx* c, int on) {
  if (!c) return HMC_ERR_INVALID;
  c->profiling = on > 0;
  c->profile_period = on > 0 ? (uint32_t)on : 1u;
  c->pass_ms = c->build_ms = 0.0;
  c->pass_n = c->build_n = 0;
  return HMC_OK;
}

int hmc_profile_read(hmc_ctx* c, double* pass_ms_total, uint64_t* pass_launches, double* build_ms_total,
                     uint64_t* builds) {
  if (!c) return HMC_ERR_INVALID;
  ApplyState s;
  int rc = read_state(c, &s);
  if (rc) return rc;
  if (pass_ms_total) *pass_ms_total = c->pass_ms;
  if (pass_launches) *pass_launches = c->pass_n;
  if (build_ms_total) *build_ms_total = c->build_ms;
  if (builds) *builds = c->build_n;
  return HMC_OK;
}

int hmc_debug_math(int device, int kind, const double* in, double* out, uint64_t n) {
  if (hipSetDevice(device) != hipSuccess) return HMC_ERR_DEVICE;
  if (n == 0) return HMC_OK;
  uint64_t nin = kind == 4 ? 2 * n : n;
  double *din = nullptr, *dout = nullptr;
  if (dalloc(&din, nin) != hipSuccess || dalloc(&dout, n) != hipSuccess) {
    (void)hipFree(din);
    return HMC_ERR_DEVICE;
  }
  int rc = HMC_OK;
  if (hipMemcpy(din, in, nin * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) rc = HMC_ERR_DEVICE;
  if (!rc) {
    k_debug_math<<<(uint32_t)((n + 255) / 256), 256>>>(kind, din, dout, n);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = HMC_ERR_DEVICE;
  }
  if (!rc && hipMemcpy(out, dout, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) rc = HMC_ERR_DEVICE;
  (void)hipFree(din);
  (void)hipFree(dout);
  return rc;
}

int hmc_debug_normals(int device, uint64_t step, uint32_t substep, uint32_t seed, uint64_t index, double* out,
                      uint64_t n) {
  if (hipSetDevice(device) != hipSuccess) return HMC_ERR_DEVICE;
  if (n == 0) return HMC_OK;
  double* d = nullptr;
  if (dalloc(&d, n) != hipSuccess) return HMC_ERR_DEVICE;
  int rc = HMC_OK;
  k_debug_normals<<<1, 64>>>(step, substep, seed, index, d, n);
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = HMC_ERR_DEVICE;
  if (!rc && hipMemcpy(out, d, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) rc = HMC_ERR_DEVICE;
  (void)hipFree(d);
  return rc;
}

const char* hmc_last_error(const hmc_ctx* c) { return c ? c->err.c_str() : "null context"; }

void hmc_destroy(hmc_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  for (void* p : {(void*)c->bodies.bpos, (void*)c->bodies.spos, (void*)c->bodies.bquat, (void*)c->bodies.squat,
                  (void*)c->cells.body, (void*)c->cells.pos32, (void*)c->cells.quat64, (void*)c->verts,
                  (void*)c->gdev,
                  (void*)c->cells.pos64, (void*)c->w.cnt, (void*)c->w.start, (void*)c->w.elem_space,
                  (void*)c->w.elem_rank, (void*)c->w.tmp_body, (void*)c->w.block_sums, (void*)c->w.block_ne,
                  (void*)c->w.acc_slots, (void*)c->w.trace, (void*)c->w.mbox, (void*)c->w.scan_status,
                  (void*)c->st})
    if (p) (void)hipFree(p);
  collect_events(c, c->pass_events, c->pass_ms, c->pass_n);
  collect_events(c, c->build_events, c->build_ms, c->build_n);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

}  // extern "C"
