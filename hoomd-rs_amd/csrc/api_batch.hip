// The C ABI of a batch of replicas: creation, state, hipGraph-captured
// steps, SDF and the tile autotuner.
#include "host.hpp"
#include <algorithm>

namespace {
bool batch_regions(hmc_batch* b, int cus);

int bfail(hmc_batch* b, int code, const std::string& msg) {
  if (b) b->err = msg;
  return code;
}

#define BHIPCHK(b, expr)                                                                     \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return bfail((b), HMC_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

// The launches of one step of every replica (ParallelSweep::apply with
// Translate, then increment_step when advance is set).
template <int D, int PREC>
void launch_batch_step(hmc_batch* b, int from_args, int advance, uint32_t launched) {
  hipStream_t s = b->stream;
  k_bsetup<<<grid_for(b->R, 64), 64, 0, s>>>(b->reps, b->sts, b->R, b->dstep, b->dseed, b->kT, b->dmove, from_args,
                                             advance);
  k_fill_u32<<<grid_for(b->s_total, 256), 256, 0, s>>>(b->w.cnt, b->s_total, 0u);
  k_bbin<D><<<grid_for(b->n_total, kBlock), kBlock, 0, s>>>(b->reps, b->sts, b->dboff, b->R, b->n_total, b->bodies,
                                                            b->w);
  k_scan_blocks<<<b->n_scan_blocks, 256, 0, s>>>(b->w.cnt, b->w.start, b->w.block_sums, b->w.block_ne, b->s_total);
  k_scan_top<<<1, 1024, 0, s>>>(b->gtot, b->sts + b->R, b->w.block_sums, b->w.block_ne, b->n_scan_blocks, b->w.start,
                                0);
  k_scan_add<<<grid_for(b->s_total, 256), 256, 0, s>>>(b->w.start, b->w.block_sums, nullptr, b->s_total);
  k_bcycles<<<grid_for(b->R, 64), 64, 0, s>>>(b->reps, b->sts, b->R);
  k_scatter<<<grid_for(b->n_total, kBlock), kBlock, 0, s>>>(b->gtot, b->w);
  if constexpr (D == 3 && PREC == HMC_MIXED) {
    if (b->persist) {
      // the persistent applies stage their sites from the FP64 copies: one
      // cooperative launch runs every cycle of every replica, round by round
      k_place<PREC, false><<<grid_for(b->n_total, kBlock), kBlock, 0, s>>>(b->gtot, b->bodies, b->cells, b->w);
      uint32_t nrounds = (uint32_t)b->grp_tile0.size(), nwg = b->nwg, chunk = (nwg + 7) / 8, probe = 0;
      void* args[] = {(void*)&b->reps, (void*)&b->dptcs, (void*)&b->dptoff, (void*)&b->dptrep, (void*)&b->sts,
                      (void*)&b->drt0, (void*)&b->drn,   (void*)&nrounds,   (void*)&nwg,       (void*)&chunk,
                      (void*)&probe,   (void*)&b->bodies, (void*)&b->cells, (void*)&b->w};
      const hipError_t e = hipLaunchCooperativeKernel((const void*)k_bapply_tile<3>, dim3(chunk * 8),
                                                      dim3(b->pthreads), args, b->plds, s);
      if (e != hipSuccess && b->launch_err == hipSuccess) b->launch_err = e;
      return;
    }
  }
  k_place<PREC><<<grid_for(b->n_total, kBlock), kBlock, 0, s>>>(b->gtot, b->bodies, b->cells, b->w);
  const uint32_t blocks = grid_for(b->pc_total, kBlock);
  const uint32_t chunk = (blocks + 7) / 8;
  const uint32_t tchunk = (b->ntiles + 7) / 8;
  for (uint32_t k = 0; k < launched; ++k)
    for (int col = 0; col < (1 << D); ++col) {
      if (b->tiled)
        k_bpass_tile<D, PREC><<<tchunk * 8, b->tile_threads, b->tile_lds, s>>>(
            b->reps, b->dtcs, b->dtoff, b->dtrep, b->sts, b->ntiles, tchunk, b->bodies, b->cells, b->w,
            (uint32_t)col, k);
      else if (b->fast)
        k_bpass<D, PREC, true><<<chunk * 8, kBlock, 0, s>>>(b->reps, b->sts, b->dcoff, b->R, b->pc_total, b->bodies,
                                                            b->cells, b->w, (uint32_t)col, k, chunk);
      else
        k_bpass<D, PREC, false><<<chunk * 8, kBlock, 0, s>>>(b->reps, b->sts, b->dcoff, b->R, b->pc_total,
                                                             b->bodies, b->cells, b->w, (uint32_t)col, k, chunk);
    }
  k_bguard<D, PREC><<<b->R, kGuardThreads, 0, s>>>(b->reps, b->sts, b->bodies, b->cells, b->w, launched);
}

void batch_step(hmc_batch* b, int from_args, int advance, uint32_t launched) {
  // regions sized for the states set since the last choice (never under
  // graph capture: persistent steps are not captured)
  if (b->persist && b->regions_stale) b->persist = batch_regions(b, b->pcus);
  if (b->dim == 2 && b->precision == HMC_FP64)
    launch_batch_step<2, HMC_FP64>(b, from_args, advance, launched);
  else if (b->dim == 2)
    launch_batch_step<2, HMC_MIXED>(b, from_args, advance, launched);
  else if (b->precision == HMC_FP64)
    launch_batch_step<3, HMC_FP64>(b, from_args, advance, launched);
  else
    launch_batch_step<3, HMC_MIXED>(b, from_args, advance, launched);
}

int batch_read(hmc_batch* b, std::vector<ApplyState>& st, std::vector<unsigned long long>& acc) {
  st.resize(b->R);
  acc.resize(b->R);
  BHIPCHK(b, hipMemcpyAsync(st.data(), b->sts, b->R * sizeof(ApplyState), hipMemcpyDeviceToHost, b->stream));
  BHIPCHK(b, hipMemcpyAsync(acc.data(), b->w.acc_slots, b->R * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                            b->stream));
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  for (uint32_t r = 0; r < b->R; ++r) {
    if (st[r].error & kErrBarrier)
      return bfail(b, HMC_ERR_DEVICE, "a neighbour wait of the persistent apply timed out (replica " +
                                          std::to_string(r) + ")");
    if (st[r].error) return bfail(b, HMC_ERR_COVER, "a body lies outside the checkerboard (replica " + std::to_string(r) + ")");
  }
  return HMC_OK;
}

// Tile shapes of every replica for the LDS-tiled batch pass: choose_tiles
// with `cus` compute units assumed per replica (1: the largest tiles that
// fit), one workgroup size for all (the row-start walk re-stepped for it)
// and each replica's FP32 band of the box-relative test.  False when a
// replica has no tile shape.
struct BatchTiles {
  std::vector<TileCfg> tcs;
  std::vector<uint32_t> toff;
  uint32_t threads = 64, lds = 0;
};

bool batch_tiles(const hmc_batch* b, int cus, BatchTiles& bt) {
  const uint32_t R = b->R;
  const int D = b->dim;
  const size_t vt = b->precision == HMC_MIXED ? sizeof(float4) : sizeof(double4);
  bt.tcs.assign(R, TileCfg{});
  bt.toff.assign(R + 1, 0);
  bt.threads = 64;
  bt.lds = 0;
  for (uint32_t r = 0; r < R; ++r) {
    const Geometry& g = b->hreps[r].g;
    TileCfg& tc = bt.tcs[r];
    if (!choose_tiles(g, D, (double)g.n_bodies / (double)g.n_spaces, vt, cus, &tc)) return false;
    if (b->precision == HMC_MIXED) {
      const double sigma = std::sqrt(g.sigma2);
      double lmax = 0.0, Rb = 0.0;
      for (int a = 0; a < D; ++a) {
        lmax = std::max(lmax, g.L[a]);
        Rb = std::max(Rb, 0.5 * tc.E[a] * g.width[a] + g.width[a]);
      }
      // as for a context's tiled pass
      const double Et = (0x1p-24 * (5.1 * lmax + 2.0 * Rb + 2.0 * sigma) + 0x1p-53 * (1.5 * lmax + 2.0 * sigma)) * 1.01;
      if (!contact_band(Et, sigma, g.sigma2, g.edge2, D, &tc.band_lo, &tc.band_hi)) return false;
    }
    bt.threads = std::max<uint32_t>(bt.threads, ((tc.trials + 63) / 64) * 64);
    bt.lds = std::max<uint32_t>(bt.lds, tc.lds_bytes);
    bt.toff[r + 1] = bt.toff[r] + tc.ntiles;
  }
  // the row-start walk steps by the batch's workgroup size
  for (uint32_t r = 0; r < R; ++r) {
    uint32_t q = bt.threads;
    for (int a = D - 1; a > 0; --a) {
      const uint32_t rad = a == D - 1 ? bt.tcs[r].E[a] + 2 : bt.tcs[r].E[a];
      bt.tcs[r].step[a] = q % rad;
      q /= rad;
    }
    bt.tcs[r].step[0] = q;
  }
  return true;
}

// Region shapes of the persistent applies of a mixed-precision 3D batch.
// For each workgroup size and LDS budget, every replica takes the cubic
// shape T (as a context's regions: 2T spaces per axis and a one-space halo
// in LDS, the persistent layout's FP32 band) with the fewest regions that
// fits, and the replicas are packed in order into rounds of at most as many
// regions as can be resident at once (occupancy of k_bapply_tile); the
// choice with the fewest rounds wins.  The launch is then probed on the
// device.  Returns false (and leaves the batch on its per-colour passes)
// when no shape fits or the probe fails.
// For each region shape T (index T = 2..6): the largest number of bodies in
// a box of E = 2T + 2 spaces per axis over every placement of the box on the
// periodic space grid (binning from the lower box corner); 0 where E exceeds
// the grid.  Sliding sums along each axis in turn.
std::array<uint32_t, 7> box_counts(const Geometry& g, const double* pos, uint32_t N) {
  std::array<uint32_t, 7> out{};
  const uint32_t S0 = (uint32_t)g.shape[0], S1 = (uint32_t)g.shape[1], S2 = (uint32_t)g.shape[2];
  const size_t S = (size_t)S0 * S1 * S2;
  if (!S || !pos) return out;
  std::vector<uint32_t> cnt(S, 0), a(S), c(S);
  for (uint32_t i = 0; i < N; ++i) {
    uint32_t m[3];
    for (int k = 0; k < 3; ++k) {
      const double q = std::floor((pos[(size_t)i * 3 + k] + 0.5 * g.L[k]) / g.width[k]);
      const int64_t v = q < 0.0 ? 0 : (int64_t)q;
      m[k] = (uint32_t)std::min<int64_t>(v, (int64_t)g.shape[k] - 1);
    }
    cnt[((size_t)m[0] * S1 + m[1]) * S2 + m[2]] += 1;
  }
  for (uint32_t T = 2; T <= 6; ++T) {
    const uint32_t E = 2 * T + 2;
    if (E > S0 || E > S1 || E > S2) continue;
    // axis 2, then 1, then 0: window sums of E consecutive spaces (wrapped)
    for (size_t r = 0; r < (size_t)S0 * S1; ++r)
      for (uint32_t z = 0; z < S2; ++z) {
        uint32_t t = 0;
        for (uint32_t k = 0; k < E; ++k) t += cnt[r * S2 + (z + k) % S2];
        a[r * S2 + z] = t;
      }
    for (uint32_t x = 0; x < S0; ++x)
      for (uint32_t y = 0; y < S1; ++y)
        for (uint32_t z = 0; z < S2; ++z) {
          uint32_t t = 0;
          for (uint32_t k = 0; k < E; ++k) t += a[((size_t)x * S1 + (y + k) % S1) * S2 + z];
          c[((size_t)x * S1 + y) * S2 + z] = t;
        }
    uint32_t best = 0;
    for (uint32_t x = 0; x < S0; ++x)
      for (uint32_t y = 0; y < S1; ++y)
        for (uint32_t z = 0; z < S2; ++z) {
          uint32_t t = 0;
          for (uint32_t k = 0; k < E; ++k) t += c[((size_t)((x + k) % S0) * S1 + y) * S2 + z];
          best = std::max(best, t);
        }
    out[T] = best;
  }
  return out;
}

bool batch_regions(hmc_batch* b, int cus) {
  const uint32_t R = b->R;
  b->pcus = cus;
  b->regions_stale = false;
  if (hipStreamSynchronize(b->stream) != hipSuccess) return false;
  for (void** p : {(void**)&b->dptcs, (void**)&b->dptoff, (void**)&b->dptrep, (void**)&b->w.mbox, (void**)&b->drt0,
                   (void**)&b->drn})
    if (*p) {
      (void)hipFree(*p);
      *p = nullptr;
    }
  constexpr int D = 3;
  const void* fn = (const void*)k_bapply_tile<3>;
  // the region shape T of replica r for a workgroup of `threads` and an LDS
  // budget `cap`: the fewest regions (then the least LDS); false if none fits
  auto shape_for = [&](uint32_t r, uint32_t threads, uint32_t cap, TileCfg& out) {
    const Geometry& g = b->hreps[r].g;
    bool found = false;
    for (uint32_t T = 2; T <= 6; ++T) {
      const uint32_t trials = T * T * T;
      if (trials > threads) continue;
      TileCfg tc;
      std::memset(&tc, 0, sizeof tc);
      bool ok = true;
      uint32_t ncell = 1, ntiles = 1;
      for (int a = 0; a < D; ++a) {
        if (g.shape[a] < 6 || g.reach[a] != 1 || T > g.half_shape[a] || T > (uint32_t)(g.shape[a] - 4) / 2) ok = false;
        tc.T[a] = T;
        tc.nt[a] = (g.half_shape[a] + T - 1) / T;
        tc.E[a] = 2 * T + 2;
        ncell *= 2 * T + 2;
        ntiles *= tc.nt[a];
      }
      const uint32_t EL = 2 * T + 2, rows = ncell / EL;
      if (rows > 64u * kRowsPerLane || rows * (EL + 2) > threads * kValsPerThread) ok = false;
      const double occ = (double)g.n_bodies / (double)g.n_spaces;
      tc.trials = trials;
      tc.ncell = ncell;
      tc.rows = rows;
      // the capacity of a region's box: 2% + 32 entries above the largest
      // box count of the state last set (every placement of the box); before
      // any state, 10% + 64 above the mean (a lattice's planes make box
      // counts vary by several percent).  A region whose box holds more takes
      // its trials from global memory (same results).
      const uint32_t known = r < b->box_max.size() ? b->box_max[r][T] : 0u;
      tc.maxb = known ? std::max((uint32_t)std::ceil(1.02 * known) + 32, (uint32_t)std::ceil(occ * ncell) + 32)
                      : (uint32_t)std::ceil(1.1 * occ * ncell) + 64;
      tc.ntiles = ntiles;
      tc.xcd_chunk = (ntiles + 7) / 8;
      const size_t lr = ((16 * ((size_t)rows + 1) + 2 * (size_t)rows * (EL + 1) + 15) & ~(size_t)15) +
                        (size_t)tc.maxb * sizeof(float4);
      if (lr > cap || lr > kTileLds) ok = false;
      tc.lds_bytes = (uint32_t)lr;
      uint32_t q = threads;
      for (int a = D - 1; a > 0; --a) {
        const uint32_t rad = a == D - 1 ? EL + 2 : tc.E[a];
        tc.step[a] = q % rad;
        q /= rad;
      }
      tc.step[0] = q;
      // the persistent layout's band, as for a context
      const double sigma = g.sigma;
      double lmax = 0.0, Rb = 0.0;
      for (int a = 0; a < D; ++a) {
        lmax = std::max(lmax, g.L[a]);
        Rb = std::max(Rb, 0.5 * tc.E[a] * g.width[a] + g.width[a]);
      }
      const double Et = (0x1p-24 * (2.0 * Rb + 2.0 * sigma) + 0x1p-53 * (4.0 * lmax + 2.0 * sigma)) * 1.01;
      if (!contact_band(Et, sigma, g.sigma2, g.edge2, D, &tc.band_lo, &tc.band_hi)) ok = false;
      if (ok && (!found || tc.ntiles < out.ntiles || (tc.ntiles == out.ntiles && tc.lds_bytes < out.lds_bytes))) {
        out = tc;
        found = true;
      }
    }
    return found;
  };
  size_t best_groups = 0;
  std::vector<TileCfg> best;
  std::vector<uint32_t> best_t0, best_nt, best_round;
  uint32_t best_threads = 0, best_lds = 0, best_per_cu = 0;
  for (uint32_t threads = 64; threads <= (uint32_t)kTileThreads; threads *= 2)
    for (uint32_t cap = 16 * 1024; cap <= (uint32_t)kTileLds; cap += 4 * 1024) {
      std::vector<TileCfg> tcs(R);
      bool ok = true;
      uint32_t lds = 0;
      for (uint32_t r = 0; r < R && ok; ++r) {
        ok = shape_for(r, threads, cap, tcs[r]);
        if (ok) lds = std::max<uint32_t>(lds, tcs[r].lds_bytes);
      }
      if (!ok) continue;
      int per_cu = 0;
      if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess ||
          hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, (int)threads, lds) != hipSuccess || per_cu <= 0)
        continue;
      const uint64_t capr = (uint64_t)per_cu * (uint64_t)cus;
      // whole replicas into rounds, first fit in decreasing order of region
      // count (7 workgroups of each round are left for the other launches)
      std::vector<uint32_t> order(R);
      for (uint32_t r = 0; r < R; ++r) order[r] = r;
      std::stable_sort(order.begin(), order.end(),
                       [&](uint32_t a, uint32_t c) { return tcs[a].ntiles > tcs[c].ntiles; });
      std::vector<uint32_t> load, round_of(R);
      for (uint32_t r : order) {
        const uint32_t n = tcs[r].ntiles;
        if ((uint64_t)n + 7 > capr) {
          ok = false;
          break;
        }
        size_t k = 0;
        while (k < load.size() && (uint64_t)load[k] + n + 7 > capr) ++k;
        if (k == load.size()) load.push_back(0);
        load[k] += n;
        round_of[r] = (uint32_t)k;
      }
      if (!ok) continue;
      if (best.empty() || load.size() < best_groups || (load.size() == best_groups && threads < best_threads)) {
        std::vector<uint32_t> t0(load.size(), 0);
        for (size_t k = 1; k < load.size(); ++k) t0[k] = t0[k - 1] + load[k - 1];
        best = tcs;
        best_groups = load.size();
        best_round = round_of;
        best_t0 = t0;
        best_nt = load;
        best_threads = threads;
        best_lds = lds;
        best_per_cu = (uint32_t)per_cu;
      }
    }
  if (best.empty()) return false;
  if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)best_lds) != hipSuccess) return false;
  // regions numbered round by round: replica r owns [toff[r], toff[r] +
  // its region count), inside its round's range [t0, t0 + nt)
  std::vector<uint32_t> toff(R + 1, 0), fill(best_t0);
  for (uint32_t r = 0; r < R; ++r) {
    toff[r] = fill[best_round[r]];
    fill[best_round[r]] += best[r].ntiles;
  }
  const uint32_t total = best_t0.back() + best_nt.back();
  toff[R] = total;
  std::vector<uint32_t> trep(total);
  for (uint32_t r = 0; r < R; ++r)
    for (uint32_t t = toff[r]; t < toff[r] + best[r].ntiles; ++t) trep[t] = r;
  const size_t fbytes = (size_t)total * kMailboxStride * sizeof(uint32_t);
  if (dalloc(&b->dptcs, R) != hipSuccess || dalloc(&b->dptoff, R + 1) != hipSuccess ||
      dalloc(&b->dptrep, total) != hipSuccess || hipMalloc((void**)&b->w.mbox, fbytes) != hipSuccess ||
      hipMemset(b->w.mbox, 0, fbytes) != hipSuccess ||
      hipMemcpy(b->dptcs, best.data(), R * sizeof(TileCfg), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->dptoff, toff.data(), (R + 1) * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->dptrep, trep.data(), total * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess)
    return false;
  b->pthreads = best_threads;
  b->plds = best_lds;
  b->pper_cu = best_per_cu;
  b->pregions = total;
  b->hptcs = best;
  b->grp_tile0 = best_t0;
  b->grp_ntiles = best_nt;
  b->nwg = 0;
  for (uint32_t n : best_nt) b->nwg = std::max(b->nwg, n);
  const uint32_t nr = (uint32_t)best_t0.size();
  if (dalloc(&b->drt0, nr) != hipSuccess || dalloc(&b->drn, nr) != hipSuccess ||
      hipMemcpy(b->drt0, best_t0.data(), nr * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->drn, best_nt.data(), nr * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess)
    return false;
  // every workgroup of the launch resident at once, checked with the kernel
  // itself (the replicas' states are kept across the probe)
  std::vector<ApplyState> keep(R + 1);
  if (hipStreamSynchronize(b->stream) != hipSuccess ||
      hipMemcpy(keep.data(), b->sts, (R + 1) * sizeof(ApplyState), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemset(b->sts, 0, (R + 1) * sizeof(ApplyState)) != hipSuccess)
    return false;
  {
    uint32_t nrounds = nr, nwg = b->nwg, chunk = (nwg + 7) / 8, probe = 1;
    void* args[] = {(void*)&b->reps, (void*)&b->dptcs, (void*)&b->dptoff, (void*)&b->dptrep, (void*)&b->sts,
                    (void*)&b->drt0, (void*)&b->drn,   (void*)&nrounds,   (void*)&nwg,       (void*)&chunk,
                    (void*)&probe,   (void*)&b->bodies, (void*)&b->cells, (void*)&b->w};
    if (hipLaunchCooperativeKernel(fn, dim3(chunk * 8), dim3(best_threads), args, best_lds, b->stream) !=
            hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
      return false;
  }
  std::vector<ApplyState> st(R);
  if (hipMemcpy(st.data(), b->sts, R * sizeof(ApplyState), hipMemcpyDeviceToHost) != hipSuccess) return false;
  bool resident = true;
  for (uint32_t r = 0; r < R; ++r) resident = resident && !(st[r].error & kErrBarrier);
  if (hipMemset(b->w.mbox, 0, fbytes) != hipSuccess ||
      hipMemcpy(b->sts, keep.data(), (R + 1) * sizeof(ApplyState), hipMemcpyHostToDevice) != hipSuccess)
    return false;
  return resident;
}

// Makes bt the batch's tile layout (the captured step graph is dropped).
int batch_set_tiles(hmc_batch* b, const BatchTiles& bt) {
  const uint32_t R = b->R;
  const void* fn = b->dim == 2 ? (b->precision == HMC_MIXED ? (const void*)k_bpass_tile<2, HMC_MIXED>
                                                             : (const void*)k_bpass_tile<2, HMC_FP64>)
                               : (b->precision == HMC_MIXED ? (const void*)k_bpass_tile<3, HMC_MIXED>
                                                             : (const void*)k_bpass_tile<3, HMC_FP64>);
  if (!b->dtcs && (dalloc(&b->dtcs, R) != hipSuccess || dalloc(&b->dtoff, R + 1) != hipSuccess))
    return HMC_ERR_DEVICE;
  const uint32_t nt = bt.toff[R];
  std::vector<uint32_t> trep(nt);
  for (uint32_t r = 0; r < R; ++r)
    for (uint32_t t = bt.toff[r]; t < bt.toff[r + 1]; ++t) trep[t] = r;
  if (nt > b->dtrep_cap) {
    if (hipStreamSynchronize(b->stream) != hipSuccess) return HMC_ERR_DEVICE;
    if (b->dtrep) (void)hipFree(b->dtrep);
    b->dtrep = nullptr;
    b->dtrep_cap = 0;
    if (dalloc(&b->dtrep, nt) != hipSuccess) return HMC_ERR_DEVICE;
    b->dtrep_cap = nt;
  }
  if (nt && hipMemcpyAsync(b->dtrep, trep.data(), nt * sizeof(uint32_t), hipMemcpyHostToDevice, b->stream) !=
                hipSuccess)
    return HMC_ERR_DEVICE;
  if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bt.lds) != hipSuccess ||
      hipMemcpyAsync(b->dtcs, bt.tcs.data(), R * sizeof(TileCfg), hipMemcpyHostToDevice, b->stream) != hipSuccess ||
      hipMemcpyAsync(b->dtoff, bt.toff.data(), (R + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, b->stream) !=
          hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess)
    return HMC_ERR_DEVICE;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->tiled = true;
  b->ntiles = bt.toff[R];
  b->tile_threads = bt.threads;
  b->tile_lds = bt.lds;
  return HMC_OK;
}

}  // namespace

extern "C" {

// ------------------------------------------------------------------ batches --
int hmc_batch_create(uint32_t n_replicas, const hmc_config* cfgs, hmc_batch** out) {
  if (!out || !cfgs || n_replicas < 1 || n_replicas > (uint32_t)kAccSlots) return HMC_ERR_INVALID;
  *out = nullptr;
  auto* b = new hmc_batch;
  b->R = n_replicas;
  b->device = cfgs[0].device;
  b->dim = (int)cfgs[0].dim;
  b->precision = cfgs[0].precision;
  b->hreps.resize(n_replicas);
  b->boff.assign(n_replicas + 1, 0);
  b->coff.assign(n_replicas + 1, 0);
  uint64_t nb = 0, ns = 0, npc = 0;
  for (uint32_t r = 0; r < n_replicas; ++r) {
    const hmc_config* cfg = cfgs + r;
    if (cfg->model != HMC_HARD_SPHERE || (int)cfg->dim != b->dim || cfg->precision != b->precision ||
        cfg->device != b->device) {
      delete b;
      return HMC_ERR_UNSUPPORTED;
    }
    bool fast = false;
    double sigma = 0.0, lmax = 0.0;
    BatchRep& rep = b->hreps[r];
    std::memset(&rep, 0, sizeof rep);
    const int rc = make_geometry(cfg, rep.g, &fast, &sigma, &lmax);
    if (rc) {
      delete b;
      return rc;
    }
    b->fast = b->fast && fast;
    rep.body_off = (uint32_t)nb;
    rep.space_off = (uint32_t)ns;
    rep.colour_off = (uint32_t)npc;
    b->boff[r] = (uint32_t)nb;
    b->coff[r] = (uint32_t)npc;
    nb += rep.g.n_bodies;
    ns += rep.g.n_spaces;
    npc += rep.g.per_colour;
  }
  if (nb >= 0x7fffffffull || ns >= 0x7fffffffull || ns > (uint64_t)kScanBlock * 1024 * 16) {
    delete b;
    return HMC_ERR_UNSUPPORTED;
  }
  b->boff[n_replicas] = (uint32_t)nb;
  b->coff[n_replicas] = (uint32_t)npc;
  b->n_total = (uint32_t)nb;
  b->s_total = (uint32_t)ns;
  b->pc_total = (uint32_t)npc;
  std::memset(&b->gtot, 0, sizeof b->gtot);
  b->gtot.dim = b->dim;
  b->gtot.n_bodies = b->n_total;
  b->gtot.n_spaces = b->s_total;
  if (hipSetDevice(b->device) != hipSuccess) {
    delete b;
    return HMC_ERR_DEVICE;
  }
  const uint32_t N = b->n_total, R = b->R;
  b->n_scan_blocks = grid_for(b->s_total, kScanBlock);
  bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess;
  ok = ok && dalloc(&b->reps, R) == hipSuccess && dalloc(&b->sts, R + 1) == hipSuccess &&
       dalloc(&b->dboff, R + 1) == hipSuccess && dalloc(&b->dcoff, R + 1) == hipSuccess &&
       dalloc(&b->dstep, R) == hipSuccess && dalloc(&b->dseed, R) == hipSuccess && dalloc(&b->dmove, R) == hipSuccess;
  ok = ok && dalloc(&b->bodies.bpos, N) == hipSuccess && dalloc(&b->bodies.spos, N) == hipSuccess &&
       dalloc(&b->cells.body, N) == hipSuccess;
  if (b->precision == HMC_MIXED)
    ok = ok && dalloc(&b->cells.pos32, N) == hipSuccess;
  else
    ok = ok && dalloc(&b->cells.pos64, N) == hipSuccess;
  ok = ok && dalloc(&b->w.cnt, b->s_total) == hipSuccess && dalloc(&b->w.start, b->s_total + 1) == hipSuccess &&
       dalloc(&b->w.elem_space, N) == hipSuccess && dalloc(&b->w.elem_rank, N) == hipSuccess &&
       dalloc(&b->w.tmp_body, N) == hipSuccess && dalloc(&b->w.block_sums, b->n_scan_blocks) == hipSuccess &&
       dalloc(&b->w.block_ne, b->n_scan_blocks) == hipSuccess && dalloc(&b->w.acc_slots, kAccSlots) == hipSuccess;
  if (ok)
    for (uint32_t r = 0; r < R; ++r) b->hreps[r].g.self = &b->reps[r].g;
  ok = ok && hipMemcpy(b->reps, b->hreps.data(), R * sizeof(BatchRep), hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(b->dboff, b->boff.data(), (R + 1) * sizeof(uint32_t), hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(b->dcoff, b->coff.data(), (R + 1) * sizeof(uint32_t), hipMemcpyHostToDevice) == hipSuccess &&
       hipMemset(b->sts, 0, (R + 1) * sizeof(ApplyState)) == hipSuccess &&
       hipMemset(b->w.acc_slots, 0, kAccSlots * sizeof(unsigned long long)) == hipSuccess;
  if (!ok) {
    hmc_batch_destroy(b);
    return HMC_ERR_DEVICE;
  }
  b->cycles_guess = 1;
  for (uint32_t r = 0; r < R; ++r) {
    const Geometry& g = b->hreps[r].g;
    b->cycles_guess = std::max<uint32_t>(b->cycles_guess, (g.n_bodies + g.n_spaces - 1) / std::max<uint32_t>(g.n_spaces, 1));
  }
  // LDS tiles for every replica (else the lane-per-space pass): the largest
  // tiles that fit (the batch fills the device by itself); hmc_batch_autotune
  // may pick another shape later
  if (b->fast && !(cfgs[0].flags & HMC_FLAG_GLOBAL_PASS)) {
    BatchTiles bt;
    if (batch_tiles(b, 1, bt) && batch_set_tiles(b, bt) != HMC_OK) {
      hmc_batch_destroy(b);
      return HMC_ERR_DEVICE;
    }
    // mixed precision in 3D: persistent applies, one cooperative launch per
    // group of replicas per step (HMC_FLAG_PER_COLOUR keeps the per-colour
    // tile passes)
    int cus = 0, coop = 0;
    if (b->tiled && b->dim == 3 && b->precision == HMC_MIXED && !(cfgs[0].flags & HMC_FLAG_PER_COLOUR) &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device) == hipSuccess && cus > 0 &&
        hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, b->device) == hipSuccess && coop)
      b->persist = batch_regions(b, cus);
  }
  *out = b;
  return HMC_OK;
}

int hmc_batch_set_state(hmc_batch* b, uint32_t replica, const double* body_pos, const double* site_pos) {
  if (!b || replica >= b->R) return HMC_ERR_INVALID;
  const Geometry& g = b->hreps[replica].g;
  const uint32_t N = g.n_bodies, off = b->boff[replica];
  const int D = b->dim;
  if (!body_pos && N) return HMC_ERR_INVALID;
  std::vector<double4> hb(N), hs(site_pos ? N : 0);
  for (uint32_t i = 0; i < N; ++i) {
    double v[3] = {0, 0, 0}, u[3] = {0, 0, 0};
    for (int a = 0; a < D; ++a) v[a] = body_pos[(size_t)i * D + a];
    hb[i] = make_double4(v[0], v[1], v[2], 0.0);
    if (site_pos) {
      for (int a = 0; a < D; ++a) u[a] = site_pos[(size_t)i * D + a];
      hs[i] = make_double4(u[0], u[1], u[2], 0.0);
    }
  }
  BHIPCHK(b, hipSetDevice(b->device));
  if (N) {
    BHIPCHK(b, hipMemcpyAsync(b->bodies.bpos + off, hb.data(), N * sizeof(double4), hipMemcpyHostToDevice, b->stream));
    if (site_pos)
      BHIPCHK(b, hipMemcpyAsync(b->bodies.spos + off, hs.data(), N * sizeof(double4), hipMemcpyHostToDevice,
                                b->stream));
    Bodies sub{b->bodies.bpos + off, b->bodies.spos + off, nullptr, nullptr};
    if (D == 2)
      k_init_sites<2><<<grid_for(N, 256), 256, 0, b->stream>>>(g, sub, site_pos ? 1 : 0);
    else
      k_init_sites<3><<<grid_for(N, 256), 256, 0, b->stream>>>(g, sub, site_pos ? 1 : 0);
    BHIPCHK(b, hipGetLastError());
  }
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  if (b->persist) {
    if (b->box_max.size() != b->R) b->box_max.assign(b->R, std::array<uint32_t, 7>{});
    b->box_max[replica] = box_counts(g, body_pos, N);
    b->regions_stale = true;
  }
  return HMC_OK;
}

int hmc_batch_get_state(hmc_batch* b, uint32_t replica, double* body_pos, double* site_pos) {
  if (!b || replica >= b->R) return HMC_ERR_INVALID;
  const uint32_t N = b->hreps[replica].g.n_bodies, off = b->boff[replica];
  const int D = b->dim;
  std::vector<double4> hb(N), hs(N);
  BHIPCHK(b, hipSetDevice(b->device));
  if (N) {
    BHIPCHK(b, hipMemcpyAsync(hb.data(), b->bodies.bpos + off, N * sizeof(double4), hipMemcpyDeviceToHost, b->stream));
    BHIPCHK(b, hipMemcpyAsync(hs.data(), b->bodies.spos + off, N * sizeof(double4), hipMemcpyDeviceToHost, b->stream));
  }
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  for (uint32_t i = 0; i < N; ++i) {
    const double bv[3] = {hb[i].x, hb[i].y, hb[i].z}, sv[3] = {hs[i].x, hs[i].y, hs[i].z};
    for (int a = 0; a < D; ++a) {
      if (body_pos) body_pos[(size_t)i * D + a] = bv[a];
      if (site_pos) site_pos[(size_t)i * D + a] = sv[a];
    }
  }
  return HMC_OK;
}

int hmc_batch_run_translate(hmc_batch* b, const uint64_t* step0, const uint32_t* seeds, double kT,
                            const double* max_distance, uint32_t n_steps, int use_graph, hmc_count* out) {
  if (!b || !step0 || !seeds || !max_distance || !(kT > 0.0)) return HMC_ERR_INVALID;
  for (uint32_t r = 0; r < b->R; ++r)
    if (!(max_distance[r] > 0.0)) return bfail(b, HMC_ERR_INVALID, "max_distance must be > 0");
  BHIPCHK(b, hipSetDevice(b->device));
  b->kT = kT;
  BHIPCHK(b, hipMemcpyAsync(b->dstep, step0, b->R * sizeof(uint64_t), hipMemcpyHostToDevice, b->stream));
  BHIPCHK(b, hipMemcpyAsync(b->dseed, seeds, b->R * sizeof(uint32_t), hipMemcpyHostToDevice, b->stream));
  BHIPCHK(b, hipMemcpyAsync(b->dmove, max_distance, b->R * sizeof(double), hipMemcpyHostToDevice, b->stream));
  std::vector<ApplyState> st0, st1;
  std::vector<unsigned long long> a0, a1;
  int rc = batch_read(b, st0, a0);
  if (rc) return rc;
  const uint32_t launched = b->cycles_guess;
  for (uint32_t i = 0; i < n_steps; ++i) {
    // (cooperative launches are not captured: persistent steps launch directly)
    if (i == 0 || !use_graph || b->persist) {
      batch_step(b, i == 0 ? 1 : 0, i == 0 ? 0 : 1, launched);
      continue;
    }
    if (!b->exec || b->graph_cycles != launched) {
      if (b->exec) (void)hipGraphExecDestroy(b->exec);
      if (b->graph) (void)hipGraphDestroy(b->graph);
      b->exec = nullptr;
      b->graph = nullptr;
      BHIPCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      batch_step(b, 0, 1, launched);
      BHIPCHK(b, hipStreamEndCapture(b->stream, &b->graph));
      BHIPCHK(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->graph_cycles = launched;
    }
    BHIPCHK(b, hipGraphLaunch(b->exec, b->stream));
  }
  BHIPCHK(b, hipGetLastError());
  if (b->launch_err != hipSuccess) {
    const hipError_t e = b->launch_err;
    b->launch_err = hipSuccess;
    return bfail(b, HMC_ERR_DEVICE, std::string("cooperative launch: ") + hipGetErrorString(e));
  }
  rc = batch_read(b, st1, a1);
  if (rc) return rc;
  uint32_t most = 1;
  for (uint32_t r = 0; r < b->R; ++r) {
    most = std::max<uint32_t>(most, st1[r].cycles);
    if (out) {
      out[r].accepted = a1[r] - a0[r];
      out[r].rejected = (st1[r].trials_total - st0[r].trials_total) - out[r].accepted;
    }
  }
  if (n_steps) b->cycles_guess = most;
  return HMC_OK;
}

int hmc_batch_sdf(hmc_batch* b, double xmax, double dx, uint32_t nbins, uint64_t* hist, uint64_t* overlaps) {
  if (!b || !hist) return HMC_ERR_INVALID;
  for (uint32_t r = 0; r < b->R; ++r)
    if (const char* why = sdf_check(b->hreps[r].g, xmax, dx, nbins)) return bfail(b, HMC_ERR_INVALID, why);
  if ((uint64_t)b->R * (nbins + 1ull) >= (1ull << 31)) return bfail(b, HMC_ERR_INVALID, "n_replicas * nbins too large");
  BHIPCHK(b, hipSetDevice(b->device));
  hipStream_t s = b->stream;
  // the replicas laid out on the SDF's checkerboards (fewer, wider spaces)
  std::vector<BatchRep> hr(b->R);
  uint32_t stot = 0;
  for (uint32_t r = 0; r < b->R; ++r) {
    hr[r] = b->hreps[r];
    hr[r].g = sdf_geometry(b->hreps[r].g, xmax);
    hr[r].space_off = stot;
    stot += hr[r].g.n_spaces;
  }
  Geometry gt = b->gtot;
  gt.n_spaces = stot;
  const uint32_t nsb = grid_for(stot, kScanBlock);
  const size_t cells = (size_t)b->R * nbins + b->R;
  char* buf = nullptr;
  BHIPCHK(b, hipMallocAsync((void**)&buf, b->R * sizeof(BatchRep) + cells * sizeof(uint32_t), s));
  BatchRep* dreps = (BatchRep*)buf;
  uint32_t* dh = (uint32_t*)(buf + b->R * sizeof(BatchRep));
  hipError_t e = hipMemcpyAsync(dreps, hr.data(), b->R * sizeof(BatchRep), hipMemcpyHostToDevice, s);
  k_fill_u32<<<grid_for((uint32_t)cells, 256), 256, 0, s>>>(dh, (uint32_t)cells, 0u);
  if (e == hipSuccess && b->n_total) {
    // cells of the current states (origins drawn from the current counters;
    // the next step rebuilds the sweep's own)
    k_bsetup<<<grid_for(b->R, 64), 64, 0, s>>>(dreps, b->sts, b->R, b->dstep, b->dseed, b->kT, b->dmove, 0, 0);
    k_fill_u32<<<grid_for(stot, 256), 256, 0, s>>>(b->w.cnt, stot, 0u);
    if (b->dim == 2)
      k_bbin<2><<<grid_for(b->n_total, kBlock), kBlock, 0, s>>>(dreps, b->sts, b->dboff, b->R, b->n_total,
                                                                b->bodies, b->w);
    else
      k_bbin<3><<<grid_for(b->n_total, kBlock), kBlock, 0, s>>>(dreps, b->sts, b->dboff, b->R, b->n_total,
                                                                b->bodies, b->w);
    k_scan_blocks<<<nsb, 256, 0, s>>>(b->w.cnt, b->w.start, b->w.block_sums, b->w.block_ne, stot);
    k_scan_top<<<1, 1024, 0, s>>>(gt, b->sts + b->R, b->w.block_sums, b->w.block_ne, nsb, b->w.start, 0);
    k_scan_add<<<grid_for(stot, 256), 256, 0, s>>>(b->w.start, b->w.block_sums, nullptr, stot);
    k_scatter<<<grid_for(b->n_total, kBlock), kBlock, 0, s>>>(b->gtot, b->w);
    if (b->dim == 2)
      k_sdf<2><<<grid_for(b->n_total, kBlock), kBlock, 0, s>>>(dreps, b->dboff, b->R, b->n_total, b->bodies, b->w,
                                                               xmax, dx, nbins, dh, dh + (size_t)b->R * nbins);
    else
      k_sdf<3><<<grid_for(b->n_total, kBlock), kBlock, 0, s>>>(dreps, b->dboff, b->R, b->n_total, b->bodies, b->w,
                                                               xmax, dx, nbins, dh, dh + (size_t)b->R * nbins);
  }
  std::vector<uint32_t> h(cells, 0u);
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), dh, cells * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFreeAsync(buf, s);
  if (e != hipSuccess) return bfail(b, HMC_ERR_DEVICE, std::string("batch sdf: ") + hipGetErrorString(e));
  std::vector<ApplyState> st;
  std::vector<unsigned long long> acc;
  int rc = batch_read(b, st, acc);  // a body outside its checkerboard is reported here
  if (rc) return rc;
  for (size_t k = 0; k < (size_t)b->R * nbins; ++k) hist[k] = h[k];
  if (overlaps)
    for (uint32_t r = 0; r < b->R; ++r) overlaps[r] = h[(size_t)b->R * nbins + r];
  return HMC_OK;
}

int hmc_batch_autotune(hmc_batch* b, const uint64_t* step0, const uint32_t* seeds, double kT,
                       const double* max_distance, uint32_t n_steps, double* best_ms, uint32_t* best_policy) {
  if (!b || !step0 || !seeds || !max_distance || !(kT > 0.0) || n_steps == 0) return HMC_ERR_INVALID;
  for (uint32_t r = 0; r < b->R; ++r)
    if (!(max_distance[r] > 0.0)) return bfail(b, HMC_ERR_INVALID, "max_distance must be > 0");
  if (!b->tiled) return bfail(b, HMC_ERR_UNSUPPORTED, "the batch does not run on LDS tiles");
  BHIPCHK(b, hipSetDevice(b->device));
  hipStream_t s = b->stream;
  const uint32_t R = b->R;
  b->kT = kT;
  BHIPCHK(b, hipMemcpyAsync(b->dstep, step0, R * sizeof(uint64_t), hipMemcpyHostToDevice, s));
  BHIPCHK(b, hipMemcpyAsync(b->dseed, seeds, R * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  BHIPCHK(b, hipMemcpyAsync(b->dmove, max_distance, R * sizeof(double), hipMemcpyHostToDevice, s));
  // the state every trial run starts from and leaves behind
  const size_t nb = (size_t)b->n_total * sizeof(double4), ns = (R + 1) * sizeof(ApplyState),
               na = kAccSlots * sizeof(unsigned long long);
  char* save = nullptr;
  BHIPCHK(b, hipMalloc((void**)&save, 2 * nb + ns + na));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = HMC_OK;
  auto copy_state = [&](bool restore) -> bool {
    void* dst[4] = {save, save + nb, save + 2 * nb, save + 2 * nb + ns};
    void* src[4] = {b->bodies.bpos, b->bodies.spos, b->sts, b->w.acc_slots};
    const size_t len[4] = {nb, nb, ns, na};
    for (int i = 0; i < 4; ++i)
      if (len[i] && hipMemcpyAsync(restore ? src[i] : dst[i], restore ? dst[i] : src[i], len[i],
                                   hipMemcpyDeviceToDevice, s) != hipSuccess)
        return false;
    return hipStreamSynchronize(s) == hipSuccess;
  };
  const int policies[] = {1, 4, 16, 64, 256};
  double best = std::numeric_limits<double>::infinity();
  int chosen = -1;
  BatchTiles best_bt;
  std::vector<std::vector<uint32_t>> seen;
  if (!copy_state(false) || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = HMC_ERR_DEVICE;
  for (int pi = 0; pi < 5 && rc == HMC_OK; ++pi) {
    BatchTiles bt;
    if (!batch_tiles(b, policies[pi], bt)) continue;
    std::vector<uint32_t> sig;
    for (const TileCfg& tc : bt.tcs) sig.insert(sig.end(), {tc.T[0], tc.T[1], tc.T[2]});
    if (std::find(seen.begin(), seen.end(), sig) != seen.end()) continue;  // same layout as an earlier policy
    seen.push_back(sig);
    if (batch_set_tiles(b, bt) != HMC_OK) {
      rc = HMC_ERR_DEVICE;
      break;
    }
    // one untimed step from (step0, seeds), then n_steps timed ones
    batch_step(b, 1, 0, b->cycles_guess);
    float ms = 0.0f;
    bool ok = hipEventRecord(e0, s) == hipSuccess;
    for (uint32_t i = 0; i < n_steps; ++i) batch_step(b, 0, 1, b->cycles_guess);
    ok = ok && hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
         hipEventElapsedTime(&ms, e0, e1) == hipSuccess && hipGetLastError() == hipSuccess;
    ok = ok && copy_state(true);
    if (!ok) {
      rc = HMC_ERR_DEVICE;
      break;
    }
    if (ms / n_steps < best) {
      best = ms / n_steps;
      chosen = policies[pi];
      best_bt = bt;
    }
  }
  if (rc == HMC_OK && chosen > 0 && batch_set_tiles(b, best_bt) != HMC_OK) rc = HMC_ERR_DEVICE;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(save);
  if (rc) return bfail(b, rc, "autotune: device error");
  if (best_ms) *best_ms = best;
  if (best_policy) *best_policy = chosen > 0 ? (uint32_t)chosen : 0u;
  return HMC_OK;
}

int hmc_batch_get_stats(hmc_batch* b, uint32_t replica, hmc_stats* out) {
  if (!b || !out || replica >= b->R) return HMC_ERR_INVALID;
  std::vector<ApplyState> st;
  std::vector<unsigned long long> acc;
  const int rc = batch_read(b, st, acc);
  if (rc) return rc;
  const ApplyState& s = st[replica];
  std::memset(out, 0, sizeof *out);
  out->cycles = s.cycles_total;
  out->band_retests = s.band_retests;
  out->guard_cycles = s.guard_cycles;
  out->tile_overflows = s.tile_overflows;
  out->tiled = b->persist ? 2u : (b->tiled ? 1u : 0u);
  if (b->persist) {
    for (int a = 0; a < 3; ++a) out->tile_shape[a] = b->hptcs[replica].T[a];
    out->regions = b->pregions;
    out->region_threads = b->pthreads;
    out->regions_per_cu = b->pper_cu;
    out->rounds = (uint32_t)b->grp_tile0.size();
  }
  return HMC_OK;
}

const char* hmc_batch_last_error(const hmc_batch* b) { return b ? b->err.c_str() : "null batch"; }

void hmc_batch_destroy(hmc_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (void* p : {(void*)b->reps, (void*)b->sts, (void*)b->dboff, (void*)b->dcoff, (void*)b->dstep, (void*)b->dseed,
                  (void*)b->dmove, (void*)b->bodies.bpos, (void*)b->bodies.spos, (void*)b->cells.body,
                  (void*)b->cells.pos32, (void*)b->cells.pos64, (void*)b->w.cnt, (void*)b->w.start,
                  (void*)b->w.elem_space, (void*)b->w.elem_rank, (void*)b->w.tmp_body, (void*)b->w.block_sums,
                  (void*)b->w.block_ne, (void*)b->w.acc_slots, (void*)b->dtcs, (void*)b->dtoff, (void*)b->dtrep,
                  (void*)b->dptcs, (void*)b->dptoff, (void*)b->dptrep, (void*)b->w.mbox, (void*)b->drt0,
                  (void*)b->drn})
    if (p) (void)hipFree(p);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

}  // extern "C"
