// The persistent apply (mixed precision, hard spheres and disks): one
// cooperative launch runs every cycle and colour of an apply.
#include "persist.hpp"

namespace hmc {


// FP64 re-test of a band pair (rare; out of line), the site given by body,
// read with device-scope loads (another region may have written it).
template <int D>
__device__ __noinline__ bool retest64_body(const Geometry* gp, const double4* spos, uint32_t body, double t0,
                                           double t1, double t2) {
  const double t[3] = {t0, t1, t2};
  return overlap64<D>(*gp, t, ld_dev_d4(&spos[body]));
}

// The trial of space s (slots [s0, s0 + n)) from global memory with exact FP64
// pair tests on the sites, for a region whose box holds more bodies than its
// LDS copy (rare).  Same proposal and decision as every other path.
template <int D>
__device__ __noinline__ TrialOut trial_global_dev(const Geometry* gp, ApplyState* st, double4* bpos, double4* spos,
                                                  const uint32_t* cbody, const uint32_t* start,
                                                  hmc_trace_record* trace, unsigned long long trace_cap,
                                                  const double* zx, const double* zf, uint32_t s, uint32_t s0,
                                                  uint32_t n, uint32_t colour, uint32_t substep) {
  const Geometry& g = *gp;
  TrialOut o{0, 0};
  int32_t m[3] = {0, 0, 0};
  for (int a = 0; a < D; ++a) m[a] = (int32_t)((s / g.stride[a]) % (uint32_t)g.shape[a]);
  Sfc64 rng = counter_rng(st->step, substep, st->seed, s);
  const uint32_t pick = choose_index(rng, n);
  double v[3] = {0.0, 0.0, 0.0};
  ball_sample<D>(rng, st->max_move, zx, zf, v);
  const uint32_t slot = s0 + pick;
  const uint32_t body = cbody[slot];
  const double4 bp = bpos[body];
  double x[3] = {bp.x, bp.y, bp.z};
  for (int a = 0; a < D; ++a) x[a] += v[a];
  for (int a = 0; a < D; ++a) x[a] = wrap1(x[a], g.L[a]);
  bool accept = false;
  if (space_of<D>(g, st->origin, x) == s) {
    double t[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < D; ++a) t[a] = wrap1(x[a], g.L[a]);
    bool hit = false;
    for (int i0 = -1; i0 <= 1 && !hit; ++i0)
      for (int i1 = -1; i1 <= 1 && !hit; ++i1)
        for (int i2 = (D == 3 ? -1 : 0); i2 <= (D == 3 ? 1 : 0) && !hit; ++i2) {
          const int di[3] = {i0, i1, i2};
          uint32_t ns = 0;
          for (int a = 0; a < D; ++a) {
            const int32_t sh = g.shape[a];
            int32_t mm = m[a] + di[a];
            mm = mm < 0 ? mm + sh : (mm >= sh ? mm - sh : mm);
            ns += (uint32_t)mm * g.stride[a];
          }
          const uint32_t q0 = start[ns], q1 = start[ns + 1];
          for (uint32_t q = q0; q < q1 && !hit; ++q) {
            if (q == slot) continue;
            hit = overlap64<D>(g, t, ld_dev_d4(&spos[cbody[q]]));
          }
        }
    accept = !hit;
    if (accept) {
      double nsp[3] = {0.0, 0.0, 0.0};
      for (int a = 0; a < D; ++a) nsp[a] = wrap1(t[a] + 0.0, g.L[a]);
      const uint32_t fl = ghost_flags<D>(g, nsp);
      bpos[body] = make_double4(t[0], t[1], t[2], move_stamp(st->epoch));
      st_dev_d4(&spos[body], pack_site(nsp, fl));
    }
  }
  o.accepted = accept ? 1u : 0u;
  if (trace) {
    unsigned long long ti = atomicAdd((unsigned long long*)&st->trace_n, 1ull);
    if (ti < trace_cap) {
      hmc_trace_record rec;
      rec.substep = substep;
      rec.color = colour;
      rec.space = s;
      rec.body = (int64_t)body;
      rec.status = accept ? 1 : 0;
      rec.pad = 0;
      trace[ti] = rec;
    }
  }
  return o;
}

// Every cycle and colour of one apply for region `tile` of the system (g, st,
// w.start, w.mbox); the region's accepted moves are added to *acc.  zx, zf:
// the ziggurat layers (an LDS copy, or the constant tables).
template <int D>
__device__ __forceinline__ void apply_tile_run(const Geometry& g, ApplyState* st, const Bodies& B, const Cells& C,
                                               const Work& w, const TileCfg& tc, uint32_t tile,
                                               unsigned long long* acc, const double* zx, const double* zf) {
  constexpr int AL = D - 1;  // the row axis (stride 1)
  constexpr uint32_t NCOL = 1u << D;
  extern __shared__ __align__(16) unsigned char smem[];
  const RegionLds l = region_lds<D>(smem, tc);
  const uint32_t CW = tc.E[AL] + 1;
  const uint16_t* coff = l.coff;
  float4* lb = l.lb;
  __shared__ unsigned blk_acc;
  const uint32_t tid = threadIdx.x;
  const uint32_t wid = tid >> 6, lane = tid & 63u;
  const uint32_t cycles = st->cycles, substep0 = st->substep0;
  const uint64_t step = st->step;
  const uint32_t seed = st->seed;
  const double max_move = st->max_move;
  const double stamp = move_stamp(st->epoch);
  double origin[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < D; ++a) origin[a] = st->origin[a];
  const Region rg = region_of<D>(g, tc, origin, tile);
  const int32_t* org = rg.org;
  const int32_t* rl = rg.rl;
  const double* c64 = rg.c64;
  // this thread's trial digits (one space of each colour)
  uint32_t hh[3] = {0, 0, 0};
  {
    uint32_t r = tid;
    for (int a = D - 1; a >= 0; --a) {
      hh[a] = r % tc.T[a];
      r /= tc.T[a];
    }
  }
  const bool has_trial = tid < tc.trials;
  // wave 0's lane l < nnb notifies neighbour region l after each pass
  const uint32_t nnb = D == 3 ? 26u : 8u;
  const uint32_t nbt = wid == 0 && lane < nnb ? region_neighbour<D>(tc, rg, lane) : tile;

  if (tid == 0) blk_acc = 0;
  const bool overflow = region_stage<D>(g, tc, rg, B, C, w, l);

  unsigned retests = 0;
  // The proposal of pass i (cycle i / 2^D, colour i mod 2^D) depends only on
  // the counter, the space's body count and the chosen body's position, which
  // no other colour changes; it is drawn while the neighbour regions finish
  // the previous pass, and with it everything of the trial that depends on
  // the proposal alone: the wrapped trial position and its space test, its
  // FP32 image, and the site and LDS entry an acceptance writes.
  bool pgo = false, pin = false;
  uint32_t ps = 0, pcrow = 0, pe0 = 0, pn = 0, ppick = 0, pbody = 0;
  uint32_t pcb[3] = {0, 0, 0};
  double pt[3] = {0.0, 0.0, 0.0};
  float ptr[3] = {0.0f, 0.0f, 0.0f};
  double4 psp = make_double4(0.0, 0.0, 0.0, 0.0);
  float4 pent = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  auto propose = [&](uint32_t i) {
    const uint32_t colour = i & (NCOL - 1), substep = substep0 + (i >> D);
    bool active = has_trial && !overflow;
    uint32_t s = 0;
    for (int a = 0; a < D; ++a) {
      const uint32_t l = 2 * hh[a] + ((colour >> (D - 1 - a)) & 1u);
      if ((int32_t)l >= rl[a]) active = false;
      s += (uint32_t)(org[a] + (int32_t)l) * g.stride[a];
      pcb[a] = l + 1;
    }
    ps = s;
    pgo = false;
    if (!active) return;
    pcrow = D == 3 ? pcb[0] * tc.E[1] + pcb[1] : pcb[0];
    pe0 = coff[pcrow * CW + pcb[AL]];
    pn = coff[pcrow * CW + pcb[AL] + 1] - pe0;
    if (!pn) return;
    pgo = true;
    // the choice first (a few draws), then the chosen body's position in
    // flight during the displacement draw
    Sfc64 rng = counter_rng(step, substep, seed, s);
    ppick = choose_index(rng, pn);
    pbody = (uint32_t)__float_as_int(lb[pe0 + ppick].w);
    const double4 bp = B.bpos[pbody];
    double v[3] = {0.0, 0.0, 0.0};
    ball_sample<D>(rng, max_move, zx, zf, v);
    double x[3] = {bp.x, bp.y, bp.z};
    for (int a = 0; a < D; ++a) x[a] += v[a];
    for (int a = 0; a < D; ++a) x[a] = wrap1(x[a], g.L[a]);  // generate_trial_moves: wrap(propose(..))
    pin = space_of<D>(g, origin, x) == s;
    for (int a = 0; a < D; ++a) pt[a] = wrap1(x[a], g.L[a]);
    const float4 te = stage_entry64<D>(g, c64, make_double4(pt[0], pt[1], D == 3 ? pt[2] : 0.0, 0.0), 0u);
    ptr[0] = te.x;
    ptr[1] = te.y;
    ptr[2] = te.z;
    // update_body_properties (microstate.rs:985-1034): site = wrap(body + 0)
    double nsp[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < D; ++a) nsp[a] = wrap1(pt[a] + 0.0, g.L[a]);
    psp = pack_site(nsp, ghost_flags<D>(g, nsp));
    pent = stage_entry64<D>(g, c64, psp, pbody);
  };

  const uint32_t npass = cycles * NCOL;
  if (npass) propose(0);
  for (uint32_t i = 0; i < npass; ++i) {
    const uint32_t colour = i & (NCOL - 1), substep = substep0 + (i >> D);
    unsigned accepted = 0;
    if (has_trial && overflow) {
      bool active = true;
      uint32_t s = 0;
      for (int a = 0; a < D; ++a) {
        const uint32_t l = 2 * hh[a] + ((colour >> (D - 1 - a)) & 1u);
        if ((int32_t)l >= rl[a]) active = false;
        s += (uint32_t)(org[a] + (int32_t)l) * g.stride[a];
      }
      const uint32_t s0 = w.start[s], n = w.start[s + 1] - s0;
      if (active && n) {
        const TrialOut o = trial_global_dev<D>(g.self, st, B.bpos, B.spos, C.body, w.start,
                                               w.tracing ? w.trace : nullptr, w.trace_cap, zx, zf, s, s0, n,
                                               colour, substep);
        accepted = o.accepted;
        retests += o.retests;
      }
    } else if (pgo) {
      // ---- the trial of this pass from LDS
      const uint32_t s = ps, self = pe0 + ppick, body = pbody;
      if (pin) {
        constexpr int NR = D == 3 ? 9 : 3;
        uint32_t rlo[NR], rhi[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const int i0 = D == 3 ? r / 3 : r;
          const int i1 = D == 3 ? r % 3 : 0;
          const uint32_t row = D == 3 ? (pcb[0] + i0 - 1) * tc.E[1] + (pcb[1] + i1 - 1) : pcb[0] + i0 - 1;
          rlo[r] = coff[row * CW + pcb[AL] - 1];
          rhi[r] = coff[row * CW + pcb[AL] + 2];
        }
        const double* t = pt;
        bool hit = false;
        uint32_t rounds = 0;
#pragma unroll
        for (int r = 0; r < NR; ++r) rounds = max(rounds, rhi[r] - rlo[r]);
        const float* tr = ptr;
        const float lo = tc.band_lo, hi = tc.band_hi;
        // rows of three cells; only the centre row (the trial's own cell in
        // its middle) can hold the trial's own site.  Two rounds per step:
        // 2 NR LDS reads in flight together.  The FP32 distance is fused
        // (three roundings, within the band's 4 u relative term).
        constexpr int RC = NR / 2;
        uint32_t len[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) len[r] = rhi[r] - rlo[r];
        for (uint32_t kk = 0; kk < rounds && !hit; kk += 2) {
          uint32_t amb = 0;
#pragma unroll
          for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int r = 0; r < NR; ++r) {
              const uint32_t k2 = kk + (uint32_t)h;
              const uint32_t idx = rlo[r] + k2;
              const bool valid = k2 < len[r] && (r != RC || idx != self);
              const float4 q = lb[valid ? idx : self];
              const float dx = q.x - tr[0], dy = q.y - tr[1];
              float d2;
              if (D == 3) {
                const float dz = q.z - tr[2];
                d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
              } else {
                d2 = fmaf(dx, dx, dy * dy);
              }
              hit |= valid && d2 < lo;
              if (valid && !(d2 < lo) && !(d2 >= hi)) amb |= 1u << (h * NR + r);
            }
          }
          while (amb && !hit) {
            const int b2 = __ffs(amb) - 1;
            amb &= amb - 1;
            const int h = b2 / NR, r = b2 % NR;
            const float4 q = lb[rlo[r] + kk + (uint32_t)h];
            ++retests;
            hit |= retest64_body<D>(g.self, B.spos, (uint32_t)__float_as_int(q.w), t[0], t[1], t[2]);
          }
        }
        if (!hit) {
          // update_body_properties (microstate.rs:985-1034), formed with the proposal
          B.bpos[body] = make_double4(t[0], t[1], t[2], stamp);
          st_dev_d4(&B.spos[body], psp);
          lb[self] = pent;
          accepted = 1;
        }
      }
      if (w.tracing) {
        unsigned long long ti = atomicAdd((unsigned long long*)&st->trace_n, 1ull);
        if (ti < w.trace_cap) {
          hmc_trace_record rec;
          rec.substep = substep;
          rec.color = colour;
          rec.space = s;
          rec.body = (int64_t)body;
          rec.status = (int32_t)accepted;
          rec.pad = 0;
          w.trace[ti] = rec;
        }
      }
    }
    const unsigned long long mk = __ballot(accepted != 0);
    if (__lane_id() == 0 && mk) atomicAdd(&blk_acc, (unsigned)__popcll(mk));
    if (i + 1 == npass) break;  // no later pass reads the halo
    region_publish(w.mbox, nbt, nnb, i & 1u);
    // the next proposal reads only static cell data and this region's own
    // bodies, none of which pass i changed: drawn while the neighbours finish
    propose(i + 1);
    region_wait(st, w.mbox, tile, nnb, i);
    // ---- reload the halo cells of this colour
    if (!overflow)
      region_reload<D>(g, tc, rg, B, l, colour);
    else
      __syncthreads();
  }
  lane_add(&st->band_retests, retests);
  if (tid == 0 && overflow) atomicAdd(&st->tile_overflows, 1ull);
  __syncthreads();
  if (tid == 0 && blk_acc) atomicAdd(acc, (unsigned long long)blk_acc);
}

template <int D>
__global__ void __launch_bounds__(kTileThreads, 2) k_apply_tile(Geometry g, ApplyState* st, Bodies B, Cells C, Work w,
                                                             TileCfg tc) {
  const uint32_t b = blockIdx.x;
  const uint32_t tile = (b & 7u) * tc.xcd_chunk + (b >> 3);  // XCD-aware region order
  if (tile >= tc.ntiles) return;  // padding workgroups (grid rounded up to 8) take part in nothing
  if (tc.probe) {
    region_probe(st, w.mbox, tc.ntiles);
    return;
  }
  __shared__ double zx[257], zf[257];
  if (D == 3) stage_zig(zx, zf);  // (read after region_stage's barriers)
  apply_tile_run<D>(g, st, B, C, w, tc, tile, &w.acc_slots[b & (kAccSlots - 1)], zx, zf);
}

// The persistent applies of every replica of a batch in one cooperative
// launch.  The batch's regions are packed, whole replicas at a time, into
// rounds of at most one region per workgroup (round k: regions rt0[k] ..
// rt0[k] + rn[k] - 1 of the batch; trep: a region's replica, toff: the
// replicas' first regions).  Workgroup b runs its region of round 0 through
// every cycle of its replica, then its region of round 1, and so on, each
// with its replica's geometry, region shape, state, spaces and mailboxes.
// A region waits only on regions of its own replica, which are all in the
// same round and each run by another resident workgroup that has finished
// the earlier rounds' regions, none of which waits on a later round: every
// round completes.  probe: only check that every workgroup of the launch is
// resident (counter: word 2 of the first region's mailbox; nwg workgroups).
template <int D>
__global__ void __launch_bounds__(kTileThreads, 2) k_bapply_tile(const BatchRep* reps, const TileCfg* tcs,
                                                              const uint32_t* toff, const uint32_t* trep,
                                                              ApplyState* sts, const uint32_t* rt0,
                                                              const uint32_t* rn, uint32_t nrounds, uint32_t nwg,
                                                              uint32_t xcd_chunk, uint32_t probe, Bodies B,
                                                              Cells C, Work w) {
  __shared__ Geometry sg;
  __shared__ TileCfg stc;
  const uint32_t b = blockIdx.x;
  const uint32_t t = (b & 7u) * xcd_chunk + (b >> 3);
  if (t >= nwg) return;  // uniform over the block
  if (probe) {
    region_probe(sts + trep[0], w.mbox, nwg);
    return;
  }
  for (uint32_t k = 0; k < nrounds; ++k) {
    if (t >= rn[k]) continue;  // uniform over the block
    const uint32_t gt = rt0[k] + t, r = trep[gt];
    const uint32_t* gsrc = (const uint32_t*)&reps[r].g;
    for (uint32_t i = threadIdx.x; i < sizeof(Geometry) / 4; i += blockDim.x) ((uint32_t*)&sg)[i] = gsrc[i];
    const uint32_t* tsrc = (const uint32_t*)&tcs[r];
    for (uint32_t i = threadIdx.x; i < sizeof(TileCfg) / 4; i += blockDim.x) ((uint32_t*)&stc)[i] = tsrc[i];
    __syncthreads();
    Work wr = w;
    wr.start = w.start + reps[r].space_off;
    wr.mbox = w.mbox + (size_t)toff[r] * kMailboxStride;
    wr.tracing = 0;
    // the ziggurat layers from the constant tables: the LDS they would take
    // is worth more as region entries here (more regions per CU)
    apply_tile_run<D>(sg, sts + r, B, C, wr, stc, gt - toff[r], &w.acc_slots[r], kZigX, kZigF);
    __syncthreads();  // sg and stc are rewritten for the next round
  }
}

// explicit instantiations (launched from the host translation units)
template __global__ void k_apply_tile<2>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, TileCfg tc);
template __global__ void k_apply_tile<3>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, TileCfg tc);
template __global__ void k_bapply_tile<3>(const BatchRep*, const TileCfg*, const uint32_t*, const uint32_t*,
                                          ApplyState*, const uint32_t*, const uint32_t*, uint32_t, uint32_t, uint32_t,
                                          uint32_t, Bodies, Cells, Work);

}  // namespace hmc
