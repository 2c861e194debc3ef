// The persistent apply (mixed precision, hard spheres and disks): one
// cooperative launch runs every cycle and colour of an apply.
#include "sweep_device.hpp"

namespace hmc {

// ------------------------------------------------ persistent apply (mixed) --
// One cooperative launch runs every cycle and colour of an apply.  Workgroup
// b owns a region of 2 T spaces per axis (the last region along an axis may
// be shorter) and keeps the region plus a one-space halo -- every cell its
// trials can read -- in LDS for the whole apply: staged once, its own cells
// updated in place by its accepted moves, and after the pass of colour c the
// halo cells of colour c (the only ones other workgroups have changed)
// reloaded from the FP64 sites.  Consecutive passes are separated by a
// neighbour synchronisation (below), not by a grid barrier.
//
// An LDS entry holds the site relative to the box centre in FP32 (NaN for a
// site without images, as in the per-colour kernel) with the body index in w.

// Neighbour synchronisation of the persistent apply, through mailboxes: two
// counters per region (one per pass parity), kMailboxStride words apart.
// After pass i a region adds 1 to the counter i mod 2 of each of its 3^D - 1
// neighbours (with multiplicity where the region grid wraps onto itself; the
// relation is symmetric, so every region receives exactly 3^D - 1 per pass)
// and then waits until its own counter reaches 3^D - 1, which it resets.  A
// region reads only its own and its neighbours' cells, so once all of them
// have finished pass i none can be a pass behind or ahead while it runs pass
// i + 1: nothing it reads is being written.  A neighbour can therefore never
// finish pass i + 2 (the next use of the same counter) before this region has
// consumed pass i, and every counter is 0 again when an apply ends.  The
// FP64 sites that neighbours read are written with device-scope stores, and
// every wave waits for its stores to complete before the increments are sent
// (no cache writeback or invalidation: those, issued by the regions that
// finish first, stalled the ones still working); the neighbours read them
// with device-scope loads.  One lane per region polls, on a counter of its
// own.  The launch is cooperative, so every
// region is resident.  The wait is bounded: past the bound the state's error
// word is set and the wait returns, so every wave still ends.
__device__ __forceinline__ void region_publish(uint32_t* mbox, uint32_t nbt, uint32_t nnb, uint32_t parity) {
  __builtin_amdgcn_s_waitcnt(0);  // this wave's stores have completed
  __syncthreads();
  if (threadIdx.x < 64) {
    if (threadIdx.x < nnb)
      __hip_atomic_fetch_add(&mbox[nbt * kMailboxStride + parity], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}
__device__ __forceinline__ void region_wait(ApplyState* st, uint32_t* mbox, uint32_t tile, uint32_t nnb,
                                            uint32_t parity) {
  if (threadIdx.x < 64) {
    if (threadIdx.x == 0) {
      uint32_t* box = &mbox[tile * kMailboxStride + parity];
      uint32_t spins = 0;
      while (__hip_atomic_load(box, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < nnb) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > kBarrierSpins) {
          atomicOr(&st->error, kErrBarrier);
          break;
        }
      }
      __hip_atomic_store(box, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  __syncthreads();
}

// LDS entry of the persistent apply from a site's FP64 copy: the image
// nearest the box centre c, relative to c, computed in FP64 and rounded once
// to FP32 (NaN for a site without images), body index in w.
template <int D>
__device__ __forceinline__ float4 stage_entry64(const Geometry& g, const double* c, double4 sp, uint32_t body) {
  const double x[3] = {sp.x, sp.y, sp.z};
  float r[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int a = 0; a < D; ++a) {
    double u = x[a] - c[a];
    if (u >= g.half[a])
      u -= g.L[a];
    else if (u < -g.half[a])
      u += g.L[a];
    r[a] = (float)u;
  }
  if ((uint32_t)__double_as_longlong(sp.w) & kNoImages) r[0] = __int_as_float(0x7fc00000);
  return make_float4(r[0], r[1], r[2], __int_as_float((int)body));
}

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
      bpos[body] = make_double4(t[0], t[1], t[2], 0.0);
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

template <int D>
__global__ void __launch_bounds__(kTileThreads, 2) k_apply_tile(Geometry g, ApplyState* st, Bodies B, Cells C, Work w,
                                                             TileCfg tc) {
  using VT = float4;
  constexpr int AL = D - 1;  // the row axis (stride 1)
  constexpr uint32_t NCOL = 1u << D;
  const uint32_t b = blockIdx.x;
  const uint32_t tile = (b & 7u) * tc.xcd_chunk + (b >> 3);  // XCD-aware region order
  const bool live = tile < tc.ntiles;  // padding workgroups (grid rounded up to 8) return at once
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t R = tc.rows, EL = tc.E[AL], CW = EL + 1;
  uint4* ri = (uint4*)smem;  // as in k_pass_tile
  uint32_t* coff = (uint32_t*)(smem + 16u * (R + 1));
  VT* lb = (VT*)(smem + ((16u * (R + 1) + 4u * R * CW + 15u) & ~15u));
  __shared__ double zx[257], zf[257];
  __shared__ unsigned blk_acc;
  const uint32_t tid = threadIdx.x, NT = blockDim.x;
  const uint32_t wid = tid >> 6, lane = tid & 63u;
  const uint32_t cycles = st->cycles, substep0 = st->substep0;
  if (!live) return;  // padding workgroups take part in nothing
  const uint64_t step = st->step;
  const uint32_t seed = st->seed;
  const double max_move = st->max_move;
  double origin[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < D; ++a) origin[a] = st->origin[a];

  // region: spaces org .. org + rl - 1 per axis; box: org - 1 .. org + 2 T
  uint32_t tcd[3] = {0, 0, 0};
  {
    uint32_t r = live ? tile : 0u;
    for (int a = D - 1; a >= 0; --a) {
      tcd[a] = r % tc.nt[a];
      r /= tc.nt[a];
    }
  }
  int32_t org[3] = {0, 0, 0}, mlo[3] = {0, 0, 0}, rl[3] = {1, 1, 1};
  for (int a = 0; a < D; ++a) {
    org[a] = (int32_t)(2 * tcd[a] * tc.T[a]);
    mlo[a] = org[a] - 1;
    rl[a] = min((int32_t)(2 * tc.T[a]), g.shape[a] - org[a]);
  }
  const int32_t SL = g.shape[AL], ml = mlo[AL];
  const uint32_t p = ml < 0 ? 1u : (uint32_t)min((int32_t)EL, SL - ml);
  const uint32_t m0 = (uint32_t)(ml < 0 ? ml + SL : ml);
  double c64[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < D; ++a) {
    double c = origin[a] + ((double)mlo[a] + 0.5 * (double)tc.E[a]) * g.width[a];
    c -= g.L[a] * floor((c + g.half[a]) / g.L[a]);
    c64[a] = c;
  }
  // this thread's trial digits (one space of each colour)
  uint32_t hh[3] = {0, 0, 0};
  {
    uint32_t r = tid;
    for (int a = D - 1; a >= 0; --a) {
      hh[a] = r % tc.T[a];
      r /= tc.T[a];
    }
  }
  const bool has_trial = live && tid < tc.trials;
  // wave 0's lane l < nnb notifies neighbour region l (3^D - 1 of them, with
  // the periodic wrap of the region grid: duplicates, and the region itself,
  // where the grid has fewer than three regions on an axis)
  const uint32_t nnb = D == 3 ? 26u : 8u;
  uint32_t nbt = tile;
  if (wid == 0 && lane < nnb) {
    uint32_t q = lane < nnb / 2 ? lane : lane + 1;  // skip the centre
    uint32_t nd[3] = {0, 0, 0};
    for (int a = D - 1; a >= 0; --a) {
      nd[a] = q % 3u;
      q /= 3u;
    }
    nbt = 0;
    for (int a = 0; a < D; ++a) nbt = nbt * tc.nt[a] + (tcd[a] + tc.nt[a] + nd[a] - 1u) % tc.nt[a];
  }

  // ---- staging (once): rows, cell offsets, sites with their body indices
  const uint32_t rpl = (R + 63) / 64;
  uint32_t rq[kRowsPerLane][4];
  if (live && wid == 0) {
#pragma unroll
    for (int u = 0; u < kRowsPerLane; ++u) {
      const uint32_t r = min(lane * rpl + (uint32_t)u, R - 1);
      uint32_t sb = 0, q = r;
      for (int a = AL - 1; a >= 0; --a) {
        const uint32_t e = a > 0 ? q % tc.E[a] : q;
        q = a > 0 ? q / tc.E[a] : 0u;
        const int32_t sh = g.shape[a];
        int32_t mm = mlo[a] + (int32_t)e;
        mm = mm < 0 ? mm + sh : (mm >= sh ? mm - sh : mm);
        sb += (uint32_t)mm * g.stride[a];
      }
      rq[u][0] = w.start[sb + m0];
      rq[u][1] = w.start[sb + m0 + p];
      rq[u][2] = w.start[sb];
      rq[u][3] = w.start[sb + (EL - p)];
    }
  }
  const uint32_t nvals = (R * (EL + 2) + NT - 1) / NT;
  uint32_t raw[kValsPerThread], key[kValsPerThread];
  {
    uint32_t d[3] = {0, 0, 0};
    uint32_t r = tid;
    for (int a = AL; a > 0; --a) {
      const uint32_t rad = a == AL ? EL + 2 : tc.E[a];
      d[a] = r % rad;
      r /= rad;
    }
    d[0] = r;
#pragma unroll
    for (int i = 0; i < kValsPerThread; ++i) {
      raw[i] = 0u;
      key[i] = 0xffffffffu;
      if (live && (uint32_t)i < nvals) {
        if (d[0] < tc.E[0]) {
          uint32_t sb = 0, row = 0;
          for (int a = 0; a < AL; ++a) {
            const int32_t sh = g.shape[a];
            int32_t mm = mlo[a] + (int32_t)d[a];
            mm = mm < 0 ? mm + sh : (mm >= sh ? mm - sh : mm);
            sb += (uint32_t)mm * g.stride[a];
            row = row * tc.E[a] + d[a];
          }
          const uint32_t j = d[AL];
          raw[i] = w.start[j <= p ? sb + m0 + j : sb + (j - p - 1)];
          key[i] = (row << 8) | j;
        }
        uint32_t carry = 0;
        for (int a = AL; a >= 0; --a) {
          d[a] += tc.step[a] + carry;
          carry = 0;
          const uint32_t rad = a == AL ? EL + 2 : tc.E[a];
          if (a > 0 && d[a] >= rad) {
            d[a] -= rad;
            carry = 1;
          }
        }
      }
    }
  }
  if (D == 3) stage_zig(zx, zf);
  if (tid == 0) blk_acc = 0;
  if (live && wid == 0) {
    uint32_t len[kRowsPerLane], sum = 0;
#pragma unroll
    for (int u = 0; u < kRowsPerLane; ++u) {
      const uint32_t r = lane * rpl + (uint32_t)u;
      len[u] = (uint32_t)u < rpl && r < R ? (rq[u][1] - rq[u][0]) + (rq[u][3] - rq[u][2]) : 0u;
      sum += len[u];
    }
    const uint32_t incl = wave_incl_scan(sum);
    uint32_t off = incl - sum;
#pragma unroll
    for (int u = 0; u < kRowsPerLane; ++u) {
      const uint32_t r = lane * rpl + (uint32_t)u;
      if ((uint32_t)u < rpl && r < R) ri[r] = make_uint4(off, rq[u][0], rq[u][1] - rq[u][0], rq[u][2]);
      off += len[u];
    }
    if (lane == 63) ri[R] = make_uint4(incl, 0u, 0u, 0u);
  }
  __syncthreads();
  const uint32_t total = live ? ri[R].x : 0u;
  const bool overflow = total > tc.maxb;  // uniform over the workgroup
  if (live && !overflow) {
    const uint32_t hw = tid >> 5, hl = tid & 31u, NH = NT >> 5;
    for (uint32_t r0 = hw; r0 < R; r0 += NH * kStageRows) {
      uint32_t dst[kStageRows], src[kStageRows], bod[kStageRows];
#pragma unroll
      for (int u = 0; u < kStageRows; ++u) {
        const uint32_t r = r0 + (uint32_t)u * NH;
        dst[u] = 0xffffffffu;
        src[u] = 0u;
        if (r < R) {
          const uint4 q = ri[r];
          const uint32_t len = ri[r + 1].x - q.x;
          if (hl < len) {
            src[u] = hl < q.z ? q.y + hl : q.w + (hl - q.z);
            dst[u] = q.x + hl;
          }
        }
        if (dst[u] != 0xffffffffu) bod[u] = C.body[src[u]];
      }
      double4 tmp[kStageRows];
#pragma unroll
      for (int u = 0; u < kStageRows; ++u)
        if (dst[u] != 0xffffffffu) tmp[u] = B.spos[bod[u]];
#pragma unroll
      for (int u = 0; u < kStageRows; ++u)
        if (dst[u] != 0xffffffffu) lb[dst[u]] = stage_entry64<D>(g, c64, tmp[u], bod[u]);
      // entries past the first 32 of a row (rare)
#pragma unroll 1
      for (int u = 0; u < kStageRows; ++u) {
        const uint32_t r = r0 + (uint32_t)u * NH;
        if (r >= R) continue;
        const uint4 q = ri[r];
        const uint32_t len = ri[r + 1].x - q.x;
        for (uint32_t e = 32u + hl; e < len; e += 32u) {
          const uint32_t sl = e < q.z ? q.y + e : q.w + (e - q.z);
          const uint32_t bd = C.body[sl];
          lb[q.x + e] = stage_entry64<D>(g, c64, B.spos[bd], bd);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < kValsPerThread; ++i) {
      if (key[i] != 0xffffffffu) {
        const uint32_t r = key[i] >> 8, j = key[i] & 0xffu;
        const uint4 q = ri[r];
        coff[r * CW + (j <= p ? j : j - 1)] = j <= p ? q.x + (raw[i] - q.y) : q.x + q.z + (raw[i] - q.w);
      }
    }
  }
  __syncthreads();

  unsigned retests = 0;
  // The proposal of pass i (cycle i / 2^D, colour i mod 2^D) depends only on
  // the counter, the space's body count and the chosen body's position, which
  // no other colour changes; it is drawn while the neighbour regions finish
  // the previous pass.
  bool pgo = false;
  uint32_t ps = 0, pcrow = 0, pe0 = 0, pn = 0, ppick = 0, pbody = 0;
  uint32_t pcb[3] = {0, 0, 0};
  double pv[3] = {0.0, 0.0, 0.0};
  double4 pbp = make_double4(0.0, 0.0, 0.0, 0.0);
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
    pbp = B.bpos[pbody];
    ball_sample<D>(rng, max_move, zx, zf, pv);
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
      double x[3] = {pbp.x, pbp.y, pbp.z};
      for (int a = 0; a < D; ++a) x[a] += pv[a];
      for (int a = 0; a < D; ++a) x[a] = wrap1(x[a], g.L[a]);  // generate_trial_moves: wrap(propose(..))
      if (space_of<D>(g, origin, x) == s) {
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
        double t[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < D; ++a) t[a] = wrap1(x[a], g.L[a]);
        bool hit = false;
        uint32_t rounds = 0;
#pragma unroll
        for (int r = 0; r < NR; ++r) rounds = max(rounds, rhi[r] - rlo[r]);
        float tr[3] = {0.0f, 0.0f, 0.0f};
        {
          const float4 te = stage_entry64<D>(g, c64, make_double4(t[0], t[1], D == 3 ? t[2] : 0.0, 0.0), 0u);
          tr[0] = te.x;
          tr[1] = te.y;
          tr[2] = te.z;
        }
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
          // update_body_properties (microstate.rs:985-1034)
          double nsp[3] = {0.0, 0.0, 0.0};
          for (int a = 0; a < D; ++a) nsp[a] = wrap1(t[a] + 0.0, g.L[a]);
          const uint32_t fl = ghost_flags<D>(g, nsp);
          const double4 sp = pack_site(nsp, fl);
          B.bpos[body] = make_double4(t[0], t[1], t[2], 0.0);
          st_dev_d4(&B.spos[body], sp);
          lb[self] = stage_entry64<D>(g, c64, sp, body);
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
    region_wait(st, w.mbox, tile, nnb, i & 1u);
    // ---- reload the halo cells of this colour: box cells b with b = 1 - off
    // (mod 2) per axis whose space lies outside the region (the first
    // kRefreshHeld entries of a thread are loaded together)
    uint32_t hdst[kRefreshHeld];
    double4 hval[kRefreshHeld];
    uint32_t nheld = 0;
    if (live && !overflow) {
      uint32_t nc = 1, cnt[3] = {1, 1, 1};
      for (int a = 0; a < D; ++a) {
        cnt[a] = tc.T[a] + 1;
        nc *= cnt[a];
      }
      for (uint32_t ci = tid; ci < nc; ci += NT) {
        uint32_t r = ci, bc[3] = {0, 0, 0};
        bool own = true;
        for (int a = D - 1; a >= 0; --a) {
          const uint32_t j = r % cnt[a];
          r /= cnt[a];
          bc[a] = 2 * j + (((colour >> (D - 1 - a)) & 1u) ? 0u : 1u);
          const int32_t l = (int32_t)bc[a] - 1;
          if (l < 0 || l >= rl[a]) own = false;
        }
        if (own) continue;
        const uint32_t row = D == 3 ? bc[0] * tc.E[1] + bc[1] : bc[0];
        const uint32_t ea = coff[row * CW + bc[AL]], eb = coff[row * CW + bc[AL] + 1];
        for (uint32_t e = ea; e < eb; ++e) {
          const uint32_t bd = (uint32_t)__float_as_int(lb[e].w);
          const double4 v = ld_dev_d4(&B.spos[bd]);
          if (nheld < (uint32_t)kRefreshHeld) {
#pragma unroll
            for (int h = 0; h < kRefreshHeld; ++h)
              if ((uint32_t)h == nheld) {
                hdst[h] = e;
                hval[h] = v;
              }
            ++nheld;
          } else {
            lb[e] = stage_entry64<D>(g, c64, v, bd);
          }
        }
      }
    }
#pragma unroll
    for (int h = 0; h < kRefreshHeld; ++h)
      if ((uint32_t)h < nheld) lb[hdst[h]] = stage_entry64<D>(g, c64, hval[h], (uint32_t)__float_as_int(lb[hdst[h]].w));
    __syncthreads();
  }
  lane_add(&st->band_retests, retests);
  if (tid == 0 && live && overflow) atomicAdd(&st->tile_overflows, 1ull);
  __syncthreads();
  if (tid == 0 && blk_acc) atomicAdd(&w.acc_slots[b & (kAccSlots - 1)], (unsigned long long)blk_acc);
}

// explicit instantiations (launched from the host translation units)
template __global__ void k_apply_tile<2>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, TileCfg tc);
template __global__ void k_apply_tile<3>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, TileCfg tc);

}  // namespace hmc
