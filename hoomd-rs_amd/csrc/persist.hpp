// The regions of the persistent applies (k_apply_tile for hard shapes,
// k_apply_lj for isotropic pair energies): region geometry, the staging of a
// region's box into LDS, the neighbour synchronisation between passes and
// the reload of the halo cells a pass changed.
#pragma once
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

// A region's box is staged as in k_pass_tile: ri (rows + 1 uint4: LDS offset
// of the row, first run start, first run length, second run start), coff
// (rows x (E_last + 1) LDS offsets of the cell boundaries, 16-bit: a box
// holds at most kTileLds / 16 entries), then the site entries.

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
                                            uint32_t pass) {
  const uint32_t parity = pass & 1u;
  if (threadIdx.x < 64) {
    if (threadIdx.x == 0) {
      uint32_t* box = &mbox[tile * kMailboxStride + parity];
      uint32_t spins = 0;
      while (__hip_atomic_load(box, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < nnb) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > kBarrierSpins) {
          atomicOr(&st->error, kErrBarrier);
          atomicCAS(&st->wait_at, 0u, (tile << 12) | ((pass & 0xfffu) + 1u));
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

// Residency probe (tc.probe): the persistent launch relies on every region
// being resident at once, which the host checks with the occupancy API; the
// probe checks it on the device with the kernel's own resources.  Thread 0 of
// every region adds 1 to a counter (word 2 of region 0's mailbox) and waits,
// bounded, until all regions have arrived; a region that times out sets
// kErrBarrier.  Every wavefront of the workgroup stays until then (a
// wavefront that ended early would free its SIMD slot and registers for a
// later workgroup, which the real apply does not).  The host clears the
// counter after the probe.
__device__ __forceinline__ void region_probe(ApplyState* st, uint32_t* mbox, uint32_t ntiles) {
  if (threadIdx.x == 0) {
    uint32_t* ctr = &mbox[2];
    __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t spins = 0;
    while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < ntiles) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > kBarrierSpins / 16) {
        atomicOr(&st->error, kErrBarrier);
        break;
      }
    }
  }
  __syncthreads();
}

// Geometry of region `tile`: its spaces org .. org + rl - 1 per axis, its box
// org - 1 .. org + 2 T (E spaces), the two slot runs of every box row along
// the last axis (p cells from m0, then the rest from 0) and the box centre.
struct Region {
  uint32_t tcd[3];
  int32_t org[3], mlo[3], rl[3];
  uint32_t p, m0;
  double c64[3];
};

template <int D>
__device__ __forceinline__ Region region_of(const Geometry& g, const TileCfg& tc, const double* origin,
                                            uint32_t tile) {
  constexpr int AL = D - 1;
  Region rg{};
  uint32_t r = tile;
  for (int a = D - 1; a >= 0; --a) {
    rg.tcd[a] = r % tc.nt[a];
    r /= tc.nt[a];
  }
  for (int a = 0; a < 3; ++a) {
    rg.org[a] = 0;
    rg.mlo[a] = 0;
    rg.rl[a] = 1;
    rg.c64[a] = 0.0;
  }
  for (int a = 0; a < D; ++a) {
    rg.org[a] = (int32_t)(2 * rg.tcd[a] * tc.T[a]);
    rg.mlo[a] = rg.org[a] - 1;
    rg.rl[a] = min((int32_t)(2 * tc.T[a]), g.shape[a] - rg.org[a]);
  }
  const int32_t SL = g.shape[AL], ml = rg.mlo[AL], EL = (int32_t)tc.E[AL];
  rg.p = ml < 0 ? 1u : (uint32_t)min(EL, SL - ml);
  rg.m0 = (uint32_t)(ml < 0 ? ml + SL : ml);
  for (int a = 0; a < D; ++a) {
    double c = origin[a] + ((double)rg.mlo[a] + 0.5 * (double)tc.E[a]) * g.width[a];
    c -= g.L[a] * floor((c + g.half[a]) / g.L[a]);
    rg.c64[a] = c;
  }
  return rg;
}

// The region that lane l < 3^D - 1 of wave 0 notifies after each pass: its
// l-th neighbour (with the periodic wrap of the region grid: duplicates, and
// the region itself, where the grid has fewer than three regions on an axis).
template <int D>
__device__ __forceinline__ uint32_t region_neighbour(const TileCfg& tc, const Region& rg, uint32_t lane) {
  constexpr uint32_t nnb = D == 3 ? 26u : 8u;
  uint32_t q = lane < nnb / 2 ? lane : lane + 1;  // skip the centre
  uint32_t nd[3] = {0, 0, 0};
  for (int a = D - 1; a >= 0; --a) {
    nd[a] = q % 3u;
    q /= 3u;
  }
  uint32_t nbt = 0;
  for (int a = 0; a < D; ++a) nbt = nbt * tc.nt[a] + (rg.tcd[a] + tc.nt[a] + nd[a] - 1u) % tc.nt[a];
  return nbt;
}

// LDS of a region's box: ri, coff, then the site entries.
struct RegionLds {
  uint4* ri;
  uint16_t* coff;
  float4* lb;
};
template <int D>
__device__ __forceinline__ RegionLds region_lds(unsigned char* smem, const TileCfg& tc) {
  const uint32_t R = tc.rows, CW = tc.E[D - 1] + 1;
  RegionLds l;
  l.ri = (uint4*)smem;
  l.coff = (uint16_t*)(smem + 16u * (R + 1));
  l.lb = (float4*)(smem + ((16u * (R + 1) + 2u * R * CW + 15u) & ~15u));
  return l;
}

// Stages the region's box into LDS once per apply: every row's runs and cell
// starts, then the sites (their FP64 copies as box-relative FP32 entries with
// the body index).  Wave 0 loads the runs and scans the row lengths; every
// thread loads cell starts with coalesced loads; a half-wave per row copies
// its runs.  Ends with a barrier; returns true (uniform over the workgroup)
// when the box holds more sites than the LDS capacity tc.maxb, in which case
// no entries were written.
template <int D>
__device__ __forceinline__ bool region_stage(const Geometry& g, const TileCfg& tc, const Region& rg, const Bodies& B,
                                             const Cells& C, const Work& w, const RegionLds& l) {
  constexpr int AL = D - 1;
  const uint32_t R = tc.rows, EL = tc.E[AL], CW = EL + 1;
  const uint32_t tid = threadIdx.x, NT = blockDim.x;
  const uint32_t wid = tid >> 6, lane = tid & 63u;
  const uint32_t p = rg.p, m0 = rg.m0;
  const int32_t* mlo = rg.mlo;
  uint4* ri = l.ri;
  uint16_t* coff = l.coff;
  float4* lb = l.lb;
  const uint32_t rpl = (R + 63) / 64;
  uint32_t rq[kRowsPerLane][4];
  if (wid == 0) {
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
      if ((uint32_t)i < nvals) {
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
  if (wid == 0) {
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
  const uint32_t total = ri[R].x;
  const bool overflow = total > tc.maxb;  // uniform over the workgroup
  if (!overflow) {
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
        if (dst[u] != 0xffffffffu) lb[dst[u]] = stage_entry64<D>(g, rg.c64, tmp[u], bod[u]);
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
          lb[q.x + e] = stage_entry64<D>(g, rg.c64, B.spos[bd], bd);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < kValsPerThread; ++i) {
      if (key[i] != 0xffffffffu) {
        const uint32_t r = key[i] >> 8, j = key[i] & 0xffu;
        const uint4 q = ri[r];
        coff[r * CW + (j <= p ? j : j - 1)] =
            (uint16_t)(j <= p ? q.x + (raw[i] - q.y) : q.x + q.z + (raw[i] - q.w));
      }
    }
  }
  __syncthreads();
  return overflow;
}

// After the wait that follows the pass of `colour`: reloads the halo cells of
// that colour -- box cells b with b = 1 - off (mod 2) per axis whose space
// lies outside the region, the only ones other regions changed -- from the
// FP64 sites, with device-scope loads (the first kRefreshHeld entries of a
// thread are loaded together).  Ends with a barrier.
template <int D>
__device__ __forceinline__ void region_reload(const Geometry& g, const TileCfg& tc, const Region& rg, const Bodies& B,
                                              const RegionLds& l, uint32_t colour) {
  constexpr int AL = D - 1;
  const uint32_t CW = tc.E[AL] + 1;
  const uint32_t tid = threadIdx.x, NT = blockDim.x;
  const int32_t* rl = rg.rl;
  const uint16_t* coff = l.coff;
  float4* lb = l.lb;
  uint32_t hdst[kRefreshHeld];
  double4 hval[kRefreshHeld];
  uint32_t nheld = 0;
  {
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
        const int32_t lc = (int32_t)bc[a] - 1;
        if (lc < 0 || lc >= rl[a]) own = false;
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
          lb[e] = stage_entry64<D>(g, rg.c64, v, bd);
        }
      }
    }
  }
#pragma unroll
  for (int h = 0; h < kRefreshHeld; ++h)
    if ((uint32_t)h < nheld) lb[hdst[h]] = stage_entry64<D>(g, rg.c64, hval[h], (uint32_t)__float_as_int(lb[hdst[h]].w));
  __syncthreads();
}

}  // namespace hmc
