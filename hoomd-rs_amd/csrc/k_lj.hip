// Isotropic pair energies (LJ, shifted LJ, the overlap penalty): one
// wavefront per trial with a Metropolis decision on dH.
#include "sweep_device.hpp"

namespace hmc {

// Sums over the candidate sites of one LJ trial (reduced over the
// wavefront) of the final and initial pair energies and of their
// magnitudes: FP64 exactly as the reference forms each term, or (F32) with
// lj_energy32, whose magnitudes are the A of its error bound.  The
// candidates are the slots of the neighbour-space segments (seg_lo, seg_pre:
// entry j in register j / 64 of lane j % 64).  The membership test r^2 <
// r_cut^2 is the FP64 one in both cases.  tab (when not null) holds the
// slot of every candidate and replaces the segment search.
template <int D, bool F32>
__device__ __forceinline__ WaveSums lj_trial_sums(const Geometry& g, const Cells& C, const uint32_t* seg_lo,
                                                  const uint32_t* seg_pre, uint32_t NS, uint32_t total,
                                                  uint32_t slot, const double* t, const double* xi,
                                                  const uint32_t* tab) {
  using acc_t = typename std::conditional<F32, float, double>::type;
  const uint32_t lane = __lane_id();
  const float sig2 = (float)(g.lj_sigma * g.lj_sigma), eps4 = (float)g.lj_eps4, shift = (float)g.lj_shift;
  acc_t ef = 0, af = 0, ei = 0, ai = 0;
  uint32_t nf = 0, ni = 0;
  // every lane runs the same number of rounds: the shuffles below read
  // segment data held by other lanes, which must all be active.  kLjBatch
  // rounds of 64 candidates at a time: their segment searches interleave and
  // their site loads are in flight together.
  for (uint32_t c0 = 0; c0 < total; c0 += 64 * kLjBatch) {
    uint32_t qv[kLjBatch];
    bool ok[kLjBatch];
#pragma unroll
    for (uint32_t r = 0; r < kLjBatch; ++r) {
      const uint32_t c = c0 + 64 * r + lane;
      if (tab) {  // uniform
        qv[r] = c < total ? tab[c] : 0u;
        ok[r] = c < total && qv[r] != slot;
        continue;
      }
      // segment of candidate c: the last j with pre[j] <= c, by a fixed-depth
      // binary search over pre[0..NS) (pre[0] = 0, NS <= 128; both registers
      // are shuffled and the reader selects)
      uint32_t jlo = 0;
#pragma unroll
      for (uint32_t stepw = 64; stepw >= 1; stepw >>= 1) {
        const uint32_t mid = jlo + stepw;
        const uint32_t pa = __shfl(seg_pre[0], (int)(mid & 63u), 64);
        const uint32_t pb = __shfl(seg_pre[1], (int)(mid & 63u), 64);
        if (mid < NS && (mid < 64 ? pa : pb) <= c) jlo = mid;
      }
      const uint32_t pa = __shfl(seg_pre[0], (int)(jlo & 63u), 64), pb = __shfl(seg_pre[1], (int)(jlo & 63u), 64);
      const uint32_t la = __shfl(seg_lo[0], (int)(jlo & 63u), 64), lb2 = __shfl(seg_lo[1], (int)(jlo & 63u), 64);
      const uint32_t plo = jlo < 64 ? pa : pb;
      const uint32_t llo = jlo < 64 ? la : lb2;
      qv[r] = llo + (c - plo);
      // lanes past the end and the body's own site (filter body_tag != site_j.body_tag)
      ok[r] = c < total && qv[r] != slot;
    }
    double4 sjv[kLjBatch];
#pragma unroll
    for (uint32_t r = 0; r < kLjBatch; ++r) sjv[r] = ok[r] ? C.pos64[qv[r]] : make_double4(0.0, 0.0, 0.0, 0.0);
#pragma unroll
    for (uint32_t r = 0; r < kLjBatch; ++r) {
      if (!ok[r]) continue;
      const double4 sj = sjv[r];
      const double r2f = lj_r2<D>(g, t, sj);
      if (r2f < g.lj_rc2) {
        if constexpr (F32) {
          float a;
          ef += lj_energy32(sig2, eps4, shift, r2f, &a);
          af += a;
        } else {
          const double e = lj_energy(g, r2f);
          ef += e;
          af += fabs(e);
        }
        nf += 1;
      }
      const double r2i = lj_r2<D>(g, xi, sj);
      if (r2i < g.lj_rc2) {
        if constexpr (F32) {
          float a;
          ei += lj_energy32(sig2, eps4, shift, r2i, &a);
          ai += a;
        } else {
          const double e = lj_energy(g, r2i);
          ei += e;
          ai += fabs(e);
        }
        ni += 1;
      }
    }
  }
  WaveSums ws;
  ws.ef = (double)wave_sum(ef);
  ws.af = (double)wave_sum(af);
  ws.ei = (double)wave_sum(ei);
  ws.ai = (double)wave_sum(ai);
  ws.nf = wave_sum(nf);
  ws.ni = wave_sum(ni);
  return ws;
}

// One LJ trial (parallel_sweep.rs:234-277) by a whole wavefront: the random
// proposal is drawn identically in every lane, the lanes share out the
// candidate sites of the 3^D (or 5^D) neighbour spaces, compute the final
// and initial pair energies of each in FP64 exactly as the reference does,
// and reduce.  The reference sums in its VecCell order; any two summation
// orders of n terms differ by at most 2 gamma_{n-1} sum|terms|, so the
// decision is exact whenever that bound cannot change it (counted otherwise).
// tab: kLjTab words of LDS for this wavefront (or null), where the
// candidates' slots are laid out when there are at most kLjTab of them.  The neighbour
// segments are listed before the proposal is drawn, so that their loads and
// the body's overlap the random draws.
template <int D, int PREC>
__device__ void trial_lj(const Geometry& g, const ApplyState* st, const Bodies& B, const Cells& C, const Work& w,
                         const double* zx, const double* zf, uint32_t s, const int32_t* m, uint32_t s0, uint32_t n,
                         uint32_t colour, uint32_t substep, LjOut& o, uint32_t* tab) {
  const uint32_t lane = __lane_id();
  Sfc64 rng = counter_rng(st->step, substep, st->seed, s);
  const uint32_t pick = choose_index(rng, n);
  const uint32_t slot = s0 + pick;
  const uint32_t body = C.body[slot];
  // neighbour spaces (de-duplicated per axis), as segments of slots
  uint32_t nb[3][5];
  uint32_t nn[3] = {1, 1, 1};
  for (int a = 0; a < D; ++a) {
    const int32_t sh = g.shape[a], r = g.reach[a];
    uint32_t k = 0;
    for (int off = -r; off <= r; ++off) {
      const int32_t mm = ((m[a] + off) % sh + sh) % sh;
      bool seen = false;
      for (uint32_t j = 0; j < k; ++j) seen |= nb[a][j] == (uint32_t)mm * g.stride[a];
      if (!seen) nb[a][k++] = (uint32_t)mm * g.stride[a];
    }
    nn[a] = k;
  }
  const uint32_t NS = nn[0] * nn[1] * nn[2];  // at most 125
  uint32_t seg_lo[2] = {0, 0}, seg_pre[2] = {0, 0}, seg_cnt[2] = {0, 0};
  uint32_t base = 0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const uint32_t j = lane + 64u * h;
    uint32_t cnt = 0, lo = 0;
    if (j < NS) {
      const uint32_t i2 = j % nn[2], i1 = (j / nn[2]) % nn[1], i0 = j / (nn[2] * nn[1]);
      const uint32_t ns = nb[0][i0] + (D > 1 ? nb[1][i1] : 0u) + (D > 2 ? nb[2][i2] : 0u);
      lo = w.start[ns];
      cnt = w.start[ns + 1] - lo;
    }
    const uint32_t incl = wave_incl_scan(cnt);
    seg_lo[h] = lo;
    seg_pre[h] = base + incl - cnt;
    seg_cnt[h] = cnt;
    base += __shfl(incl, 63, 64);
  }
  const uint32_t total = base;
  // the candidates' slots in LDS (a run of consecutive slots per segment)
  const bool tabbed = tab != nullptr && total <= kLjTab;
  if (tabbed) {
#pragma unroll
    for (int h = 0; h < 2; ++h)
      for (uint32_t k = 0; k < seg_cnt[h]; ++k) tab[seg_pre[h] + k] = seg_lo[h] + k;
    wave_lds_order();
  }
  const double4 bp = B.bpos[body];
  const double4 sp = B.spos[body];
  double v[3] = {0.0, 0.0, 0.0};
  ball_sample<D>(rng, st->max_move, zx, zf, v);
  double x[3] = {bp.x, bp.y, bp.z};
  for (int a = 0; a < D; ++a) x[a] += v[a];
  for (int a = 0; a < D; ++a) x[a] = wrap1(x[a], g.L[a]);  // generate_trial_moves: wrap(propose(..))
  bool accept = false;
  if (space_of<D>(g, st->origin, x) == s) {
    double t[3] = {0.0, 0.0, 0.0};  // wrap(x + 0) == wrap(x)
    for (int a = 0; a < D; ++a) t[a] = wrap1(x[a], g.L[a]);
    const double xi[3] = {sp.x, sp.y, sp.z};
    const uint32_t* tb = tabbed ? tab : nullptr;
    const uint32_t total = base;
    bool amb = false;
    bool decided = false;
    if constexpr (PREC == HMC_MIXED) {
      // FP32 energies: the difference of the two FP32 sums is exact in FP64;
      // |dH32 - dH| <= (48 u + gamma_(rounds + 6)) (A_f + A_i) for the exact
      // dH (each lane sums at most `rounds` terms, the reduction adds 6
      // levels), plus 2^-40 (A_f + A_i) for the reference's own roundings.
      const WaveSums w32 = lj_trial_sums<D, true>(g, C, seg_lo, seg_pre, NS, total, slot, t, xi, tb);
      const double dH32 = w32.ef - w32.ei;
      const double rounds = (double)((total + 63) / 64);
      const double delta32 = ((48.0 + rounds + 8.0) * 0x1p-24 + 0x1p-40) * (w32.af + w32.ai) * (1.0 + 0x1p-10) +
                             0x1p-1070;
      if (g.lj_fp32_error && isfinite(dH32) && isfinite(delta32)) {
        const WaveSums w64 = lj_trial_sums<D, false>(g, C, seg_lo, seg_pre, NS, total, slot, t, xi, tb);
        o.err = fabs(dH32 - (w64.ef - w64.ei));
        o.ratio = o.err / delta32;
      }
      const Sfc64 rng0 = rng;
      bool open = false;
      accept = lj_metropolis(rng, dH32, delta32, st->kT, &open) != 0;
      decided = !open;
      if (open) {
        rng = rng0;  // the FP64 decision draws the same u
        o.retests = lane == 0 ? 1u : 0u;
      }
    }
    if (!decided) {
      const WaveSums ws = lj_trial_sums<D, false>(g, C, seg_lo, seg_pre, NS, total, slot, t, xi, tb);
      // energy_final - energy_initial (pairwise_cutoff.rs:919-941): an
      // infinite final energy is returned as is and rejected; an infinite
      // initial one makes dH = -inf, accepted
      if (ws.ef == INFINITY) {
        accept = false;
      } else if (ws.ei == INFINITY) {
        accept = true;
      } else {
        const double dH = ws.ef - ws.ei;
        const double nmax = (double)max(max(ws.nf, ws.ni), 1u);
        const double gam = nmax * 0x1p-53 / (1.0 - nmax * 0x1p-53);
        const double delta = (2.0 * gam * (ws.af + ws.ai) + 0x1p-52 * fabs(dH)) * (1.0 + 0x1p-20) + 0x1p-1070;
        accept = lj_metropolis(rng, dH, delta, st->kT, &amb) != 0;
      }
    }
    o.ambiguous = (amb && lane == 0) ? 1u : 0u;
    if (accept && lane == 0) {
      // update_body_properties (microstate.rs:985-1034)
      double nsp[3] = {0.0, 0.0, 0.0};
      for (int a = 0; a < D; ++a) nsp[a] = wrap1(t[a] + 0.0, g.L[a]);
      const uint32_t fl = ghost_flags<D>(g, nsp);
      B.bpos[body] = make_double4(t[0], t[1], t[2], move_stamp(st->epoch));
      const double4 nspk = pack_site(nsp, fl);
      B.spos[body] = nspk;
      C.pos64[slot] = nspk;
    }
  }
  o.accepted = (accept && lane == 0) ? 1u : 0u;
  if (w.tracing && lane == 0) {
    unsigned long long ti = atomicAdd((unsigned long long*)&st->trace_n, 1ull);
    if (ti < w.trace_cap) {
      hmc_trace_record rec;
      rec.substep = substep;
      rec.color = colour;
      rec.space = s;
      rec.body = (int64_t)body;
      rec.status = accept ? 1 : 0;
      rec.pad = 0;
      w.trace[ti] = rec;
    }
  }
}

// One colour of one cycle for LJ: one wavefront per space of the colour.
// (four wavefronts per SIMD: measured faster than three, five spills)
template <int D, int PREC>
__global__ void __launch_bounds__(kBlock, 4) k_pass_lj(Geometry g, ApplyState* st, Bodies B, Cells C, Work w,
                                                    uint32_t colour, uint32_t cycle, uint32_t xcd_chunk) {
  if (cycle >= st->cycles) return;  // uniform over the grid
  __shared__ double zx[257], zf[257];
  __shared__ unsigned blk_acc;
  __shared__ uint32_t tab[kBlock / 64][kLjTab];
  if (D == 3) stage_zig(zx, zf);
  if (threadIdx.x == 0) blk_acc = 0;
  __syncthreads();
  const uint32_t b = blockIdx.x;
  const uint32_t lb = (b & 7u) * xcd_chunk + (b >> 3);
  const uint32_t t = lb * (kBlock / 64) + (threadIdx.x >> 6);  // one space per wavefront
  LjOut o{0, 0, 0, 0.0, 0.0};
  if (t < g.per_colour) {
    int32_t m[3] = {0, 0, 0};
    const uint32_t s = colour_space<D>(g, colour, t, m);
    const uint32_t s0 = w.start[s];
    const uint32_t n = w.start[s + 1] - s0;
    if (n) trial_lj<D, PREC>(g, st, B, C, w, zx, zf, s, m, s0, n, colour, st->substep0 + cycle, o, tab[threadIdx.x >> 6]);
  }
  if (o.accepted) atomicAdd(&blk_acc, 1u);
  lane_add(&st->ambiguous, o.ambiguous);
  lane_add(&st->band_retests, o.retests);
  if (g.lj_fp32_error && __lane_id() == 0 && o.err > 0.0) {  // non-negative doubles order as their bits
    atomicMax(&st->fp32_err_max, (unsigned long long)__double_as_longlong(o.err));
    atomicMax(&st->fp32_err_ratio, (unsigned long long)__double_as_longlong(o.ratio));
  }
  __syncthreads();
  if (threadIdx.x == 0 && blk_acc) atomicAdd(&w.acc_slots[b & (kAccSlots - 1)], (unsigned long long)blk_acc);
}

// LJ cycles beyond the ones the host launched, in one workgroup.
template <int D>
__global__ void __launch_bounds__(kGuardThreads) k_guard_lj(Geometry g, ApplyState* st, Bodies B, Cells C, Work w,
                                                            uint32_t first_cycle) {
  const uint32_t cycles = st->cycles;
  if (first_cycle >= cycles) return;
  __shared__ double zx[257], zf[257];
  if (D == 3) stage_zig(zx, zf);
  __syncthreads();
  const uint32_t waves = kGuardThreads / 64, wid = threadIdx.x >> 6;
  for (uint32_t k = first_cycle; k < cycles; ++k) {
    for (uint32_t c = 0; c < (1u << D); ++c) {
      for (uint32_t t = wid; t < g.per_colour; t += waves) {
        LjOut o{0, 0, 0, 0.0, 0.0};
        int32_t m[3] = {0, 0, 0};
        const uint32_t s = colour_space<D>(g, c, t, m);
        const uint32_t s0 = w.start[s];
        const uint32_t n = w.start[s + 1] - s0;
        if (n) trial_lj<D, HMC_FP64>(g, st, B, C, w, zx, zf, s, m, s0, n, c, st->substep0 + k, o, nullptr);
        if (o.accepted) atomicAdd(&w.acc_slots[0], 1ull);
        lane_add(&st->ambiguous, o.ambiguous);
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) st->guard_cycles += 1;
  }
}

// explicit instantiations (launched from the host translation units)
template __global__ void k_pass_lj<2, HMC_MIXED>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, uint32_t xcd_chunk);
template __global__ void k_pass_lj<2, HMC_FP64>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, uint32_t xcd_chunk);
template __global__ void k_pass_lj<3, HMC_MIXED>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, uint32_t xcd_chunk);
template __global__ void k_pass_lj<3, HMC_FP64>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, uint32_t xcd_chunk);
template __global__ void k_guard_lj<2>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t first_cycle);
template __global__ void k_guard_lj<3>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t first_cycle);

}  // namespace hmc
