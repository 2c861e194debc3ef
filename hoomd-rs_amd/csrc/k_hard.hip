// Hard-shape colour passes (spheres and disks): the lane-per-space pass,
// the LDS-staged tile pass and the guard, for one system and for a batch.
#include "sweep_device.hpp"

namespace hmc {

// One trial in the non-empty space s (multi-index m) with slots [s0, s0 + n)
// (parallel_sweep.rs:234-277).
//
// FAST (every axis has >= 3 spaces and the scan reaches one space out): the
// 3^D neighbour ranges and the first body of the space are loaded before the
// random draws and the proposal arithmetic, so their latency is hidden; the
// neighbours are then tested in rounds -- round k tests the k-th slot of
// every neighbour space -- whose loads are all issued before any is tested.
// Otherwise a general loop visits the de-duplicated neighbour spaces (reach 1
// or 2) one by one.
template <int D, int PREC, bool FAST>
__device__ __forceinline__ void trial(const Geometry& g, const ApplyState* st, const Bodies& B, const Cells& C,
                                      const Work& w, const double* zx, const double* zf, uint32_t s,
                                      const int32_t* m, uint32_t s0, uint32_t n, uint32_t colour, uint32_t substep,
                                      TrialOut& o) {
  constexpr int NC = D == 3 ? 27 : 9;
  using VT = typename CellCopy<PREC>::type;
  uint32_t nb[3][5];
  int nn[3] = {1, 1, 1};
  for (int a = 0; a < D; ++a) {
    const int32_t sh = g.shape[a];
    const int32_t r = g.reach[a];
    if (FAST || (r == 1 && sh >= 3)) {
      nb[a][0] = (uint32_t)(m[a] == 0 ? sh - 1 : m[a] - 1) * g.stride[a];
      nb[a][1] = (uint32_t)m[a] * g.stride[a];
      nb[a][2] = (uint32_t)(m[a] == sh - 1 ? 0 : m[a] + 1) * g.stride[a];
      nn[a] = 3;
    } else {
      int k = 0;
      for (int off = -r; off <= r; ++off) {
        int32_t mm = ((m[a] + off) % sh + sh) % sh;
        bool seen = false;
        for (int j = 0; j < k; ++j) seen |= nb[a][j] == (uint32_t)mm * g.stride[a];
        if (!seen) nb[a][k++] = (uint32_t)mm * g.stride[a];
      }
      nn[a] = k;
    }
  }
  uint32_t lo[FAST ? NC : 1], cnt[FAST ? NC : 1];
  if (FAST) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int i0 = D == 3 ? c / 9 : c / 3;
      const int i1 = D == 3 ? (c / 3) % 3 : c % 3;
      const int i2 = D == 3 ? c % 3 : 0;
      const uint32_t ns = nb[0][i0] + nb[1][i1] + (D == 3 ? nb[2][i2] : 0u);
      lo[c] = w.start[ns];
      cnt[c] = w.start[ns + 1];
    }
  }
  // the first body of the space (chosen whenever the space holds one body)
  const uint32_t body0 = C.body[s0];
  const double4 bp0 = B.bpos[body0];

  Sfc64 rng = counter_rng(st->step, substep, st->seed, s);
  const uint32_t pick = choose_index(rng, n);
  double v[3] = {0.0, 0.0, 0.0};
  ball_sample<D>(rng, st->max_move, zx, zf, v);
  const uint32_t slot = s0 + pick;
  uint32_t body = body0;
  double4 bp = bp0;
  if (pick != 0) {
    body = C.body[slot];
    bp = B.bpos[body];
  }
  double x[3] = {bp.x, bp.y, bp.z};
  for (int a = 0; a < D; ++a) x[a] += v[a];
  for (int a = 0; a < D; ++a) x[a] = wrap1(x[a], g.L[a]);  // generate_trial_moves: wrap(propose(..))

  bool accept = false;
  if (space_of<D>(g, st->origin, x) == s) {
    // trial site: wrap(body.transform(site)) = wrap(x + 0) (pairwise_cutoff.rs:436);
    // x + 0.0 differs from x only for -0.0, which wraps to the same value
    double t[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < D; ++a) t[a] = wrap1(x[a], g.L[a]);
    float t32[3] = {0.0f, 0.0f, 0.0f};
    for (int a = 0; a < D; ++a) t32[a] = (float)t[a];
    bool hit = false;
    unsigned retests = 0;
    if (FAST) {
      uint32_t rounds = 0;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        cnt[c] -= lo[c];
        rounds = max(rounds, cnt[c]);
      }
      // loads in flight per batch: all neighbours for the FP32 copy, a third
      // of them for the FP64 copy (register budget)
      constexpr int CH = 9;
      for (uint32_t k = 0; k < rounds && !hit; ++k) {
#pragma unroll
        for (int c0 = 0; c0 < NC; c0 += CH) {
          VT vq[CH];
#pragma unroll
          for (int j = 0; j < CH; ++j) {
            const int c = c0 + j;
            const uint32_t q = k < cnt[c] ? lo[c] + k : slot;  // a valid slot either way
            vq[j] = CellCopy<PREC>::load(C, q);
          }
#pragma unroll
          for (int j = 0; j < CH; ++j) {
            const int c = c0 + j;
            const uint32_t q = lo[c] + k;
            const bool valid = k < cnt[c] && q != slot;
            if (PREC == HMC_MIXED) {
              const int r = overlap32<D>(g, t32, *(const float4*)&vq[j]);
              hit |= valid && r == 1;
              if (valid && r == 2) {
                ++retests;
                hit |= retest64<D>(g.self, B.spos, C.body, q, t[0], t[1], t[2]);
              }
            } else {
              hit |= valid && overlap64<D>(g, t, *(const double4*)&vq[j]);
            }
          }
        }
      }
    } else {
      for (int i0 = 0; i0 < nn[0] && !hit; ++i0) {
        for (int i1 = 0; i1 < nn[1] && !hit; ++i1) {
          for (int i2 = 0; i2 < (D == 3 ? nn[2] : 1) && !hit; ++i2) {
            const uint32_t ns = nb[0][i0] + nb[1][i1] + (D == 3 ? nb[2][i2] : 0u);
            const uint32_t q0 = w.start[ns], q1 = w.start[ns + 1];
            for (uint32_t q = q0; q < q1 && !hit; ++q)
              if (q != slot) hit = pair_hits<D, PREC>(g, B, C, q, t, t32, retests);
          }
        }
      }
    }
    o.retests = retests;
    // Hard spheres: dH is 0 or +inf, so u is never drawn (parallel_sweep.rs:267-273).
    accept = !hit;
    if (accept) {
      // update_body_properties (microstate.rs:985-1034): body = wrap(props),
      // which is t; site = wrap(body + 0)
      double nsp[3] = {0.0, 0.0, 0.0};
      for (int a = 0; a < D; ++a) nsp[a] = wrap1(t[a] + 0.0, g.L[a]);
      const uint32_t fl = ghost_flags<D>(g, nsp);
      B.bpos[body] = make_double4(t[0], t[1], t[2], move_stamp(st->epoch));
      const double4 sp = pack_site(nsp, fl);
      B.spos[body] = sp;
      if (PREC == HMC_MIXED)
        C.pos32[slot] = pack_site32(nsp, fl);
      else
        C.pos64[slot] = sp;
    }
  }
  o.accepted = accept ? 1u : 0u;
  if (w.tracing) {
    unsigned long long idx = atomicAdd((unsigned long long*)&st->trace_n, 1ull);
    if (idx < w.trace_cap) {
      hmc_trace_record r;
      r.substep = substep;
      r.color = colour;
      r.space = s;
      r.body = (int64_t)body;
      r.status = accept ? 1 : 0;
      r.pad = 0;
      w.trace[idx] = r;
    }
  }
}

// One colour of one cycle: one lane per space of the colour.
template <int D, int PREC, bool FAST>
__global__ void __launch_bounds__(kBlock) k_pass(Geometry g, ApplyState* st, Bodies B, Cells C, Work w,
                                                 uint32_t colour, uint32_t cycle, uint32_t xcd_chunk) {
  if (cycle >= st->cycles) return;  // uniform over the grid
  __shared__ unsigned blk_acc;
  __shared__ double zx[257], zf[257];
  if (threadIdx.x == 0) blk_acc = 0;
  if (D == 3) stage_zig(zx, zf);
  __syncthreads();
  // XCD-aware block order: consecutive logical blocks (neighbouring spaces)
  // land on the same XCD and share its L2.
  const uint32_t b = blockIdx.x;
  const uint32_t lb = (b & 7u) * xcd_chunk + (b >> 3);
  const uint32_t t = lb * kBlock + threadIdx.x;
  TrialOut o{0, 0};
  if (t < g.per_colour) {
    int32_t m[3] = {0, 0, 0};
    const uint32_t s = colour_space<D>(g, colour, t, m);
    const uint32_t s0 = w.start[s];
    const uint32_t n = w.start[s + 1] - s0;
    if (n) trial<D, PREC, FAST>(g, st, B, C, w, zx, zf, s, m, s0, n, colour, st->substep0 + cycle, o);
  }
  const unsigned long long mk = __ballot(o.accepted != 0);
  if (__lane_id() == 0 && mk) atomicAdd(&blk_acc, (unsigned)__popcll(mk));
  lane_add(&st->band_retests, o.retests);
  __syncthreads();
  if (threadIdx.x == 0 && blk_acc) atomicAdd(&w.acc_slots[b & (kAccSlots - 1)], (unsigned long long)blk_acc);
}

// The trial of space s from global memory, for a tile whose box holds more
// bodies than its LDS copy (rare).
template <int D, int PREC>
__device__ __forceinline__ TrialOut trial_from_global(const Geometry& g, const ApplyState* st, const Bodies& B,
                                                   const Cells& C, const Work& w, const double* zx, const double* zf,
                                                   uint32_t s, uint32_t s0, uint32_t n, uint32_t colour,
                                                   uint32_t substep) {
  int32_t m[3] = {0, 0, 0};
  for (int a = 0; a < D; ++a) m[a] = (int32_t)((s / g.stride[a]) % (uint32_t)g.shape[a]);
  TrialOut o{0, 0};
  trial<D, PREC, false>(g, st, B, C, w, zx, zf, s, m, s0, n, colour, substep, o);
  return o;
}

// The pass of one tile (the body of k_pass_tile and k_bpass_tile): g and tc
// are the system's, w.start its spaces' offsets; the block's accepted moves
// are added to *acc.  FP32_ERROR (mixed): every pair within 1.01 sigma is
// also measured in FP64 and the largest |d2_32 - d2_64| and its ratio to the
// band's error bound are recorded in the state.
template <int D, int PREC, bool FP32_ERROR = false>
__device__ __forceinline__ void pass_tile(const Geometry& g, ApplyState* st, const Bodies& B, const Cells& C,
                                          const Work& w, uint32_t colour, uint32_t cycle, const TileCfg& tc,
                                          uint32_t tile, unsigned long long* acc) {
  using VT = typename CellCopy<PREC>::type;
  constexpr int AL = D - 1;  // the row axis (stride 1)
  if (cycle >= st->cycles) return;  // uniform over the grid
  if (tile >= tc.ntiles) return;    // uniform over the block
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t R = tc.rows, EL = tc.E[AL], CW = EL + 1;
  // per row: LDS offset, first run's start slot and length, second run's
  // start slot; ri[R].x = the number of staged sites
  uint4* ri = (uint4*)smem;
  uint32_t* coff = (uint32_t*)(smem + 16u * (R + 1));  // per row: LDS offsets of its E + 1 cell boundaries
  VT* lb = (VT*)(smem + ((16u * (R + 1) + 4u * R * CW + 15u) & ~15u));
  __shared__ double zx[257], zf[257];
  __shared__ unsigned blk_acc;
  const uint32_t tid = threadIdx.x, NT = blockDim.x;
  const uint32_t wid = tid >> 6, lane = tid & 63u;
  const uint32_t substep = st->substep0 + cycle;
  // the apply's parameters, read once here: the state is written by atomics,
  // so these are vector loads, and read later they would wait behind the
  // staging loads
  const uint64_t step = st->step;
  const uint32_t seed = st->seed;
  const double max_move = st->max_move;
  double origin[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < D; ++a) origin[a] = st->origin[a];

  uint32_t tcd[3] = {0, 0, 0};
  {
    uint32_t r = tile;
    for (int a = D - 1; a >= 0; --a) {
      tcd[a] = r % tc.nt[a];
      r /= tc.nt[a];
    }
  }
  int32_t mlo[3] = {0, 0, 0};  // box origin: 2 T tcd + offset - 1 per axis
  for (int a = 0; a < D; ++a) mlo[a] = (int32_t)(2 * tcd[a] * tc.T[a] + ((colour >> (D - 1 - a)) & 1u)) - 1;
  // this thread's trial space and its centre cell in the box
  bool active = tid < tc.trials;
  uint32_t cc[3] = {0, 0, 0};
  uint32_t s = 0;
  {
    uint32_t r = tid;
    for (int a = D - 1; a >= 0; --a) {
      const uint32_t l = r % tc.T[a];
      r /= tc.T[a];
      const uint32_t h = tcd[a] * tc.T[a] + l;
      if (h >= g.half_shape[a]) active = false;
      s += (2 * h + ((colour >> (D - 1 - a)) & 1u)) * g.stride[a];
      cc[a] = 2 * l + 1;
    }
  }
  // along a row, cells [0, p) are spaces m0 .. m0 + p - 1 and cells [p, E)
  // wrap round to spaces 0 .. E - p - 1 (p = E: no wrap)
  const int32_t SL = g.shape[AL], ml = mlo[AL];
  const uint32_t p = ml < 0 ? 1u : (uint32_t)min((int32_t)EL, SL - ml);
  const uint32_t m0 = (uint32_t)(ml < 0 ? ml + SL : ml);

  // (A) the own range, the row starts, the runs of each row (wave 0), the
  // ziggurat tables
  uint32_t s0 = 0, s1 = 0;
  if (active) {
    s0 = w.start[s];
    s1 = w.start[s + 1];
  }
  // wave 0: the two runs of each of its rows (all loads issued before any is
  // used; rows past the box load row R - 1 again)
  const uint32_t rpl = (R + 63) / 64;  // rows per lane of wave 0 (uniform)
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
  // value j of row r is the start of space m0 + j (j <= p) or j - p - 1
  // (j > p) of the row; thread tid takes values tid + i NT, walking the
  // digits (row digits, j) by the workgroup size
  const uint32_t nvals = (R * (EL + 2) + NT - 1) / NT;  // uniform
  uint32_t raw[kValsPerThread], key[kValsPerThread];
  {
    uint32_t d[3] = {0, 0, 0};
    uint32_t r = tid;
    for (int a = AL; a > 0; --a) {
      const uint32_t rad = a == AL ? EL + 2 : tc.E[a];
      d[a] = r % rad;
      r /= rad;
    }
    d[0] = r;  // the leading digit is not reduced: d[0] >= E[0] past the box
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
  if (D == 3) stage_zig(zx, zf);
  if (tid == 0) blk_acc = 0;
  const uint32_t n = s1 - s0;
  const bool go = active && n != 0;
  // body indices of the own space (the chosen one is among them unless
  // pick >= 3); three scalars, so that the loads stay in flight together
  uint32_t ob0 = 0, ob1 = 0, ob2 = 0;
  if (go) {
    ob0 = C.body[s0];
    if (n > 1) ob1 = C.body[s0 + 1];
    if (n > 2) ob2 = C.body[s0 + 2];
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
  // everything of (A) arrives before the barrier, so that no wait in (B) has
  // to count these loads
  uint32_t held = ob0 ^ ob1 ^ ob2 ^ seed ^ (uint32_t)step ^ (uint32_t)__double2hiint(max_move);
#pragma unroll
  for (int i = 0; i < kValsPerThread; ++i) held ^= raw[i];
  asm volatile("" ::"v"(held));
  __syncthreads();

  // (B) staging, with the proposal drawn and the cell offsets written while
  // the first loads are in flight
  const uint32_t total = ri[R].x;
  const bool overflow = total > tc.maxb;  // uniform over the block
  float c32[3] = {0.0f, 0.0f, 0.0f};
  if (PREC == HMC_MIXED) {
    for (int a = 0; a < D; ++a) {
      double c = origin[a] + ((double)mlo[a] + 0.5 * (double)tc.E[a]) * g.width[a];
      c -= g.L[a] * floor((c + g.half[a]) / g.L[a]);
      c32[a] = (float)c;
    }
  }
  // a half-wave per row: entry e of row r comes from slot b0 + e (e < l0) or
  // b1 + e - l0; the first 32 entries of RB rows are loaded together
  constexpr int RB = kRowBatch<PREC>;
  const uint32_t hw = tid >> 5, hl = tid & 31u, NH = NT >> 5;
  VT tmp[RB];
  uint32_t dst[RB], src[RB];
  uint32_t longrows = 0;  // rows of the batch with more than 32 entries
  auto issue = [&](uint32_t r0) {
    uint4 q[RB];
    uint32_t end[RB];
#pragma unroll
    for (int u = 0; u < RB; ++u) {
      const uint32_t r = min(r0 + (uint32_t)u * NH, R - 1);
      q[u] = ri[r];
      end[u] = ri[r + 1].x;
    }
    longrows = 0;
#pragma unroll
    for (int u = 0; u < RB; ++u) {
      const uint32_t len = r0 + (uint32_t)u * NH < R ? end[u] - q[u].x : 0u;
      dst[u] = 0xffffffffu;
      src[u] = 0u;
      if (hl < len) {
        src[u] = hl < q[u].z ? q[u].y + hl : q[u].w + (hl - q[u].z);
        dst[u] = q[u].x + hl;
      }
      if (len > 32u) longrows |= 1u << u;
      if (dst[u] != 0xffffffffu) tmp[u] = CellCopy<PREC>::load(C, src[u]);
    }
  };
  auto finish = [&](uint32_t r0) {
#pragma unroll
    for (int u = 0; u < RB; ++u)
      if (dst[u] != 0xffffffffu) lb[dst[u]] = stage_entry<D, PREC>(g, c32, tmp[u], src[u]);
    // entries past the first 32 of a row (rare)
    while (longrows) {
      const int u = __ffs(longrows) - 1;
      longrows &= longrows - 1;
      const uint32_t r = r0 + (uint32_t)u * NH;
      const uint4 q = ri[r];
      const uint32_t len = ri[r + 1].x - q.x;
      for (uint32_t e = 32u + hl; e < len; e += 32u) {
        const uint32_t sl = e < q.z ? q.y + e : q.w + (e - q.z);
        lb[q.x + e] = stage_entry<D, PREC>(g, c32, CellCopy<PREC>::load(C, sl), sl);
      }
    }
  };
  if (!overflow) issue(hw);
  // the proposal (parallel_sweep.rs:240-251): choose, then Translate
  Sfc64 rng = counter_rng(step, substep, seed, s);
  uint32_t pick = 0;
  double v[3] = {0.0, 0.0, 0.0};
  if (go && !overflow) {
    pick = choose_index(rng, n);
    ball_sample<D>(rng, max_move, zx, zf, v);
  }
  uint32_t body = pick == 0 ? ob0 : (pick == 1 ? ob1 : ob2);
  double4 bp = make_double4(0.0, 0.0, 0.0, 0.0);
  if (!overflow) {
    // LDS offsets of the cell boundaries: start - run start + run offset
#pragma unroll
    for (int i = 0; i < kValsPerThread; ++i) {
      if (key[i] != 0xffffffffu) {
        const uint32_t r = key[i] >> 8, j = key[i] & 0xffu;
        const uint4 q = ri[r];
        coff[r * CW + (j <= p ? j : j - 1)] = j <= p ? q.x + (raw[i] - q.y) : q.x + q.z + (raw[i] - q.w);
      }
    }
    finish(hw);
    // the chosen body's position, needed after the barrier
    if (go && pick > 2) body = C.body[s0 + pick];
    if (go) bp = B.bpos[body];
    for (uint32_t r0 = hw + NH * RB; r0 < R; r0 += NH * RB) {
      issue(r0);
      finish(r0);
    }
  }
  __syncthreads();

  // (C) the trials
  TrialOut o{0, 0};
  double err_max = 0.0, err_ratio = 0.0;  // FP32_ERROR
  if (go && overflow) {
    o = trial_from_global<D, PREC>(g, st, B, C, w, zx, zf, s, s0, n, colour, substep);
  } else if (go) {
    constexpr int NR = D == 3 ? 9 : 3;  // rows of three cells along the last axis
    uint32_t rlo[NR], rhi[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int i0 = D == 3 ? r / 3 : r;
      const int i1 = D == 3 ? r % 3 : 0;
      const uint32_t row = D == 3 ? (cc[0] + i0 - 1) * tc.E[1] + (cc[1] + i1 - 1) : cc[0] + i0 - 1;
      rlo[r] = coff[row * CW + cc[AL] - 1];
      rhi[r] = coff[row * CW + cc[AL] + 2];
    }
    const uint32_t crow = D == 3 ? cc[0] * tc.E[1] + cc[1] : cc[0];
    const uint32_t self = coff[crow * CW + cc[AL]] + pick;
    const uint32_t slot = s0 + pick;
    double x[3] = {bp.x, bp.y, bp.z};
    for (int a = 0; a < D; ++a) x[a] += v[a];
    for (int a = 0; a < D; ++a) x[a] = wrap1(x[a], g.L[a]);  // generate_trial_moves: wrap(propose(..))

    bool accept = false;
    if (space_of<D>(g, st->origin, x) == s) {
      // the trial site wrap(x + 0) (pairwise_cutoff.rs:436); x + 0.0 differs
      // from x only for -0.0, which wraps to the same value
      double t[3] = {0.0, 0.0, 0.0};
      for (int a = 0; a < D; ++a) t[a] = wrap1(x[a], g.L[a]);
      bool hit = false;
      unsigned retests = 0;
      uint32_t rounds = 0;
#pragma unroll
      for (int r = 0; r < NR; ++r) rounds = max(rounds, rhi[r] - rlo[r]);
      if (PREC == HMC_MIXED) {
        float t32[3] = {0.0f, 0.0f, 0.0f}, tr[3] = {0.0f, 0.0f, 0.0f};
        for (int a = 0; a < D; ++a) t32[a] = (float)t[a];
        box_relative<D>(g, c32, t32, tr);
        const float lo = tc.band_lo, hi = tc.band_hi;
        for (uint32_t k = 0; k < rounds && !hit; ++k) {
          uint32_t amb = 0;  // rows whose pair lies inside the error band
#pragma unroll
          for (int r = 0; r < NR; ++r) {
            const uint32_t idx = rlo[r] + k;
            const bool valid = idx < rhi[r] && idx != self;
            const float4 p = *(const float4*)&lb[valid ? idx : self];
            const float dx = p.x - tr[0], dy = p.y - tr[1];
            float d2 = dx * dx + dy * dy;
            if (D == 3) {
              const float dz = p.z - tr[2];
              d2 = d2 + dz * dz;
            }
            hit |= valid && d2 < lo;
            if (valid && !(d2 < lo) && !(d2 >= hi)) amb |= 1u << r;
            if constexpr (FP32_ERROR) {
              if (valid) {
                const double e2 = dist2_64<D>(g, t, B.spos[C.body[(uint32_t)__float_as_int(p.w)]]);
                if (e2 < 1.0201 * g.sigma2) {
                  const double err = fabs((double)d2 - e2);
                  err_max = fmax(err_max, err);
                  err_ratio = fmax(err_ratio, err / (g.sigma2 - (double)lo));
                }
              }
            }
          }
          // FP64 re-test of the band pairs (rare): one call site
          while (amb && !hit) {
            const int r = __ffs(amb) - 1;
            amb &= amb - 1;
            const float4 p = *(const float4*)&lb[rlo[r] + k];
            ++retests;
            hit |= retest64<D>(g.self, B.spos, C.body, (uint32_t)__float_as_int(p.w), t[0], t[1], t[2]);
          }
        }
      } else {
        for (uint32_t k = 0; k < rounds && !hit; ++k) {
#pragma unroll
          for (int r = 0; r < NR; ++r) {
            const uint32_t idx = rlo[r] + k;
            const bool valid = idx < rhi[r] && idx != self;
            const double4 p = *(const double4*)&lb[valid ? idx : self];
            hit |= valid && overlap64<D>(g, t, p);
          }
        }
      }
      o.retests = retests;
      accept = !hit;  // hard spheres: dH is 0 or +inf (parallel_sweep.rs:267-273)
      if (accept) {
        // update_body_properties (microstate.rs:985-1034): body = wrap(x),
        // which is t; site = wrap(body + 0)
        double nsp[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < D; ++a) nsp[a] = wrap1(t[a] + 0.0, g.L[a]);
        const uint32_t fl = ghost_flags<D>(g, nsp);
        B.bpos[body] = make_double4(t[0], t[1], t[2], move_stamp(st->epoch));
        const double4 sp = pack_site(nsp, fl);
        B.spos[body] = sp;
        if (PREC == HMC_MIXED)
          C.pos32[slot] = pack_site32(nsp, fl);
        else
          C.pos64[slot] = sp;
      }
    }
    o.accepted = accept ? 1u : 0u;
    if (w.tracing) {
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
  const unsigned long long mk = __ballot(o.accepted != 0);
  if (__lane_id() == 0 && mk) atomicAdd(&blk_acc, (unsigned)__popcll(mk));
  lane_add(&st->band_retests, o.retests);
  if (tid == 0 && overflow) atomicAdd(&st->tile_overflows, 1ull);
  if constexpr (FP32_ERROR) {
    if (err_max > 0.0) {  // non-negative doubles order as their bits
      atomicMax(&st->fp32_err_max, (unsigned long long)__double_as_longlong(err_max));
      atomicMax(&st->fp32_err_ratio, (unsigned long long)__double_as_longlong(err_ratio));
    }
  }
  __syncthreads();
  if (tid == 0 && blk_acc) atomicAdd(acc, (unsigned long long)blk_acc);
}

template <int D, int PREC, bool FP32_ERROR = false>
__global__ void __launch_bounds__(kTileThreads) k_pass_tile(Geometry g, ApplyState* st, Bodies B, Cells C, Work w,
                                                            uint32_t colour, uint32_t cycle, TileCfg tc) {
  const uint32_t b = blockIdx.x;
  const uint32_t tile = (b & 7u) * tc.xcd_chunk + (b >> 3);  // XCD-aware tile order
  pass_tile<D, PREC, FP32_ERROR>(g, st, B, C, w, colour, cycle, tc, tile, &w.acc_slots[b & (kAccSlots - 1)]);
}

// Runs the cycles beyond the ones the host launched, in one workgroup.
template <int D, int PREC>
__global__ void __launch_bounds__(kGuardThreads) k_guard(Geometry g, ApplyState* st, Bodies B, Cells C, Work w,
                                                         uint32_t first_cycle) {
  const uint32_t cycles = st->cycles;
  if (first_cycle >= cycles) return;
  __shared__ double zx[257], zf[257];
  if (D == 3) stage_zig(zx, zf);
  __syncthreads();
  const int ncol = 1 << D;
  const uint32_t rounds = (g.per_colour + kGuardThreads - 1) / kGuardThreads;
  for (uint32_t k = first_cycle; k < cycles; ++k) {
    for (int c = 0; c < ncol; ++c) {
      for (uint32_t r = 0; r < rounds; ++r) {
        const uint32_t t = r * kGuardThreads + threadIdx.x;
        TrialOut o{0, 0};
        if (t < g.per_colour) {
          int32_t m[3] = {0, 0, 0};
          const uint32_t s = colour_space<D>(g, (uint32_t)c, t, m);
          const uint32_t s0 = w.start[s];
          const uint32_t n = w.start[s + 1] - s0;
          if (n) trial<D, PREC, false>(g, st, B, C, w, zx, zf, s, m, s0, n, (uint32_t)c, st->substep0 + k, o);
        }
        wave_count(&w.acc_slots[0], o.accepted != 0);
        lane_add(&st->band_retests, o.retests);
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) st->guard_cycles += 1;
  }
}

// One colour of one cycle of every replica: one lane per space.
template <int D, int PREC, bool FAST>
__global__ void __launch_bounds__(kBlock) k_bpass(const BatchRep* reps, ApplyState* sts, const uint32_t* coff,
                                                  uint32_t R, uint32_t per_colour_total, Bodies B, Cells C, Work w,
                                                  uint32_t colour, uint32_t cycle, uint32_t xcd_chunk) {
  __shared__ double zx[257], zf[257];
  if (D == 3) stage_zig(zx, zf);
  __syncthreads();
  const uint32_t b = blockIdx.x;
  const uint32_t lb = (b & 7u) * xcd_chunk + (b >> 3);
  const uint32_t t = lb * kBlock + threadIdx.x;
  if (t >= per_colour_total) return;
  const uint32_t r = find_replica(coff, R, t);
  ApplyState* st = sts + r;
  if (cycle >= st->cycles) return;
  const BatchRep& rep = reps[r];
  const Geometry& g = rep.g;
  int32_t m[3] = {0, 0, 0};
  const uint32_t s = colour_space<D>(g, colour, t - rep.colour_off, m);
  Work wr = w;
  wr.start = w.start + rep.space_off;
  const uint32_t s0 = wr.start[s];
  const uint32_t n = wr.start[s + 1] - s0;
  if (!n) return;
  TrialOut o{0, 0};
  trial<D, PREC, FAST>(g, st, B, C, wr, zx, zf, s, m, s0, n, colour, st->substep0 + cycle, o);
  if (o.accepted) atomicAdd(&w.acc_slots[r], 1ull);
  lane_add(&st->band_retests, o.retests);
}

// One colour of one cycle of every replica on LDS tiles: block b takes tile
// b' = (b mod 8) chunk + b / 8 of the concatenated tile list (trep: each
// tile's replica, toff: the replicas' first tiles), copies that replica's
// geometry and tile shape into LDS and runs pass_tile on its spaces.
template <int D, int PREC>
__global__ void __launch_bounds__(kTileThreads) k_bpass_tile(const BatchRep* reps, const TileCfg* tcs,
                                                             const uint32_t* toff, const uint32_t* trep,
                                                             ApplyState* sts, uint32_t ntiles, uint32_t xcd_chunk,
                                                             Bodies B, Cells C, Work w, uint32_t colour,
                                                             uint32_t cycle) {
  __shared__ Geometry sg;
  __shared__ TileCfg stc;
  const uint32_t b = blockIdx.x;
  const uint32_t t = (b & 7u) * xcd_chunk + (b >> 3);
  if (t >= ntiles) return;  // uniform over the block
  const uint32_t r = trep[t];
  const uint32_t* gsrc = (const uint32_t*)&reps[r].g;
  for (uint32_t i = threadIdx.x; i < sizeof(Geometry) / 4; i += blockDim.x) ((uint32_t*)&sg)[i] = gsrc[i];
  const uint32_t* tsrc = (const uint32_t*)&tcs[r];
  for (uint32_t i = threadIdx.x; i < sizeof(TileCfg) / 4; i += blockDim.x) ((uint32_t*)&stc)[i] = tsrc[i];
  __syncthreads();
  Work wr = w;
  wr.start = w.start + reps[r].space_off;
  pass_tile<D, PREC>(sg, sts + r, B, C, wr, colour, cycle, stc, t - toff[r], &w.acc_slots[r]);
}

// Cycles of one replica beyond the ones launched (one workgroup per replica).
template <int D, int PREC>
__global__ void __launch_bounds__(kGuardThreads) k_bguard(const BatchRep* reps, ApplyState* sts, Bodies B, Cells C,
                                                          Work w, uint32_t first_cycle) {
  const uint32_t r = blockIdx.x;
  ApplyState* st = sts + r;
  const uint32_t cycles = st->cycles;
  if (first_cycle >= cycles) return;
  __shared__ double zx[257], zf[257];
  if (D == 3) stage_zig(zx, zf);
  __syncthreads();
  const Geometry& g = reps[r].g;
  Work wr = w;
  wr.start = w.start + reps[r].space_off;
  const uint32_t rounds = (g.per_colour + kGuardThreads - 1) / kGuardThreads;
  for (uint32_t k = first_cycle; k < cycles; ++k) {
    for (int c = 0; c < (1 << D); ++c) {
      for (uint32_t q = 0; q < rounds; ++q) {
        const uint32_t t = q * kGuardThreads + threadIdx.x;
        TrialOut o{0, 0};
        if (t < g.per_colour) {
          int32_t m[3] = {0, 0, 0};
          const uint32_t s = colour_space<D>(g, (uint32_t)c, t, m);
          const uint32_t s0 = wr.start[s];
          const uint32_t n = wr.start[s + 1] - s0;
          if (n) trial<D, PREC, false>(g, st, B, C, wr, zx, zf, s, m, s0, n, (uint32_t)c, st->substep0 + k, o);
        }
        wave_count(&w.acc_slots[r], o.accepted != 0);
        lane_add(&st->band_retests, o.retests);
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) st->guard_cycles += 1;
  }
}

// explicit instantiations (launched from the host translation units)
template __global__ void k_pass<2, HMC_MIXED, true>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, uint32_t xcd_chunk);
template __global__ void k_pass<2, HMC_MIXED, false>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, uint32_t xcd_chunk);
template __global__ void k_pass<2, HMC_FP64, true>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, uint32_t xcd_chunk);
template __global__ void k_pass<2, HMC_FP64, false>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, uint32_t xcd_chunk);
template __global__ void k_pass<3, HMC_MIXED, true>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, uint32_t xcd_chunk);
template __global__ void k_pass<3, HMC_MIXED, false>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, uint32_t xcd_chunk);
template __global__ void k_pass<3, HMC_FP64, true>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, uint32_t xcd_chunk);
template __global__ void k_pass<3, HMC_FP64, false>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, uint32_t xcd_chunk);
template __global__ void k_pass_tile<2, HMC_MIXED, false>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, TileCfg tc);
template __global__ void k_pass_tile<2, HMC_FP64, false>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, TileCfg tc);
template __global__ void k_pass_tile<3, HMC_MIXED, false>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, TileCfg tc);
template __global__ void k_pass_tile<3, HMC_FP64, false>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, TileCfg tc);
template __global__ void k_pass_tile<2, HMC_MIXED, true>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, TileCfg tc);
template __global__ void k_pass_tile<3, HMC_MIXED, true>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, TileCfg tc);
template __global__ void k_guard<2, HMC_MIXED>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t first_cycle);
template __global__ void k_guard<2, HMC_FP64>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t first_cycle);
template __global__ void k_guard<3, HMC_MIXED>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t first_cycle);
template __global__ void k_guard<3, HMC_FP64>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t first_cycle);
template __global__ void k_bpass<2, HMC_MIXED, true>(const BatchRep* reps, ApplyState* sts, const uint32_t* coff, uint32_t R, uint32_t per_colour_total, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, uint32_t xcd_chunk);
template __global__ void k_bpass<2, HMC_MIXED, false>(const BatchRep* reps, ApplyState* sts, const uint32_t* coff, uint32_t R, uint32_t per_colour_total, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, uint32_t xcd_chunk);
template __global__ void k_bpass<2, HMC_FP64, true>(const BatchRep* reps, ApplyState* sts, const uint32_t* coff, uint32_t R, uint32_t per_colour_total, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, uint32_t xcd_chunk);
template __global__ void k_bpass<2, HMC_FP64, false>(const BatchRep* reps, ApplyState* sts, const uint32_t* coff, uint32_t R, uint32_t per_colour_total, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, uint32_t xcd_chunk);
template __global__ void k_bpass<3, HMC_MIXED, true>(const BatchRep* reps, ApplyState* sts, const uint32_t* coff, uint32_t R, uint32_t per_colour_total, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, uint32_t xcd_chunk);
template __global__ void k_bpass<3, HMC_MIXED, false>(const BatchRep* reps, ApplyState* sts, const uint32_t* coff, uint32_t R, uint32_t per_colour_total, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, uint32_t xcd_chunk);
template __global__ void k_bpass<3, HMC_FP64, true>(const BatchRep* reps, ApplyState* sts, const uint32_t* coff, uint32_t R, uint32_t per_colour_total, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, uint32_t xcd_chunk);
template __global__ void k_bpass<3, HMC_FP64, false>(const BatchRep* reps, ApplyState* sts, const uint32_t* coff, uint32_t R, uint32_t per_colour_total, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, uint32_t xcd_chunk);
template __global__ void k_bpass_tile<2, HMC_MIXED>(const BatchRep* reps, const TileCfg* tcs, const uint32_t* toff, const uint32_t* trep, ApplyState* sts, uint32_t ntiles, uint32_t xcd_chunk, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle);
template __global__ void k_bpass_tile<2, HMC_FP64>(const BatchRep* reps, const TileCfg* tcs, const uint32_t* toff, const uint32_t* trep, ApplyState* sts, uint32_t ntiles, uint32_t xcd_chunk, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle);
template __global__ void k_bpass_tile<3, HMC_MIXED>(const BatchRep* reps, const TileCfg* tcs, const uint32_t* toff, const uint32_t* trep, ApplyState* sts, uint32_t ntiles, uint32_t xcd_chunk, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle);
template __global__ void k_bpass_tile<3, HMC_FP64>(const BatchRep* reps, const TileCfg* tcs, const uint32_t* toff, const uint32_t* trep, ApplyState* sts, uint32_t ntiles, uint32_t xcd_chunk, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle);
template __global__ void k_bguard<2, HMC_MIXED>(const BatchRep* reps, ApplyState* sts, Bodies B, Cells C, Work w, uint32_t first_cycle);
template __global__ void k_bguard<2, HMC_FP64>(const BatchRep* reps, ApplyState* sts, Bodies B, Cells C, Work w, uint32_t first_cycle);
template __global__ void k_bguard<3, HMC_MIXED>(const BatchRep* reps, ApplyState* sts, Bodies B, Cells C, Work w, uint32_t first_cycle);
template __global__ void k_bguard<3, HMC_FP64>(const BatchRep* reps, ApplyState* sts, Bodies B, Cells C, Work w, uint32_t first_cycle);

}  // namespace hmc
