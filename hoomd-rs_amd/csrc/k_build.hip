// Per-apply setup and the checkerboard build: origin draw, counting sort of
// the bodies into spaces (bin, scan, scatter, place), for one system and
// for a batch of replicas.
#include "sweep_device.hpp"

namespace hmc {

// ----------------------------------------------------------------- kernels --
// advance: 0 keep the counter, 1 increment_step first, 2 continue after
// the previous apply's substeps (the next trial of the same step).
// new_epoch: 1 for an apply (its build tags the look-back scan's status words
// with the new epoch), 0 for the other users of the build (energies, SDF,
// insertion energies), which scan with three kernels and leave the words and
// the epoch alone: a word tagged by an older apply must never match.
__global__ void k_setup(Geometry g, ApplyState* st, uint64_t step, uint32_t substep, uint32_t seed, double kT,
                        double max_move, int from_args, int advance, uint32_t move, int new_epoch) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  st->move = move;
  if (from_args) {
    st->step = step;
    st->substep0 = substep;
    st->seed = seed;
  } else if (advance == 1) {
    // Microstate::increment_step (microstate.rs:412-415)
    st->step += 1;
    st->substep0 = 0;
  } else if (advance == 2) {
    // the next trial of the same step: the substeps the previous apply used
    // (one per colour cycle, parallel_sweep.rs:414-419) follow on
    st->substep0 += st->cycles;
  }
  st->kT = kT;
  st->max_move = max_move;
  Sfc64 rng = counter_rng(st->step, st->substep0, st->seed, UINT64_MAX);
  for (int a = 0; a < g.dim; ++a) {
    double u = uniform01(rng);
    st->origin[a] = -g.L[a] / 2.0 - u * g.width[a];
  }
  st->cycles = 0;
  st->nonempty = 0;
  if (new_epoch) st->epoch += 1;
}

// After an apply of move `m` of a several-move step (hmc_run): the accepted
// moves (the sum of the accept slots) and trials since the last tally are
// added to that move's counters (m = kMaxMoves: clear them and only record
// the totals).  One workgroup of kAccSlots = 1024 threads.
__global__ void __launch_bounds__(1024) k_tally(ApplyState* st, const unsigned long long* acc_slots, uint32_t m) {
  __shared__ unsigned long long part[16];
  unsigned long long v = acc_slots[threadIdx.x];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long acc = 0;
    for (int i = 0; i < 16; ++i) acc += part[i];
    if (m < (uint32_t)kMaxMoves) {
      st->move_acc[m] += acc - st->tally_acc;
      st->move_trials[m] += st->trials_total - st->tally_trials;
    } else {  // m = kMaxMoves: start of a run, the per-move counters start at 0
      for (int i = 0; i < kMaxMoves; ++i) st->move_acc[i] = st->move_trials[i] = 0;
    }
    st->tally_acc = acc;
    st->tally_trials = st->trials_total;
  }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Space of every body (in body order) and its arrival rank in that space.
template <int D>
__global__ void __launch_bounds__(kBlock) k_bin(Geometry g, ApplyState* st, Bodies B, Work w) {
  uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= g.n_bodies) return;
  double4 p4 = B.bpos[b];
  double x[3] = {p4.x, p4.y, p4.z};
  uint32_t s = space_of<D>(g, st->origin, x);
  if (s == UINT32_MAX) {
    atomicOr(&st->error, 1u);
    s = 0;
  }
  w.elem_space[b] = s;
  w.elem_rank[b] = atomicAdd(&w.cnt[s], 1u);
}

// Exclusive scan within a workgroup through LDS.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds, uint32_t* total) {
  const int tid = threadIdx.x;
  const int nt = blockDim.x;
  lds[tid] = v;
  __syncthreads();
  for (int off = 1; off < nt; off <<= 1) {
    uint32_t t = tid >= off ? lds[tid - off] : 0;
    __syncthreads();
    lds[tid] += t;
    __syncthreads();
  }
  uint32_t incl = lds[tid];
  *total = lds[nt - 1];
  __syncthreads();
  return incl - v;
}

// Scan of cnt in blocks of 1024; also counts the non-empty spaces per block.
__global__ void __launch_bounds__(256) k_scan_blocks(const uint32_t* cnt, uint32_t* start, uint32_t* block_sums,
                                                    uint32_t* block_ne, uint32_t n) {
  __shared__ uint32_t lds[256];
  uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * 4;
  uint32_t v[4];
  uint32_t sum = 0, ne = 0;
  for (int k = 0; k < 4; ++k) {
    v[k] = (base + k < n) ? cnt[base + k] : 0;
    sum += v[k];
    ne += v[k] ? 1u : 0u;
  }
  uint32_t total;
  uint32_t off = block_exclusive_scan(sum, lds, &total);
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) start[base + k] = off;
    off += v[k];
  }
  uint32_t ne_total;
  block_exclusive_scan(ne, lds, &ne_total);
  if (threadIdx.x == 0) {
    block_sums[blockIdx.x] = total;
    block_ne[blockIdx.x] = ne_total;
  }
}

// Single workgroup: exclusive scan of the block sums (up to 1024 * 16 blocks),
// start[n] = N, and the colour-cycle count ceil(N / non-empty spaces).
__global__ void __launch_bounds__(1024) k_scan_top(Geometry g, ApplyState* st, uint32_t* block_sums,
                                                  const uint32_t* block_ne, uint32_t nblocks, uint32_t* start,
                                                  int count) {
  __shared__ uint32_t lds[1024];
  const int per = 16;
  uint32_t base = threadIdx.x * per;
  uint32_t v[per];
  uint32_t sum = 0, ne = 0;
  for (int k = 0; k < per; ++k) {
    const bool in = base + k < nblocks;
    v[k] = in ? block_sums[base + k] : 0;
    sum += v[k];
    ne += in ? block_ne[base + k] : 0;
  }
  uint32_t total;
  uint32_t off = block_exclusive_scan(sum, lds, &total);
  for (int k = 0; k < per; ++k) {
    if (base + k < nblocks) block_sums[base + k] = off;
    off += v[k];
  }
  uint32_t ne_total;
  block_exclusive_scan(ne, lds, &ne_total);
  if (threadIdx.x == 0) {
    start[g.n_spaces] = total;
    st->nonempty = ne_total;
    if (!count) return;  // a rebuild for an energy evaluation: no trials follow
    // ParallelSweep::apply loops colour cycles until count.total() >= N
    // (parallel_sweep.rs:402); each cycle counts one trial per non-empty space.
    const uint32_t cycles = ne_total ? (g.n_bodies + ne_total - 1) / ne_total : 0;
    st->cycles = cycles;
    st->cycles_total += cycles;
    st->trials_total += (unsigned long long)cycles * ne_total;
  }
}

// Single-pass scan of the per-space counts (decoupled look-back): block b
// scans its kScanBlock counts, publishes its aggregate, adds its
// predecessors' aggregates back to the first inclusive prefix and publishes
// its own.  Status word: bits 56-63 the apply's epoch (older words read as
// not ready), 54-55 the state (1 aggregate, 2 inclusive prefix), 27-53 the
// count sum, 0-26 the non-empty spaces (so n and the body count stay below
// 2^27).  Also clears the counts for the next build and, in the last block,
// forms the totals and the colour cycles as k_scan_top does.  Every
// predecessor of a waiting block was dispatched before it and never waits
// on a later block, so the waits end; a bounded spin reports
// kErrBarrier all the same.
__global__ void __launch_bounds__(256) k_scan_lookback(Geometry g, ApplyState* st, uint32_t* cnt, uint32_t* start,
                                                       unsigned long long* status, uint32_t n, uint32_t nblocks) {
  __shared__ uint32_t lds[256];
  __shared__ uint32_t pre[2];
  const uint32_t b = blockIdx.x;
  const unsigned long long tag = (unsigned long long)(st->epoch & 0xffu) << 56;
  const unsigned long long field = (1ull << kLookbackBits) - 1ull;
  const uint32_t base = b * kScanBlock + threadIdx.x * 4;
  uint32_t v[4];
  uint32_t sum = 0, ne = 0;
  for (int k = 0; k < 4; ++k) {
    v[k] = (base + k < n) ? cnt[base + k] : 0;
    sum += v[k];
    ne += v[k] ? 1u : 0u;
  }
  uint32_t total, ne_total;
  const uint32_t off = block_exclusive_scan(sum, lds, &total);
  block_exclusive_scan(ne, lds, &ne_total);
  if (threadIdx.x == 0 && b > 0)
    __hip_atomic_store(&status[b], tag | (1ull << 54) | ((unsigned long long)total << kLookbackBits) | ne_total,
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (threadIdx.x < 64) {
    // the first wavefront looks back over 64 predecessors at a time: lane l
    // reads block hi - l; the window's words up to the nearest inclusive
    // prefix are summed once they are all ready
    const uint32_t lane = threadIdx.x;
    uint32_t ps = 0, pn = 0, spins = 0;
    int32_t hi = (int32_t)b - 1;
    while (hi >= 0) {
      const int32_t j = hi - (int32_t)lane;
      unsigned long long w = tag | (2ull << 54);  // before block 0: an inclusive prefix of 0
      if (j >= 0) w = __hip_atomic_load(&status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t state = (uint32_t)(w >> 54) & 3u;
      const bool ready = (w & 0xff00000000000000ull) == tag && state != 0u;
      const unsigned long long incl = __ballot(ready && state == 2u);
      const unsigned long long notready = __ballot(!ready);
      const uint32_t first = incl ? (uint32_t)__ffsll((long long)incl) - 1u : 64u;  // nearest inclusive
      const unsigned long long need = first >= 63u ? ~0ull : ((2ull << first) - 1ull);  // lanes 0..first
      if (notready & need) {
        if (++spins > kBarrierSpins) {
          if (lane == 0) atomicOr(&st->error, kErrBarrier);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        continue;  // read the same window again
      }
      uint32_t a = lane <= first ? (uint32_t)((w >> kLookbackBits) & field) : 0u;
      uint32_t c = lane <= first ? (uint32_t)(w & field) : 0u;
#pragma unroll
      for (int d = 32; d >= 1; d >>= 1) {
        a += __shfl_xor(a, d, 64);
        c += __shfl_xor(c, d, 64);
      }
      ps += a;
      pn += c;
      if (first < 64u) break;
      hi -= 64;
    }
    if (lane == 0) {
      __hip_atomic_store(&status[b],
                         tag | (2ull << 54) | ((unsigned long long)(ps + total) << kLookbackBits) | (pn + ne_total),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      pre[0] = ps;
      pre[1] = pn;
    }
  }
  __syncthreads();
  uint32_t o = pre[0] + off;
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) {
      start[base + k] = o;
      cnt[base + k] = 0;  // the next build counts from zero
    }
    o += v[k];
  }
  if (b == nblocks - 1 && threadIdx.x == 0) {
    const uint32_t all = pre[0] + total, ne_all = pre[1] + ne_total;
    start[n] = all;
    st->nonempty = ne_all;
    // ParallelSweep::apply loops colour cycles until count.total() >= N
    // (parallel_sweep.rs:402); each cycle counts one trial per non-empty space.
    const uint32_t cycles = ne_all ? (g.n_bodies + ne_all - 1) / ne_all : 0;
    st->cycles = cycles;
    st->cycles_total += cycles;
    st->trials_total += (unsigned long long)cycles * ne_all;
  }
}

// Adds the block offsets; with cnt given, also clears the per-space counts
// (their last reader is k_scan_blocks) for the next build.
__global__ void k_scan_add(uint32_t* start, const uint32_t* block_sums, uint32_t* cnt, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    start[i] += block_sums[i / kScanBlock];
    if (cnt) cnt[i] = 0u;
  }
}

__global__ void k_scatter(Geometry g, Work w) {
  uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= g.n_bodies) return;
  w.tmp_body[w.start[w.elem_space[b]] + w.elem_rank[b]] = b;
}

// Order each space by body index and write the cell arrays (COPY: also the
// per-slot site copies the per-colour passes read).
template <int PREC, bool COPY = true>
__global__ void __launch_bounds__(kBlock) k_place(Geometry g, Bodies B, Cells C, Work w) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.n_bodies) return;
  const uint32_t b = w.tmp_body[p];
  const uint32_t s = w.elem_space[b];
  const uint32_t s0 = w.start[s], s1 = w.start[s + 1];
  uint32_t rank = 0;
  for (uint32_t q = s0; q < s1; ++q) rank += w.tmp_body[q] < b ? 1u : 0u;
  const uint32_t dst = s0 + rank;
  C.body[dst] = b;
  if (!COPY) return;
  if (C.quat64) C.quat64[dst] = B.squat[b];
  const double4 sp = B.spos[b];
  if (PREC == HMC_MIXED)
    C.pos32[dst] = site32_of(sp);
  else
    C.pos64[dst] = sp;
}

__global__ void k_bsetup(const BatchRep* reps, ApplyState* sts, uint32_t R, const uint64_t* step,
                         const uint32_t* seed, double kT, const double* max_move, int from_args, int advance) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const Geometry& g = reps[r].g;
  ApplyState* st = sts + r;
  st->move = 0;
  if (from_args) {
    st->step = step[r];
    st->substep0 = 0;
    st->seed = seed[r];
  } else if (advance) {
    st->step += 1;  // Microstate::increment_step (microstate.rs:412-415)
    st->substep0 = 0;
  }
  st->kT = kT;
  st->max_move = max_move[r];
  Sfc64 rng = counter_rng(st->step, st->substep0, st->seed, UINT64_MAX);
  for (int a = 0; a < g.dim; ++a) {
    const double u = uniform01(rng);
    st->origin[a] = -g.L[a] / 2.0 - u * g.width[a];
  }
  st->cycles = 0;
  st->nonempty = 0;
}

template <int D>
__global__ void __launch_bounds__(kBlock) k_bbin(const BatchRep* reps, ApplyState* sts, const uint32_t* boff,
                                                 uint32_t R, uint32_t n_total, Bodies B, Work w) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_total) return;
  const uint32_t r = find_replica(boff, R, b);
  const double4 p4 = B.bpos[b];
  const double x[3] = {p4.x, p4.y, p4.z};
  uint32_t s = space_of<D>(reps[r].g, sts[r].origin, x);
  if (s == UINT32_MAX) {
    atomicOr(&sts[r].error, 1u);
    s = 0;
  }
  const uint32_t gs = reps[r].space_off + s;
  w.elem_space[b] = gs;
  const uint32_t rank = atomicAdd(&w.cnt[gs], 1u);
  w.elem_rank[b] = rank;
  // a space's first body makes it non-empty: counted per wavefront when the
  // wavefront's first bodies all belong to one replica (bodies are in
  // replica order), else per lane
  const unsigned long long fm = __ballot(rank == 0);
  if (fm) {
    const int lead = __ffsll((long long)fm) - 1;
    const uint32_t rl = __shfl(r, lead, 64);
    if (__ballot(rank == 0 && r != rl) == 0ull) {
      if ((int)__lane_id() == lead) atomicAdd(&sts[rl].nonempty, (unsigned)__popcll(fm));
    } else if (rank == 0) {
      atomicAdd(&sts[r].nonempty, 1u);
    }
  }
}

// Colour cycles of each replica from its non-empty spaces (counted by k_bbin):
// ParallelSweep::apply loops until count.total() >= N (parallel_sweep.rs:402),
// each cycle one trial per non-empty space.
__global__ void k_bcycles(const BatchRep* reps, ApplyState* sts, uint32_t R) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  ApplyState* st = sts + r;
  const uint32_t ne_total = st->nonempty;
  const uint32_t n = reps[r].g.n_bodies;
  const uint32_t cycles = ne_total ? (n + ne_total - 1) / ne_total : 0;
  st->cycles = cycles;
  st->cycles_total += cycles;
  st->trials_total += (unsigned long long)cycles * ne_total;
}

// explicit instantiations (launched from the host translation units)
template __global__ void k_bin<2>(Geometry g, ApplyState* st, Bodies B, Work w);
template __global__ void k_bin<3>(Geometry g, ApplyState* st, Bodies B, Work w);
template __global__ void k_place<HMC_MIXED, true>(Geometry g, Bodies B, Cells C, Work w);
template __global__ void k_place<HMC_MIXED, false>(Geometry g, Bodies B, Cells C, Work w);
template __global__ void k_place<HMC_FP64, true>(Geometry g, Bodies B, Cells C, Work w);
template __global__ void k_place<HMC_FP64, false>(Geometry g, Bodies B, Cells C, Work w);
template __global__ void k_bbin<2>(const BatchRep* reps, ApplyState* sts, const uint32_t* boff, uint32_t R, uint32_t n_total, Bodies B, Work w);
template __global__ void k_bbin<3>(const BatchRep* reps, ApplyState* sts, const uint32_t* boff, uint32_t R, uint32_t n_total, Bodies B, Work w);

}  // namespace hmc
