// MI355X (gfx950) implementation of hoomd-mc's checkerboard-parallel local
// trial-move sweep for single-site point bodies in a periodic cuboid with a
// hard-sphere pair energy.  Exposed through the C ABI in include/hmc.h.
//
// Per apply (ParallelSweep::apply, hoomd-mc/src/parallel_sweep.rs:390-421):
//   setup   draw the checkerboard origin from Counter(step, substep, seed)
//           .index(u64::MAX) (parallel_sweep.rs:182, hypercuboid.rs:315-341)
//   build   counting sort of the bodies into checkerboard spaces: bin (atomic
//           histogram, body order) -> exclusive scan (which also counts the
//           non-empty spaces and so the number of colour cycles) -> scatter ->
//           place.  place orders each space by body index (the order of
//           ParallelSweep's per-space lists, parallel_sweep.rs:194-200) and
//           writes the cell arrays the passes read: the body index and the
//           site position (FP32 in mixed mode, FP64 otherwise) of every slot.
//   passes  for each colour cycle and colour: one lane per space of that
//           colour runs the whole trial (choose, propose, wrap, space check,
//           overlap scan of the 3^D neighbour spaces, accept, in-place write);
//           lanes of empty spaces idle.  Same-colour spaces never touch each
//           other's neighbour spaces, so in-place writes need no barriers.
//   guard   the host launches its best guess of the cycle count; one
//           single-workgroup kernel runs any cycles beyond it.
//
// Data in HBM:
//   bodies (body-index order, authoritative, FP64): body position double4,
//       site position double4 with the ghost-image flags in .w
//   cells (slot order = space order, rebuilt each apply): body index u32 and
//       the neighbour copy of the site (float4 mixed / double4 FP64)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/hmc.h"
#include "dev_math.hpp"
#include "dev_convex.hpp"
#include "dev_rng.hpp"

namespace hmc {

constexpr int kBlock = 256;
constexpr int kScanBlock = 1024;  // elements per scan block (256 threads x 4)
constexpr uint32_t kBarrierSpins = 1u << 22;  // bound of a device-side wait (about 0.2 s of sleeps)
constexpr uint32_t kErrBarrier = 2u;          // ApplyState::error: a device-side wait timed out
constexpr int kGuardThreads = 1024;
constexpr int kAccSlots = 1024;
constexpr int kMaxVertices = 32;  // ConvexPolyhedron = ConvexPolytope<3, 32>

// Constant description of one system, passed by value to the kernels.
struct Geometry {
  int dim;
  uint32_t n_bodies;
  uint32_t n_spaces;
  uint32_t per_colour;
  double L[3];
  double half[3];
  double ghost_hi[3];  // half - boundary_range: above it the image is at x - L
  double ghost_lo[3];  // -half + boundary_range: below it the image is at x + L
  float Lf[3];
  double width[3];
  int32_t shape[3];
  uint32_t stride[3];
  uint32_t half_shape[3];  // spaces of one colour along each axis
  int32_t reach[3];        // neighbour spaces scanned on each side (1, or 2 when width ~ sigma)
  double sigma2;
  double sigma;     // hard-sphere diameter (the interaction range)
  double edge2;     // sigma^2 (1 - 2^-30): at or above it the spatial-structure rule is applied
  double vc_width;  // VecCell cell width of the reference microstate
  double vc_k;      // VecCell stencil radius ceil(sigma / vc_width)
  float band_lo, band_hi;  // mixed mode: d2 < band_lo overlaps, d2 >= band_hi does not
  int model;               // HMC_HARD_SPHERE, HMC_LJ, HMC_LJ_SHIFTED, HMC_OVERLAP_PENALTY, ...
  double lj_eps4;          // 4 epsilon (lennard_jones.rs:110)
  double lj_sigma;
  double lj_rc2;           // r_cut^2 (isotropic.rs:151)
  double lj_rc2_edge;      // r_cut^2 (1 - 2^-30): at or above it the spatial-structure rule is applied
  double lj_shift;         // E(r_shift) for the shifted potential, else 0
  double op_half_k;        // overlap penalty: 0.5 k, maximum_allowed_overlap, epsilon_shoulder, delta
  double op_dmax, op_eps, op_delta;
  int lj_mixed;            // LJ pair energies in FP32 with an FP64 re-test inside the error bound
  int lj_fp32_error;       // mixed: also measure in FP64 (LJ: every trial's dH; hard spheres: near-contact
                           // pairs of the per-colour tile pass) and record the largest FP32 error
  const double* verts;     // convex polyhedron: vertices (x, y, z triples), device memory
  uint32_t nverts;
  double cv_rcut2;         // (R + R)^2, R the circumsphere radius (convex.rs:163-166)
  double cv_edge2;         // cv_rcut2 (1 - 2^-30): at or above it the spatial-structure rule is applied
  // A copy of this struct in device memory.  Out-of-line helpers receive this
  // pointer instead of the address of a kernel argument, which would make
  // every lane copy the whole struct to scratch memory at kernel entry.
  const Geometry* self;
};

// The isotropic pair energies, which run one trial per wavefront with a
// Metropolis decision on dH (the others are hard overlap tests).
__host__ __device__ inline bool is_isotropic(int model) {
  return model == HMC_LJ || model == HMC_LJ_SHIFTED || model == HMC_OVERLAP_PENALTY;
}

// Device-resident per-apply state.
struct ApplyState {
  uint64_t step;
  uint32_t substep0;
  uint32_t seed;
  double kT;
  double max_move;
  double origin[3];
  uint32_t cycles;
  uint32_t nonempty;
  uint32_t error;
  uint32_t move;      // 0 Translate, 1 Rotate<Versor>
  uint32_t epoch;     // applies since creation: the tag of the look-back scan's status words
  uint32_t pad;
  unsigned long long accepted_total;  // filled on the host from the slot counters
  unsigned long long trials_total;
  unsigned long long cycles_total;
  unsigned long long band_retests;
  unsigned long long guard_cycles;
  unsigned long long tile_overflows;
  unsigned long long ambiguous;  // decisions the error bound could not settle
  unsigned long long trace_n;
  unsigned long long fp32_err_max;    // bits of the largest |dH32 - dH64| (HMC_FLAG_FP32_ERROR)
  unsigned long long fp32_err_ratio;  // bits of the largest |dH32 - dH64| / bound
};

struct Bodies {
  double4* bpos;
  double4* spos;
  double4* bquat;  // oriented bodies: (s, x, y, z); null otherwise
  double4* squat;
};

struct Cells {
  uint32_t* body;
  float4* pos32;
  double4* pos64;
  double4* quat64;  // site orientation copy of oriented bodies
};

struct Work {
  uint32_t* cnt;
  uint32_t* start;
  uint32_t* elem_space;
  uint32_t* elem_rank;
  uint32_t* tmp_body;
  uint32_t* block_sums;
  uint32_t* block_ne;
  unsigned long long* acc_slots;  // accepted moves, one counter per pass block slot
  hmc_trace_record* trace;
  uint64_t trace_cap;
  int tracing;
  uint32_t* mbox;  // persistent apply: per region two pass counters (by parity), kMailboxStride words apart
  unsigned long long* scan_status;  // single-pass scan: per block its tagged aggregate or inclusive prefix
};

// ----------------------------------------------------------- small helpers --
// Periodic<Hypercuboid>::wrap (hoomd-microstate/src/boundary/periodic/cuboid.rs:90-103)
__device__ __forceinline__ double wrap1(double x, double L) {
  double lambda = x / L;
  lambda = lambda - round(lambda);
  if (lambda == 0.5) lambda = -0.5;
  return lambda * L;
}

// Ghost image flags of a site (cuboid.rs:108-161): 2 bits per axis,
// 1 = image at x + L, 2 = image at x - L; kNoImages alone when the site lies
// outside [-L/2, L/2) on any axis (the reference then places no ghosts).
constexpr uint32_t kNoImages = 0x100u;
template <int D>
__device__ __forceinline__ uint32_t ghost_flags(const Geometry& g, const double* x) {
  uint32_t f = 0;
  for (int a = 0; a < D; ++a) {
    if (x[a] < -g.half[a] || x[a] >= g.half[a]) return kNoImages;
    if (x[a] > g.ghost_hi[a])
      f |= 2u << (2 * a);
    else if (x[a] < g.ghost_lo[a])
      f |= 1u << (2 * a);
  }
  return f;
}

// HypercuboidCheckerboard::point_to_space_index for a periodic cuboid
// (hoomd-mc/src/hypercuboid.rs:86-127); returns UINT32_MAX for None.
template <int D>
__device__ __forceinline__ uint32_t space_of(const Geometry& g, const double* origin, const double* x) {
  uint32_t index = 0;
  for (int a = 0; a < D; ++a) {
    double p = x[a] - origin[a];
    double q = floor(p / g.width[a]);
    if (!(q >= 0.0)) return UINT32_MAX;
    if (q > 4.0e9) return UINT32_MAX;
    uint32_t m = (uint32_t)q;
    if (m == (uint32_t)g.shape[a]) m = 0;
    if (m > (uint32_t)g.shape[a]) return UINT32_MAX;
    index += m * g.stride[a];
  }
  return index;
}

// The t-th space of colour c, in the order of
// HypercuboidCheckerboard::construct_space_indices_by_color (hypercuboid.rs:231-255):
// multi-index 2h + offset(c), h running over the half grid, last axis fastest.
template <int D>
__device__ __forceinline__ uint32_t colour_space(const Geometry& g, uint32_t c, uint32_t t, int32_t* m) {
  uint32_t s = 0;
  for (int a = D - 1; a >= 0; --a) {
    const uint32_t h = t % g.half_shape[a];
    t /= g.half_shape[a];
    const uint32_t off = (c >> (D - 1 - a)) & 1u;
    m[a] = (int32_t)(2 * h + off);
    s += (2 * h + off) * g.stride[a];
  }
  return s;
}

__device__ __forceinline__ double4 pack_site(const double* x, uint32_t flags) {
  return make_double4(x[0], x[1], x[2], __longlong_as_double((long long)flags));
}
__device__ __forceinline__ float4 pack_site32(const double* x, uint32_t flags) {
  return make_float4((float)x[0], (float)x[1], (float)x[2], __int_as_float((int)flags));
}
__device__ __forceinline__ float4 site32_of(double4 s) {
  return make_float4((float)s.x, (float)s.y, (float)s.z, __int_as_float((int)__double_as_longlong(s.w)));
}

// Wave-aggregated count of a per-lane 0/1 flag.
__device__ __forceinline__ void wave_count(unsigned long long* dst, bool flag) {
  unsigned long long m = __ballot(flag);
  int leader = __ffsll((long long)__ballot(1)) - 1;
  if (m && (int)__lane_id() == leader) atomicAdd(dst, (unsigned long long)__popcll(m));
}
// Per-lane add of a rare count.
__device__ __forceinline__ void lane_add(unsigned long long* dst, unsigned v) {
  if (v) atomicAdd(dst, (unsigned long long)v);
}

// ----------------------------------------------------------------- kernels --
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
  } else if (advance) {
    // Microstate::increment_step (microstate.rs:412-415)
    st->step += 1;
    st->substep0 = 0;
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
constexpr uint32_t kLookbackBits = 27;
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

// Ball::sample (hoomd-vector/src/distribution.rs:39-70): the displacement of
// Translate::propose (hoomd-mc/src/translate/cartesian.rs:39-52).
template <int D>
__device__ __forceinline__ void ball_sample(Sfc64& rng, double r, const double* zx, const double* zf, double* v) {
  if (D == 2) {
    for (;;) {
      double c0 = uniform01(rng);
      double c1 = uniform01(rng);
      v[0] = (c0 * 2.0 - 1.0) * r;
      v[1] = (c1 * 2.0 - 1.0) * r;
      double n2 = fma(v[1], v[1], fma(v[0], v[0], 0.0));
      if (n2 < r * r) break;
    }
  } else {
    // three draws through one copy of the sampler (shift register, no array)
    double p0 = 0.0, p1 = 0.0, p2 = 0.0;
#pragma unroll 1
    for (int i = 0; i < 3; ++i) {
      double z = standard_normal(rng, zx, zf);
      p0 = p1;
      p1 = p2;
      p2 = z;
    }
    double norm = sqrt(fma(p2, p2, fma(p1, p1, fma(p0, p0, 0.0))));
    double q = r / norm;
    double sc = q * hmc_cbrt(uniform01(rng));
    v[0] = p0 * sc;
    v[1] = p1 * sc;
    v[2] = p2 * sc;
  }
}

// Does site j (FP64) overlap the trial site t?  Image per axis: the direct one
// or, where the reference holds a ghost, x + T; the smaller |d| (the overlap
// decision over all images equals the decision for that image).
//
// Within 2^-30 (relative) of contact the reference's answer also depends on
// which images its VecCell query visits: a key is visited only if its cell
// floor(c / w) is within the stencil radius of the trial site's cell on every
// axis (hoomd-spatial/src/vec_cell.rs:356-358, 710-733).  Away from contact
// the rule cannot change the answer, so it is evaluated only there.
template <int D>
__device__ __noinline__ double edge_r2(const Geometry* gp, double t0, double t1, double t2, double sx, double sy,
                                       double sz, double sw) {
  const Geometry& g = *gp;
  const double t[3] = {t0, t1, t2};
  uint32_t fl = (uint32_t)__double_as_longlong(sw);
  double xj[3] = {sx, sy, sz};
  double d2 = 0.0;
  for (int a = 0; a < D; ++a) {
    double ct = floor(t[a] / g.vc_width);
    bool any = false;
    double best = 0.0;
    if (fabs(floor(xj[a] / g.vc_width) - ct) <= g.vc_k) {
      best = t[a] - xj[a];
      any = true;
    }
    uint32_t f = (fl >> (2 * a)) & 3u;
    if (f) {
      double gx = xj[a] + (f == 1u ? g.L[a] : -g.L[a]);
      if (fabs(floor(gx / g.vc_width) - ct) <= g.vc_k) {
        double dg = t[a] - gx;
        if (!any || fabs(dg) < fabs(best)) best = dg;
        any = true;
      }
    }
    if (!any) return __longlong_as_double(0x7ff0000000000000ll);  // no visited image
    d2 = d2 + best * best;
  }
  return d2;
}

template <int D>
__device__ __forceinline__ bool overlap64_edge(const Geometry& g, const double* t, double4 sj) {
  return edge_r2<D>(g.self, t[0], t[1], D == 3 ? t[2] : 0.0, sj.x, sj.y, sj.z, sj.w) < g.sigma2;
}

template <int D>
__device__ __forceinline__ bool overlap64(const Geometry& g, const double* t, double4 sj) {
  uint32_t fl = (uint32_t)__double_as_longlong(sj.w);
  double xj[3] = {sj.x, sj.y, sj.z};
  double d2 = 0.0;
  for (int a = 0; a < D; ++a) {
    double d = t[a] - xj[a];
    uint32_t f = (fl >> (2 * a)) & 3u;
    if (f) {
      double gx = xj[a] + (f == 1u ? g.L[a] : -g.L[a]);
      double dg = t[a] - gx;
      if (fabs(dg) < fabs(d)) d = dg;
    }
    d2 = d2 + d * d;
  }
  if (d2 < g.edge2) return true;
  if (!(d2 < g.sigma2)) return false;
  return overlap64_edge<D>(g, t, sj);
}

// The squared distance of overlap64 (the image its test uses), for the FP32
// error report.
template <int D>
__device__ __forceinline__ double dist2_64(const Geometry& g, const double* t, double4 sj) {
  const uint32_t fl = (uint32_t)__double_as_longlong(sj.w);
  const double xj[3] = {sj.x, sj.y, sj.z};
  double d2 = 0.0;
  for (int a = 0; a < D; ++a) {
    double d = t[a] - xj[a];
    const uint32_t f = (fl >> (2 * a)) & 3u;
    if (f) {
      const double gx = xj[a] + (f == 1u ? g.L[a] : -g.L[a]);
      const double dg = t[a] - gx;
      if (fabs(dg) < fabs(d)) d = dg;
    }
    d2 = d2 + d * d;
  }
  return d2;
}

// FP32 filter: 1 overlap, 0 no overlap, 2 inside the error band.
template <int D>
__device__ __forceinline__ int overlap32(const Geometry& g, const float* t, float4 sj) {
  uint32_t fl = (uint32_t)__float_as_int(sj.w);
  float xj[3] = {sj.x, sj.y, sj.z};
  float d2 = 0.0f;
  for (int a = 0; a < D; ++a) {
    float d = t[a] - xj[a];
    uint32_t f = (fl >> (2 * a)) & 3u;
    if (f) {
      float gx = xj[a] + (f == 1u ? g.Lf[a] : -g.Lf[a]);
      float dg = t[a] - gx;
      if (fabsf(dg) < fabsf(d)) d = dg;
    }
    d2 = d2 + d * d;
  }
  if (d2 < g.band_lo) return 1;
  if (d2 >= g.band_hi) return 0;
  return 2;
}

// FP64 re-test of a pair inside the FP32 error band (rare; kept out of line
// so the unrolled neighbour scan stays small).
template <int D>
__device__ __noinline__ bool retest64(const Geometry* gp, const double4* spos, const uint32_t* cbody, uint32_t q,
                                      double t0, double t1, double t2) {
  const double t[3] = {t0, t1, t2};
  return overlap64<D>(*gp, t, spos[cbody[q]]);
}

// Overlap of the trial site with the site in slot q, in the context's
// precision (the FP32 filter re-tests its error band in FP64).
template <int D, int PREC>
__device__ __forceinline__ bool pair_hits(const Geometry& g, const Bodies& B, const Cells& C, uint32_t q,
                                          const double* t, const float* t32, unsigned& retests) {
  if (PREC == HMC_MIXED) {
    int r = overlap32<D>(g, t32, C.pos32[q]);
    if (r == 2) {
      ++retests;
      r = retest64<D>(g.self, B.spos, C.body, q, t[0], t[1], t[2]) ? 1 : 0;
    }
    return r != 0;
  }
  return overlap64<D>(g, t, C.pos64[q]);
}

struct TrialOut {
  unsigned accepted;
  unsigned retests;
};

template <int PREC>
struct CellCopy;
template <>
struct CellCopy<HMC_MIXED> {
  using type = float4;
  __device__ static type load(const Cells& C, uint32_t q) { return C.pos32[q]; }
};
template <>
struct CellCopy<HMC_FP64> {
  using type = double4;
  __device__ static type load(const Cells& C, uint32_t q) { return C.pos64[q]; }
};

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
      B.bpos[body] = make_double4(t[0], t[1], t[2], 0.0);
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

// Copy of the ziggurat tables in LDS (the normal sampler's table reads are
// data dependent; LDS keeps them off the global-memory latency path).
__device__ __forceinline__ void stage_zig(double* zx, double* zf) {
  for (int i = threadIdx.x; i < 257; i += blockDim.x) {
    zx[i] = kZigX[i];
    zf[i] = kZigF[i];
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

// ------------------------------------------------------- LDS-staged tiles --
// A workgroup runs the trials of a T0 x T1 (x T2) block of one colour's
// spaces.  Their 3^D neighbourhoods form a box of (2T+1)^D spaces.  A trial
// reads only the neighbours of its own space, which are never spaces of the
// running colour (every neighbour differs by one in some axis and the space
// counts are even), so nothing a trial reads changes during the pass and a
// snapshot of the box taken at its start is exactly what every trial must
// see.  (In the last, partial tile along an axis the box wraps onto spaces
// of other tiles; those cells are staged but no active trial reads them.)
//
// Spaces are numbered with the last axis fastest and the slots are grouped
// by space in that order, so every row of the box along the last axis is one
// contiguous run of slots, or two where the row wraps round the periodic
// edge.  The wrap falls at the same cell of every row of a box, and a box
// spans at most shape - 2 spaces, so there is at most one.  The pass is a
// chain of dependent phases, kept to two barriers:
//  (A) every row's E + 2 cell starts are loaded (the E + 1 boundaries of its
//      cells and the start of the second run), one coalesced sweep over the
//      box, while wave 0 loads the two runs of each row and scans the row
//      lengths into LDS offsets;
//  (B) a half-wave per row copies the row's runs into LDS with loads of
//      consecutive slots; the proposal is drawn while the first of those
//      loads are in flight; the loaded starts become LDS cell offsets;
//  (C) the trials test their neighbours from LDS.
//
// Mixed precision stages each site relative to the box centre c, as the
// image nearest c (the box spans less than L - 2 w on every axis, so that
// image is the only one within reach of any trial in the box and the
// difference of two staged positions is the minimum-image displacement).
// The FP32 test is then a plain difference; pairs inside its error band are
// re-tested with the FP64 sites and the reference's image rules.  A site the
// reference gives no images (outside [-L/2, L/2) after rounding) is staged
// as NaN, which always lands in the band.
struct TileCfg {
  uint32_t T[3];     // colour spaces per tile along each axis (1 for unused axes)
  uint32_t nt[3];    // tiles along each axis
  uint32_t E[3];     // staged box edge 2T + 1 (1 for unused axes)
  uint32_t trials;   // T0 T1 T2
  uint32_t ncell;    // E0 E1 E2
  uint32_t rows;     // rows of the box along the last axis (ncell / E[D-1])
  uint32_t maxb;     // body capacity of the LDS copy
  uint32_t ntiles;
  uint32_t xcd_chunk;
  uint32_t lds_bytes;
  float band_lo, band_hi;  // FP32 band of the box-relative test
  uint32_t step[3];        // the workgroup size in the mixed radix (E0, .., E[D-2], E[D-1] + 2) of the row starts
};

constexpr int kTileThreads = 256;       // upper bound of the workgroup size
constexpr int kValsPerThread = 16;      // row starts per thread: rows (E + 2) <= 16 x threads
constexpr int kRowsPerLane = 4;         // rows per lane of wave 0: rows <= 256
// rows per half-wave whose loads are in flight together (the kernel runs at
// two waves per SIMD, so registers are not the limit)
template <int PREC>
constexpr int kRowBatch = PREC == HMC_MIXED ? 16 : 8;
constexpr size_t kTileLds = 72 * 1024;  // dynamic LDS per workgroup

// Orders the LDS accesses of one wavefront (lanes read what other lanes of
// the same wavefront wrote; LDS operations of a wavefront complete in order).
__device__ __forceinline__ void wave_lds_order() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inclusive scan over the 64 lanes of a wavefront.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = (int)__lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

// Image of an FP32 position nearest the box centre c, relative to c.
template <int D>
__device__ __forceinline__ void box_relative(const Geometry& g, const float* c32, const float* x, float* r) {
#pragma unroll
  for (int a = 0; a < D; ++a) {
    float u = x[a] - c32[a];
    const float h = 0.5f * g.Lf[a];
    if (u >= h)
      u -= g.Lf[a];
    else if (u < -h)
      u += g.Lf[a];
    r[a] = u;
  }
}

// The LDS entry of a staged site: box-relative FP32 position with the source
// slot in w (mixed), or the FP64 copy itself.
template <int D, int PREC>
__device__ __forceinline__ typename CellCopy<PREC>::type stage_entry(const Geometry& g, const float* c32,
                                                                   typename CellCopy<PREC>::type v, uint32_t slot) {
  if constexpr (PREC == HMC_MIXED) {
    const float x[3] = {v.x, v.y, v.z};
    float r[3] = {0.0f, 0.0f, 0.0f};
    box_relative<D>(g, c32, x, r);
    if ((uint32_t)__float_as_int(v.w) & kNoImages) r[0] = __int_as_float(0x7fc00000);
    return make_float4(r[0], r[1], r[2], __int_as_float((int)slot));
  } else {
    return v;
  }
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
        B.bpos[body] = make_double4(t[0], t[1], t[2], 0.0);
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

// ------------------------------------------------ persistent apply (mixed) --
// One cooperative launch runs every cycle and colour of an apply.  Workgroup
// b owns a region of 2 T spaces per axis (the last region along an axis may
// be shorter) and keeps the region plus a one-space halo -- every cell its
// trials can read -- in LDS for the whole apply: staged once, its own cells
// updated in place by its accepted moves, and after the pass of colour c the
// halo cells of colour c (the only ones other workgroups have changed)
// reloaded from global memory.  A grid barrier separates the colour passes.
//
// The XCDs have separate L2 caches.  Writes are plain (they combine in the
// writer's L2); at each grid barrier every wave waits for its stores, and one
// lane per workgroup writes its XCD's L2 back before arriving (an agent-scope
// release fence) and invalidates its caches after the last arrival (an
// agent-scope acquire fence), so the site copies and FP64 sites written in
// one colour pass are what every workgroup reads in the next.
//
// An LDS entry holds the site relative to the box centre in FP32 (NaN for a
// site without images, as in the per-colour kernel) with the body index in w.
constexpr uint32_t kMailboxStride = 64;       // words between region mailboxes (256 B: separate channels)
constexpr int kRefreshHeld = 2;               // halo entries per thread loaded together in a refresh
constexpr int kStageRows = 4;                 // rows per half-wave staged together by the persistent apply

// Device-scope (agent) accesses to the FP64 sites, the only data one region
// writes and another reads within the apply: the XCDs have separate L2
// caches, and these bypass them where needed.
__device__ __forceinline__ uint64_t ld_dev(const void* p) {
  return __hip_atomic_load((uint64_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double4 ld_dev_d4(const double4* p) {
  const uint64_t* q = (const uint64_t*)p;
  return make_double4(__longlong_as_double((long long)ld_dev(q)), __longlong_as_double((long long)ld_dev(q + 1)),
                      __longlong_as_double((long long)ld_dev(q + 2)), __longlong_as_double((long long)ld_dev(q + 3)));
}
// Two 16-byte device-scope (sc1) stores.  Their completion is waited for
// explicitly (s_waitcnt 0) before the region publishes the pass.
__device__ __forceinline__ void st_dev_d4(double4* p, double4 v) {
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  const unsigned long long a = (unsigned long long)__double_as_longlong(v.x),
                           b = (unsigned long long)__double_as_longlong(v.y),
                           c = (unsigned long long)__double_as_longlong(v.z),
                           d = (unsigned long long)__double_as_longlong(v.w);
  const u32x4 lo = {(unsigned)a, (unsigned)(a >> 32), (unsigned)b, (unsigned)(b >> 32)};
  const u32x4 hi = {(unsigned)c, (unsigned)(c >> 32), (unsigned)d, (unsigned)(d >> 32)};
  asm volatile("global_store_dwordx4 %0, %1, off sc1" ::"v"(p), "v"(lo) : "memory");
  asm volatile("global_store_dwordx4 %0, %1, off offset:16 sc1" ::"v"(p), "v"(hi) : "memory");
}

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
  const bool live = tile < tc.ntiles;  // padding workgroups still take part in every grid barrier
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

// ------------------------------------------------------------ Lennard-Jones --
// Isotropic<LennardJones<12, 6>> and Isotropic<Shifted<LennardJones>>
// (isotropic.rs:147-156, lennard_jones.rs:107-111, shifted.rs:92-94):
// r = sqrt(r^2), s = sigma / r, E = (4 eps) (s^12 - s^6) - E(r_shift), with
// s^12 = s^4 s^8 and s^6 = s^2 s^4 as powi's square chain evaluates them
// (lj_shift is 0 for the unshifted potential, which subtracts exactly).
//
// Isotropic<Expanded<OverlapPenalty>> (overlap_penalty.rs:109-115,
// expanded.rs:62-66): x = r - delta, E = inf for x < -d_max,
// (0.5 k) x^2 + eps_shoulder for x < 0, else 0 (r_cut = delta).
__device__ __forceinline__ double lj_energy(const Geometry& g, double r2) {
  const double r = sqrt(r2);
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

// Squared distance from q to site sj over the images the reference's
// neighbour query visits: the per-axis nearest image (only one image can lie
// within r_cut <= L/2), and within 2^-30 of the cutoff the VecCell rule.
template <int D>
__device__ __forceinline__ double lj_r2(const Geometry& g, const double* q, double4 sj) {
  const uint32_t fl = (uint32_t)__double_as_longlong(sj.w);
  const double xj[3] = {sj.x, sj.y, sj.z};
  double d2 = 0.0;
  for (int a = 0; a < D; ++a) {
    double d = q[a] - xj[a];
    const uint32_t f = (fl >> (2 * a)) & 3u;
    if (f) {
      const double dg = q[a] - (xj[a] + (f == 1u ? g.L[a] : -g.L[a]));
      if (fabs(dg) < fabs(d)) d = dg;
    }
    d2 = d2 + d * d;
  }
  if (d2 >= g.lj_rc2_edge && d2 < g.lj_rc2) d2 = edge_r2<D>(g.self, q[0], q[1], D == 3 ? q[2] : 0.0, sj.x, sj.y, sj.z, sj.w);
  return d2;
}

struct WaveSums {
  double ef, af, ei, ai;  // final / initial energy and sum of |terms|
  uint32_t nf, ni;        // number of non-zero terms
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// FP32 pair energy of the mixed-precision LJ path: s2 = sigma^2 / r2 as
// fl(sigma^2) rcp(fl(r2)) (v_rcp_f32, within 1 ulp), then the reference's
// powers and form (lennard_jones.rs:107-111; shifted.rs:92-94).  Relative
// errors: s2 5u, s4 11u, s8 23u, s12 35u, s6 17u (u = 2^-24), so
// |e - e_exact| <= 40 u A with A = 4 eps (s12 + s6) + |shift|, and the
// computed *a (A in FP32) gives |e - e_exact| <= 48 u *a with room to spare.
__device__ __forceinline__ float lj_energy32(float sig2, float eps4, float shift, double r2, float* a) {
  const float s2 = sig2 * __builtin_amdgcn_rcpf((float)r2);
  const float s4 = s2 * s2, s8 = s4 * s4;
  const float s12 = s4 * s8, s6 = s2 * s4;
  *a = eps4 * (s12 + s6) + fabsf(shift);
  return eps4 * (s12 - s6) - shift;
}

// Metropolis decision of parallel_sweep.rs:267-273 for dH known to within
// +-delta of the reference's value (same terms, other summation order):
// accept iff dH <= 0 or u < exp(-dH / kT), u drawn only when dH > 0.
// Returns 1 accept, 0 reject.  A u clear of exp(q) for every q of the
// interval [dH - delta, dH + delta] is decided by the device's fast exp;
// otherwise the reference's own comparison u < exp(-dH / kT), with glibc's
// exp, decides at the GPU's dH, and *ambiguous is set when delta > 0 (the
// reference's dH may then lie on the other side).
__device__ __forceinline__ int lj_metropolis(Sfc64& rng, double dH, double delta, double kT, bool* ambiguous) {
  const double hi = dH + delta, lo = dH - delta;
  *ambiguous = false;
  if (hi <= 0.0) return 1;
  const double u = uniform01(rng);
  // exp(q) for q within the interval, with the roundings of -dH / kT and of
  // exp itself covered by a relative margin of 2^-45 (1 + |q|)
  const double qa = -hi / kT;
  const double ea = exp(qa);
  if (u < ea * (1.0 - 0x1p-45 * (1.0 + fabs(qa)))) return 1;
  if (lo > 0.0) {
    const double qb = -lo / kT;
    const double eb = exp(qb);
    if (u >= eb * (1.0 + 0x1p-45 * (1.0 + fabs(qb)))) return 0;
  }
  *ambiguous = delta > 0.0;
  if (dH <= 0.0) return 1;
  return u < glibc_exp(-dH / kT) ? 1 : 0;
}

struct LjOut {
  unsigned accepted;
  unsigned ambiguous;
  unsigned retests;  // mixed precision: decisions the FP32 bound left open, re-taken in FP64
  double err, ratio; // with lj_fp32_error: |dH32 - dH64| and its ratio to the bound
};

constexpr uint32_t kLjBatch = 4;  // candidate rounds of an LJ trial with their loads in flight together
constexpr uint32_t kLjTab = 1024;  // candidate slots of an LJ trial laid out in LDS (else searched)

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
      B.bpos[body] = make_double4(t[0], t[1], t[2], 0.0);
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

// ------------------------------------------------------ convex polyhedra --
// Image of site sj nearest q per axis (the one the reference's neighbour
// query can find within 2 R <= L/2), its squared distance
// distance_squared(q, image) (cartesian.rs:339-342) and, within 2^-30 of the
// circumsphere cutoff, the VecCell visitation rule as for hard spheres.
__device__ __noinline__ double4 convex_edge_image(const Geometry* gp, double q0, double q1, double q2, double sx,
                                                  double sy, double sz, double sw) {
  const Geometry& g = *gp;
  const double q[3] = {q0, q1, q2};
  const uint32_t fl = (uint32_t)__double_as_longlong(sw);
  const double xj[3] = {sx, sy, sz};
  double img[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < 3; ++a) {
    const double ct = floor(q[a] / g.vc_width);
    bool any = false;
    double best = 0.0, bx = 0.0;
    if (fabs(floor(xj[a] / g.vc_width) - ct) <= g.vc_k) {
      best = q[a] - xj[a];
      bx = xj[a];
      any = true;
    }
    const uint32_t f = (fl >> (2 * a)) & 3u;
    if (f) {
      const double gx = xj[a] + (f == 1u ? g.L[a] : -g.L[a]);
      if (fabs(floor(gx / g.vc_width) - ct) <= g.vc_k) {
        const double dg = q[a] - gx;
        if (!any || fabs(dg) < fabs(best)) {
          best = dg;
          bx = gx;
        }
        any = true;
      }
    }
    if (!any) return make_double4(__longlong_as_double(0x7ff0000000000000ll), 0.0, 0.0, 0.0);
    img[a] = bx;
  }
  double d2 = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double d = q[a] - img[a];
    d2 = d2 + d * d;
  }
  return make_double4(d2, img[0], img[1], img[2]);  // squared distance, image
}

__device__ __forceinline__ double convex_image(const Geometry& g, const double* q, double4 sj, v3* rj) {
  const uint32_t fl = (uint32_t)__double_as_longlong(sj.w);
  const double xj[3] = {sj.x, sj.y, sj.z};
  double img[3] = {xj[0], xj[1], xj[2]};
  for (int a = 0; a < 3; ++a) {
    const uint32_t f = (fl >> (2 * a)) & 3u;
    if (f) {
      const double gx = xj[a] + (f == 1u ? g.L[a] : -g.L[a]);
      if (fabs(q[a] - gx) < fabs(q[a] - xj[a])) img[a] = gx;
    }
  }
  double d2 = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double d = q[a] - img[a];
    d2 = d2 + d * d;
  }
  *rj = {img[0], img[1], img[2]};
  if (d2 >= g.cv_edge2 && d2 < g.cv_rcut2) {
    const double4 e = convex_edge_image(g.self, q[0], q[1], q[2], sj.x, sj.y, sj.z, sj.w);
    d2 = e.x;
    *rj = {e.y, e.z, e.w};
  }
  return d2;
}

// One trial on a hard convex polyhedron (parallel_sweep.rs:234-277) with
// Translate (move 0) or Rotate<Versor> (move 1; rotate/versor.rs:111-122:
// orientation = delta * old, position wrapped again).  One lane per trial.
__device__ void trial_convex(const Geometry& g, const ApplyState* st, const Bodies& B, const Cells& C,
                             const Work& w, const double* zx, const double* zf, const double* vx, uint32_t s,
                             const int32_t* m, uint32_t s0, uint32_t n, uint32_t colour, uint32_t substep,
                             TrialOut& o) {
  Sfc64 rng = counter_rng(st->step, substep, st->seed, s);
  const uint32_t pick = choose_index(rng, n);
  const uint32_t slot = s0 + pick;
  const uint32_t body = C.body[slot];
  const double4 bp = B.bpos[body];
  const double4 bq = B.bquat[body];
  double x[3] = {bp.x, bp.y, bp.z};
  quat q = {bq.x, {bq.y, bq.z, bq.w}};
  if (st->move == 0) {
    double v[3] = {0.0, 0.0, 0.0};
    ball_sample<3>(rng, st->max_move, zx, zf, v);
    for (int a = 0; a < 3; ++a) x[a] += v[a];
  } else {
    q = qmul(versor_displacement(rng, st->max_move, zx, zf), q);
  }
  for (int a = 0; a < 3; ++a) x[a] = wrap1(x[a], g.L[a]);  // wrap(propose(..))
  bool accept = false;
  if (space_of<3>(g, st->origin, x) == s) {
    // the trial site: wrap(x + rotate(q, 0)) with orientation q * identity
    // (property/oriented_point.rs:107-113)
    const v3 r0 = qrotate(q, v3{0.0, 0.0, 0.0});
    const double t[3] = {wrap1(x[0] + r0.x, g.L[0]), wrap1(x[1] + r0.y, g.L[1]), wrap1(x[2] + r0.z, g.L[2])};
    const quat qt = qmul(q, quat{1.0, {0.0, 0.0, 0.0}});
    const v3 tv = {t[0], t[1], t[2]};
    uint32_t nb[3][5];
    int nn[3] = {1, 1, 1};
    for (int a = 0; a < 3; ++a) {
      const int32_t sh = g.shape[a], r = g.reach[a];
      int k = 0;
      for (int off = -r; off <= r; ++off) {
        const int32_t mm = ((m[a] + off) % sh + sh) % sh;
        bool seen = false;
        for (int j = 0; j < k; ++j) seen |= nb[a][j] == (uint32_t)mm * g.stride[a];
        if (!seen) nb[a][k++] = (uint32_t)mm * g.stride[a];
      }
      nn[a] = k;
    }
    bool hit = false;
    for (int i0 = 0; i0 < nn[0] && !hit; ++i0)
      for (int i1 = 0; i1 < nn[1] && !hit; ++i1)
        for (int i2 = 0; i2 < nn[2] && !hit; ++i2) {
          const uint32_t ns = nb[0][i0] + nb[1][i1] + nb[2][i2];
          const uint32_t q0 = w.start[ns], q1 = w.start[ns + 1];
          for (uint32_t qs = q0; qs < q1 && !hit; ++qs) {
            if (qs == slot) continue;
            v3 rj;
            const double d2 = convex_image(g, t, C.pos64[qs], &rj);
            if (d2 < g.cv_rcut2) {
              const double4 qj4 = C.quat64[qs];
              hit = convex_overlap(vx, g.nverts, g.cv_rcut2, d2, tv, qt, rj, quat{qj4.x, {qj4.y, qj4.z, qj4.w}});
            }
          }
        }
    accept = !hit;  // hard shapes: dH is 0 or +inf (parallel_sweep.rs:267-273)
    if (accept) {
      // update_body_properties (microstate.rs:985-1034): body = wrap(props),
      // site = wrap(body.transform(site))
      double nb3[3], ns3[3];
      for (int a = 0; a < 3; ++a) nb3[a] = wrap1(x[a], g.L[a]);
      const double r0a[3] = {r0.x, r0.y, r0.z};
      for (int a = 0; a < 3; ++a) ns3[a] = wrap1(nb3[a] + r0a[a], g.L[a]);
      const uint32_t fl = ghost_flags<3>(g, ns3);
      B.bpos[body] = make_double4(nb3[0], nb3[1], nb3[2], 0.0);
      B.bquat[body] = make_double4(q.s, q.v.x, q.v.y, q.v.z);
      const double4 sp = pack_site(ns3, fl);
      const double4 sq = make_double4(qt.s, qt.v.x, qt.v.y, qt.v.z);
      B.spos[body] = sp;
      B.squat[body] = sq;
      C.pos64[slot] = sp;
      C.quat64[slot] = sq;
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

// Convex pass layout: kCvTrials trials per wavefront, one per group of
// kCvLanes lanes (the group draws its trial's proposal and lists its
// neighbour spaces); the candidate sites of all the wavefront's trials are
// tested against the circumsphere cutoff 64 at a time (kCvBatch rounds with
// their loads in flight together), the pairs within it are gathered in a
// per-wavefront list, and XenoCollide runs over that list 64 pairs at a time.
// Any overlap rejects a hard-shape trial, so the order of the pair tests does
// not change a decision (parallel_sweep.rs:267-273).
constexpr uint32_t kCvTrials = 4;
constexpr uint32_t kCvLanes = 64 / kCvTrials;
constexpr uint32_t kCvSegs = 128;  // neighbour spaces of one trial: at most 5^3
constexpr uint32_t kCvBatch = 4;
constexpr uint32_t kCvPairs = 384;  // listed pairs; tested once more than kCvPairs - 64 kCvBatch are pending
struct CvWave {
  uint32_t pre[kCvTrials][kCvSegs];  // index of the first candidate of each neighbour space, per trial
  uint32_t lo[kCvTrials][kCvSegs];   // its first slot
  uint32_t qs[kCvPairs];             // the listed pair's neighbour slot
  uint32_t tk[kCvPairs];             // and trial
  double t[kCvTrials][3];            // trial site position
  double qt[kCvTrials][4];           // trial site orientation
};

// XenoCollide over the np listed pairs; returns the trials (bits) that
// found an overlap, skipping pairs of trials already in `hits`.  NV: the
// vertex count when it is one of the specialised ones, else 0.
template <int NV>
__device__ __forceinline__ uint32_t cv_pair_tests(const Geometry& g, const Cells& C, const CvWave& W,
                                                  const double* vx, uint32_t np, uint32_t hits) {
  const uint32_t lane = __lane_id();
  for (uint32_t p0 = 0; p0 < np; p0 += 64) {
    const uint32_t p = p0 + lane;
    bool h = false;
    uint32_t kp = 0;
    if (p < np) {
      kp = W.tk[p];
      if (!((hits >> kp) & 1u)) {
        const uint32_t qs = W.qs[p];
        const double4 sj = C.pos64[qs];
        const double4 qj4 = C.quat64[qs];
        const double tq[3] = {W.t[kp][0], W.t[kp][1], W.t[kp][2]};
        v3 rj;
        (void)convex_image(g, tq, sj, &rj);  // the image found when the pair was listed
        const quat qt = {W.qt[kp][0], {W.qt[kp][1], W.qt[kp][2], W.qt[kp][3]}};
        // listed pairs are within the cutoff: 0 stands for their distance
        h = convex_overlap<NV>(vx, g.nverts, g.cv_rcut2, 0.0, v3{tq[0], tq[1], tq[2]}, qt, rj,
                           quat{qj4.x, {qj4.y, qj4.z, qj4.w}});
      }
    }
#pragma unroll
    for (uint32_t i = 0; i < kCvTrials; ++i)
      if (__ballot(h && kp == i)) hits |= 1u << i;
  }
  return hits;
}

__device__ __forceinline__ void stage_verts(const Geometry& g, double* vx) {
  for (uint32_t i = threadIdx.x; i < 3 * g.nverts; i += blockDim.x) vx[i] = g.verts[i];
}

// One colour of one cycle for convex polyhedra: kCvTrials spaces per
// wavefront.  Each trial is the one of trial_convex (parallel_sweep.rs:234-277
// with Translate or Rotate<Versor>).
// (three wavefronts per SIMD: 22% faster on C3 than the two its free
// register use allows)
__global__ void __launch_bounds__(kBlock, 3) k_pass_convex(Geometry g, ApplyState* st, Bodies B, Cells C, Work w,
                                                        uint32_t colour, uint32_t cycle, uint32_t xcd_chunk) {
  if (cycle >= st->cycles) return;  // uniform over the grid
  __shared__ unsigned blk_acc;
  __shared__ double zx[257], zf[257];
  __shared__ double vx[3 * kMaxVertices];
  __shared__ CvWave cw[kBlock / 64];
  if (threadIdx.x == 0) blk_acc = 0;
  stage_zig(zx, zf);
  stage_verts(g, vx);
  __syncthreads();
  const uint32_t b = blockIdx.x;
  const uint32_t lb = (b & 7u) * xcd_chunk + (b >> 3);
  const uint32_t lane = __lane_id(), k = lane / kCvLanes, kl = lane % kCvLanes;
  CvWave& W = cw[threadIdx.x >> 6];
  const uint32_t tix = (lb * (kBlock / 64) + (threadIdx.x >> 6)) * kCvTrials + k;
  const uint32_t substep = st->substep0 + cycle;
  // the trial of lane group k (all its lanes draw the same proposal)
  bool tried = false, active = false;
  uint32_t s = 0, slot = 0xFFFFFFFFu, body = 0;
  int32_t m[3] = {0, 0, 0};
  double x[3] = {0.0, 0.0, 0.0};
  quat q = {1.0, {0.0, 0.0, 0.0}};
  Sfc64 rng = {0, 0, 0, 0};
  if (tix < g.per_colour) {
    s = colour_space<3>(g, colour, tix, m);
    const uint32_t s0 = w.start[s];
    const uint32_t n = w.start[s + 1] - s0;
    if (n) {
      tried = true;
      rng = counter_rng(st->step, substep, st->seed, s);
      slot = s0 + choose_index(rng, n);
      body = C.body[slot];
    }
  }
  // the trial's neighbour spaces (distinct ones, as the reference's cell
  // list visits each once) and the first candidate index of each; listed
  // before the proposal is drawn so that their loads overlap the body's
  uint32_t nb[3][5];
  uint32_t nn[3] = {0, 0, 0};
  if (tried) {
    for (int a = 0; a < 3; ++a) {
      const int32_t sh = g.shape[a], r = g.reach[a];
      uint32_t c = 0;
      for (int off = -r; off <= r; ++off) {
        const int32_t mm = ((m[a] + off) % sh + sh) % sh;
        bool seen = false;
        for (uint32_t j = 0; j < c; ++j) seen |= nb[a][j] == (uint32_t)mm * g.stride[a];
        if (!seen) nb[a][c++] = (uint32_t)mm * g.stride[a];
      }
      nn[a] = c;
    }
  }
  const uint32_t NS = nn[0] * nn[1] * nn[2];  // at most 125
  uint32_t tot = 0;
  for (uint32_t j0 = 0; __ballot(j0 < NS) != 0ull; j0 += kCvLanes) {
    const uint32_t j = j0 + kl;
    uint32_t cnt = 0, lo = 0;
    if (j < NS) {
      const uint32_t i2 = j % nn[2], i1 = (j / nn[2]) % nn[1], i0 = j / (nn[2] * nn[1]);
      const uint32_t ns = nb[0][i0] + nb[1][i1] + nb[2][i2];
      lo = w.start[ns];
      cnt = w.start[ns + 1] - lo;
    }
    uint32_t incl = cnt;
#pragma unroll
    for (uint32_t d = 1; d < kCvLanes; d <<= 1) {
      const uint32_t u = __shfl_up(incl, d, kCvLanes);
      if (kl >= d) incl += u;
    }
    if (j < NS) {
      W.pre[k][j] = tot + incl - cnt;
      W.lo[k][j] = lo;
    }
    tot += __shfl(incl, (int)(lane | (kCvLanes - 1)), 64);
  }
  // the proposal (the random draws overlap the loads of the body)
  double4 bp = make_double4(0.0, 0.0, 0.0, 0.0), bq = make_double4(1.0, 0.0, 0.0, 0.0);
  if (tried) {
    bp = B.bpos[body];
    bq = B.bquat[body];
  }
  double v[3] = {0.0, 0.0, 0.0};
  quat dq = {1.0, {0.0, 0.0, 0.0}};
  if (tried) {
    if (st->move == 0)
      ball_sample<3>(rng, st->max_move, zx, zf, v);
    else
      dq = versor_displacement(rng, st->max_move, zx, zf);
  }
  if (tried) {
    x[0] = bp.x;
    x[1] = bp.y;
    x[2] = bp.z;
    q = {bq.x, {bq.y, bq.z, bq.w}};
    if (st->move == 0) {
      for (int a = 0; a < 3; ++a) x[a] += v[a];
    } else {
      q = qmul(dq, q);
    }
    for (int a = 0; a < 3; ++a) x[a] = wrap1(x[a], g.L[a]);  // wrap(propose(..))
    active = space_of<3>(g, st->origin, x) == s;
  }
  if (!active) tot = 0;  // a trial outside its space is rejected without pair tests
  // the trial site: wrap(x + rotate(q, 0)) with orientation q * identity
  // (property/oriented_point.rs:107-113)
  const v3 r0 = qrotate(q, v3{0.0, 0.0, 0.0});
  const quat qt = qmul(q, quat{1.0, {0.0, 0.0, 0.0}});
  if (active && kl == 0) {
    W.t[k][0] = wrap1(x[0] + r0.x, g.L[0]);
    W.t[k][1] = wrap1(x[1] + r0.y, g.L[1]);
    W.t[k][2] = wrap1(x[2] + r0.z, g.L[2]);
    W.qt[k][0] = qt.s;
    W.qt[k][1] = qt.v.x;
    W.qt[k][2] = qt.v.y;
    W.qt[k][3] = qt.v.z;
  }
  uint32_t base[kCvTrials + 1];
  base[0] = 0;
#pragma unroll
  for (uint32_t i = 0; i < kCvTrials; ++i) base[i + 1] = base[i] + __shfl(tot, (int)(i * kCvLanes), 64);
  const uint32_t total = base[kCvTrials];
  wave_lds_order();
  uint32_t hits = 0;  // trials with an overlap (uniform over the wavefront)
  uint32_t np = 0;    // listed pairs (uniform)
  for (uint32_t c0 = 0; c0 < total; c0 += 64 * kCvBatch) {
    uint32_t qsv[kCvBatch], kcv[kCvBatch];
    bool ok[kCvBatch];
#pragma unroll
    for (uint32_t r = 0; r < kCvBatch; ++r) {
      const uint32_t c = c0 + 64 * r + lane;
      uint32_t kc = 0, bk = 0;
#pragma unroll
      for (uint32_t i = 1; i < kCvTrials; ++i)
        if (c >= base[i]) {
          kc = i;
          bk = base[i];
        }
      const uint32_t nsk = __shfl(NS, (int)(kc * kCvLanes), 64);
      const uint32_t slk = __shfl(slot, (int)(kc * kCvLanes), 64);
      ok[r] = c < total && !((hits >> kc) & 1u);
      uint32_t qs = 0;
      if (ok[r]) {
        // neighbour space of candidate c: last one whose first index is <= c
        const uint32_t cl = c - bk;
        uint32_t jlo = 0;
#pragma unroll
        for (uint32_t sw = kCvSegs / 2; sw >= 1; sw >>= 1) {
          const uint32_t mid = jlo + sw;
          if (mid < nsk && W.pre[kc][mid] <= cl) jlo = mid;
        }
        qs = W.lo[kc][jlo] + (cl - W.pre[kc][jlo]);
        ok[r] = qs != slk;
      }
      qsv[r] = qs;
      kcv[r] = kc;
    }
    double4 sj[kCvBatch];
#pragma unroll
    for (uint32_t r = 0; r < kCvBatch; ++r) sj[r] = ok[r] ? C.pos64[qsv[r]] : make_double4(0.0, 0.0, 0.0, 0.0);
#pragma unroll
    for (uint32_t r = 0; r < kCvBatch; ++r) {
      bool inr = false;
      if (ok[r]) {
        const double tq[3] = {W.t[kcv[r]][0], W.t[kcv[r]][1], W.t[kcv[r]][2]};
        v3 rj;
        inr = convex_image(g, tq, sj[r], &rj) < g.cv_rcut2;
      }
      const unsigned long long mk = __ballot(inr);
      if (inr) {
        const uint32_t p = np + (uint32_t)__popcll(mk & ((1ull << lane) - 1ull));
        W.qs[p] = qsv[r];
        W.tk[p] = kcv[r];
      }
      np += (uint32_t)__popcll(mk);
    }
    if (np + 64 * kCvBatch > kCvPairs || c0 + 64 * kCvBatch >= total) {  // uniform
      wave_lds_order();
      if (g.nverts == 8)
        hits = cv_pair_tests<8>(g, C, W, vx, np, hits);
      else if (g.nverts == 6)
        hits = cv_pair_tests<6>(g, C, W, vx, np, hits);
      else if (g.nverts == 4)
        hits = cv_pair_tests<4>(g, C, W, vx, np, hits);
      else
        hits = cv_pair_tests<0>(g, C, W, vx, np, hits);
      np = 0;
      wave_lds_order();
    }
  }
  const bool accept = active && !((hits >> k) & 1u);  // hard shapes: dH is 0 or +inf
  if (kl == 0 && tried) {
    if (accept) {
      // update_body_properties (microstate.rs:985-1034): body = wrap(props),
      // site = wrap(body.transform(site))
      double nb3[3], ns3[3];
      for (int a = 0; a < 3; ++a) nb3[a] = wrap1(x[a], g.L[a]);
      const double r0a[3] = {r0.x, r0.y, r0.z};
      for (int a = 0; a < 3; ++a) ns3[a] = wrap1(nb3[a] + r0a[a], g.L[a]);
      const uint32_t fl = ghost_flags<3>(g, ns3);
      B.bpos[body] = make_double4(nb3[0], nb3[1], nb3[2], 0.0);
      B.bquat[body] = make_double4(q.s, q.v.x, q.v.y, q.v.z);
      const double4 sp = pack_site(ns3, fl);
      const double4 sq = make_double4(qt.s, qt.v.x, qt.v.y, qt.v.z);
      B.spos[body] = sp;
      B.squat[body] = sq;
      C.pos64[slot] = sp;
      C.quat64[slot] = sq;
      atomicAdd(&blk_acc, 1u);
    }
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
  __syncthreads();
  if (threadIdx.x == 0 && blk_acc) atomicAdd(&w.acc_slots[b & (kAccSlots - 1)], (unsigned long long)blk_acc);
}

// Convex cycles beyond the ones the host launched, in one workgroup.
__global__ void __launch_bounds__(kGuardThreads) k_guard_convex(Geometry g, ApplyState* st, Bodies B, Cells C,
                                                                Work w, uint32_t first_cycle) {
  const uint32_t cycles = st->cycles;
  if (first_cycle >= cycles) return;
  __shared__ double zx[257], zf[257];
  __shared__ double vx[3 * kMaxVertices];
  stage_zig(zx, zf);
  stage_verts(g, vx);
  __syncthreads();
  const uint32_t rounds = (g.per_colour + kGuardThreads - 1) / kGuardThreads;
  for (uint32_t k = first_cycle; k < cycles; ++k) {
    for (uint32_t c = 0; c < 8; ++c) {
      for (uint32_t r = 0; r < rounds; ++r) {
        const uint32_t t = r * kGuardThreads + threadIdx.x;
        TrialOut o{0, 0};
        if (t < g.per_colour) {
          int32_t m[3] = {0, 0, 0};
          const uint32_t s = colour_space<3>(g, c, t, m);
          const uint32_t s0 = w.start[s];
          const uint32_t n = w.start[s + 1] - s0;
          if (n) trial_convex(g, st, B, C, w, zx, zf, vx, s, m, s0, n, c, st->substep0 + k, o);
        }
        wave_count(&w.acc_slots[0], o.accepted != 0);
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) st->guard_cycles += 1;
  }
}

// Site orientations: given, or body * identity (oriented_point.rs:107-113).
__global__ void k_init_site_quats(Geometry g, Bodies B, int have_sites) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.n_bodies || have_sites) return;
  const double4 b = B.bquat[e];
  const quat qs = qmul(quat{b.x, {b.y, b.z, b.w}}, quat{1.0, {0.0, 0.0, 0.0}});
  B.squat[e] = make_double4(qs.s, qs.v.x, qs.v.y, qs.v.z);
}

// Test hook: collide3d (kind 0, in = v_ij, q_ij: 7 values), the global pair
// test (kind 1, in = r_i, q_i, r_j, q_j: 14 values) or the rotation sampler
// (kind 2, in = step, substep, seed, index, sigma: out 4 values).
__global__ void k_debug_convex(int kind, const double* vx, uint32_t nv, double rcut2, const double* in,
                               double* out, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (kind == 0) {
    const double* p = in + 7 * i;
    out[i] = collide3d(vx, nv, p[0], p[1], p[2], p[3], p[4], p[5], p[6]) ? 1.0 : 0.0;
  } else if (kind == 1) {
    const double* p = in + 14 * i;
    const v3 ri = {p[0], p[1], p[2]}, rj = {p[7], p[8], p[9]};
    double d2 = 0.0;
    const double dd[3] = {p[0] - p[7], p[1] - p[8], p[2] - p[9]};
    for (int a = 0; a < 3; ++a) d2 = d2 + dd[a] * dd[a];
    out[i] = convex_overlap(vx, nv, rcut2, d2, ri, quat{p[3], {p[4], p[5], p[6]}}, rj,
                            quat{p[10], {p[11], p[12], p[13]}})
                 ? 1.0
                 : 0.0;
  } else {
    const double* p = in + 5 * i;
    Sfc64 rng = counter_rng((uint64_t)p[0], (uint32_t)p[1], (uint32_t)p[2], (uint64_t)p[3]);
    const quat q = versor_displacement(rng, p[4], kZigX, kZigF);
    out[4 * i] = q.s;
    out[4 * i + 1] = q.v.x;
    out[4 * i + 2] = q.v.y;
    out[4 * i + 3] = q.v.z;
  }
}


// --------------------------------------------------------- replica batches --
// Many independent systems (each with its own box, checkerboard, counter and
// seed) advanced by the same launches: bodies, spaces and one colour's spaces
// of all replicas are concatenated, and every thread finds its replica from
// the running offsets.  Each replica's trials are exactly those of a context
// of its own (same spaces, same streams), so the replay bar carries over.
struct BatchRep {
  Geometry g;
  uint32_t body_off, space_off, colour_off, pad;
};

// the last r with off[r] <= x (off[0] = 0, R entries)
__device__ __forceinline__ uint32_t find_replica(const uint32_t* off, uint32_t R, uint32_t x) {
  uint32_t lo = 0, hi = R;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= x)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
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

// ------------------------------------------------------------- total energy --
// PairwiseCutoff::total_energy (pairwise_cutoff.rs:720-745): the sum over
// sites i and neighbours j with site_tag_i < site_tag_j (body index = site
// tag for single-site bodies) of the pair energy, infinite on any hard
// overlap.  Per block: the partial sum, the sum of |terms|, the number of
// non-zero terms and an overlap flag, reduced in a fixed order.
constexpr int kEnergyBlock = 256;

template <int D>
__global__ void __launch_bounds__(kEnergyBlock) k_energy(Geometry g, const ApplyState* st, Bodies B, Work w,
                                                         double* part) {
  __shared__ double vx[3 * kMaxVertices];
  __shared__ double re[kEnergyBlock], ra[kEnergyBlock], rn[kEnergyBlock];
  __shared__ int any_inf;
  if (threadIdx.x == 0) any_inf = 0;
  if (g.model == HMC_CONVEX_POLYHEDRON) stage_verts(g, vx);
  __syncthreads();
  const uint32_t i = blockIdx.x * kEnergyBlock + threadIdx.x;
  double e = 0.0, a = 0.0, cnt = 0.0;
  bool inf = false;
  if (i < g.n_bodies) {
    const double4 si = B.spos[i];
    const double t[3] = {si.x, si.y, si.z};
    const double4 bp = B.bpos[i];
    const double x[3] = {bp.x, bp.y, bp.z};
    uint32_t s = space_of<D>(g, st->origin, x);
    if (s == UINT32_MAX) s = 0;  // outside the checkerboard: reported as an error by the host
    int32_t m[3] = {0, 0, 0};
    for (int ax = 0; ax < D; ++ax) m[ax] = (int32_t)((s / g.stride[ax]) % (uint32_t)g.shape[ax]);
    uint32_t nb[3][5];
    int nn[3] = {1, 1, 1};
    for (int ax = 0; ax < D; ++ax) {
      const int32_t sh = g.shape[ax], r = g.reach[ax];
      int k = 0;
      for (int off = -r; off <= r; ++off) {
        const int32_t mm = ((m[ax] + off) % sh + sh) % sh;
        bool seen = false;
        for (int j = 0; j < k; ++j) seen |= nb[ax][j] == (uint32_t)mm * g.stride[ax];
        if (!seen) nb[ax][k++] = (uint32_t)mm * g.stride[ax];
      }
      nn[ax] = k;
    }
    quat qi = {1.0, {0.0, 0.0, 0.0}};
    if (g.model == HMC_CONVEX_POLYHEDRON) {
      const double4 q4 = B.squat[i];
      qi = {q4.x, {q4.y, q4.z, q4.w}};
    }
    for (int i0 = 0; i0 < nn[0] && !inf; ++i0)
      for (int i1 = 0; i1 < nn[1] && !inf; ++i1)
        for (int i2 = 0; i2 < (D == 3 ? nn[2] : 1) && !inf; ++i2) {
          const uint32_t ns = nb[0][i0] + nb[1][i1] + (D == 3 ? nb[2][i2] : 0u);
          const uint32_t q0 = w.start[ns], q1 = w.start[ns + 1];
          for (uint32_t q = q0; q < q1 && !inf; ++q) {
            const uint32_t j = w.tmp_body[q];
            if (j <= i) continue;
            const double4 sj = B.spos[j];
            if (g.model == HMC_HARD_SPHERE) {
              inf = overlap64<D>(g, t, sj);
            } else if (g.model == HMC_CONVEX_POLYHEDRON) {
              v3 rj;
              const double d2 = convex_image(g, t, sj, &rj);
              if (d2 < g.cv_rcut2) {
                const double4 q4 = B.squat[j];
                inf = convex_overlap(vx, g.nverts, g.cv_rcut2, d2, v3{t[0], t[1], t[2]}, qi, rj,
                                     quat{q4.x, {q4.y, q4.z, q4.w}});
              }
            } else {
              const double r2 = lj_r2<D>(g, t, sj);
              if (r2 < g.lj_rc2) {
                const double one = lj_energy(g, r2);
                if (one == INFINITY) {  // pairwise_cutoff.rs:720-745 returns at once
                  inf = true;
                  break;
                }
                e += one;
                a += fabs(one);
                cnt += 1.0;
              }
            }
          }
        }
  }
  if (inf) any_inf = 1;
  re[threadIdx.x] = e;
  ra[threadIdx.x] = a;
  rn[threadIdx.x] = cnt;
  __syncthreads();
  for (int h = kEnergyBlock / 2; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h) {
      re[threadIdx.x] += re[threadIdx.x + h];
      ra[threadIdx.x] += ra[threadIdx.x + h];
      rn[threadIdx.x] += rn[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[4 * blockIdx.x] = re[0];
    part[4 * blockIdx.x + 1] = ra[0];
    part[4 * blockIdx.x + 2] = rn[0];
    part[4 * blockIdx.x + 3] = any_inf ? 1.0 : 0.0;
  }
}

// Sum of the block partials by one workgroup of kEnergyFinal threads: thread t
// adds partials t, t + kEnergyFinal, ... in order, then a fixed tree in LDS
// (a deterministic order; the energy's bound holds for any order of its terms).
constexpr int kEnergyFinal = 256;
__global__ void __launch_bounds__(kEnergyFinal) k_energy_final(const double* part, uint32_t nblocks, double* out) {
  __shared__ double r[4][kEnergyFinal];
  double e = 0.0, a = 0.0, n = 0.0, inf = 0.0;
  for (uint32_t b = threadIdx.x; b < nblocks; b += kEnergyFinal) {
    e += part[4 * b];
    a += part[4 * b + 1];
    n += part[4 * b + 2];
    inf = fmax(inf, part[4 * b + 3]);
  }
  r[0][threadIdx.x] = e;
  r[1][threadIdx.x] = a;
  r[2][threadIdx.x] = n;
  r[3][threadIdx.x] = inf;
  __syncthreads();
  for (int h = kEnergyFinal / 2; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h) {
      r[0][threadIdx.x] += r[0][threadIdx.x + h];
      r[1][threadIdx.x] += r[1][threadIdx.x + h];
      r[2][threadIdx.x] += r[2][threadIdx.x + h];
      r[3][threadIdx.x] = fmax(r[3][threadIdx.x], r[3][threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = r[0][0];
    out[1] = r[1][0];
    out[2] = r[2][0];
    out[3] = r[3][0];
  }
}

// PairwiseCutoff::delta_energy_insert (pairwise_cutoff.rs:1017-1025 with
// :420-488): the energy of a new body's site wrap(body + 0) with every site of
// the microstate (the new body has no tag yet, so nothing is filtered),
// infinite on a hard overlap or an infinite pair term.  One thread per
// candidate over the 3^D (or 5^D) neighbour spaces of the current cells;
// out[2 k] is the energy, out[2 k + 1] the bound 2 gamma_n sum |terms| on its
// difference from the reference's summation order.
template <int D>
__global__ void __launch_bounds__(kEnergyBlock) k_insert_energy(Geometry g, const ApplyState* st, Bodies B, Work w,
                                                                const double* __restrict__ cand, uint32_t n,
                                                                double* __restrict__ out, uint32_t* outside) {
  const uint32_t k = blockIdx.x * kEnergyBlock + threadIdx.x;
  if (k >= n) return;
  double t[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < D; ++a) t[a] = wrap1(cand[(size_t)k * D + a] + 0.0, g.L[a]);
  const uint32_t s = space_of<D>(g, st->origin, t);
  if (s == UINT32_MAX) {  // outside the checkerboard: reported as an error by the host
    atomicOr(outside, 1u);
    out[2 * (size_t)k] = 0.0;
    out[2 * (size_t)k + 1] = 0.0;
    return;
  }
  int32_t m[3] = {0, 0, 0};
  for (int ax = 0; ax < D; ++ax) m[ax] = (int32_t)((s / g.stride[ax]) % (uint32_t)g.shape[ax]);
  uint32_t nb[3][5];
  int nn[3] = {1, 1, 1};
  for (int ax = 0; ax < D; ++ax) {
    const int32_t sh = g.shape[ax], r = g.reach[ax];
    int kk = 0;
    for (int off = -r; off <= r; ++off) {
      const int32_t mm = ((m[ax] + off) % sh + sh) % sh;
      bool seen = false;
      for (int j = 0; j < kk; ++j) seen |= nb[ax][j] == (uint32_t)mm * g.stride[ax];
      if (!seen) nb[ax][kk++] = (uint32_t)mm * g.stride[ax];
    }
    nn[ax] = kk;
  }
  double e = 0.0, a = 0.0, cnt = 0.0;
  bool inf = false;
  for (int i0 = 0; i0 < nn[0] && !inf; ++i0)
    for (int i1 = 0; i1 < nn[1] && !inf; ++i1)
      for (int i2 = 0; i2 < (D == 3 ? nn[2] : 1) && !inf; ++i2) {
        const uint32_t ns = nb[0][i0] + nb[1][i1] + (D == 3 ? nb[2][i2] : 0u);
        const uint32_t q0 = w.start[ns], q1 = w.start[ns + 1];
        for (uint32_t q = q0; q < q1 && !inf; ++q) {
          const double4 sj = B.spos[w.tmp_body[q]];
          if (g.model == HMC_HARD_SPHERE) {
            inf = overlap64<D>(g, t, sj);
          } else {
            const double r2 = lj_r2<D>(g, t, sj);
            if (r2 < g.lj_rc2) {
              const double one = lj_energy(g, r2);
              if (one == INFINITY) {
                inf = true;
                break;
              }
              e += one;
              a += fabs(one);
              cnt += 1.0;
            }
          }
        }
      }
  const double nt = fmax(cnt, 1.0);
  const double gam = nt * 0x1p-53 / (1.0 - nt * 0x1p-53);
  out[2 * (size_t)k] = inf ? INFINITY : e;
  out[2 * (size_t)k + 1] = inf ? 0.0 : 2.0 * gam * a;
}

// ------------------------------------------------ scale distribution function --
// The SDF of hard spheres/disks as HOOMD-blue's hpmc compute.SDF defines it
// (no counterpart in the reference): for each body, the smallest compression
// x = 1 - sigma / r over the other bodies (r the minimum-image distance
// between sites) at which the scaled system would first overlap, counted in
// bin floor(x / dx) when 0 <= x < xmax and the bin is below nbins; x < 0 (an
// overlap) is counted apart.  x is monotone in r^2, so the minimum is taken
// over r^2 = (d_x^2 + d_y^2) + d_z^2 and x is formed once.  Bodies within
// r_max = sigma / (1 - xmax) of a body lie in its neighbour spaces: one space
// on each side when the space width is at least r_max (1 + 2^-30), else two
// (the width is at least sigma and xmax < 1/2).  reps/boff describe one or
// several systems laid out as in a batch.
template <int D>
__global__ void __launch_bounds__(kBlock) k_sdf(const BatchRep* reps, const uint32_t* boff, uint32_t R,
                                                uint32_t n_total, Bodies B, Work w, double xmax, double dx,
                                                uint32_t nbins, uint32_t* hist, uint32_t* ovl) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_total) return;
  const uint32_t r = R > 1 ? find_replica(boff, R, b) : 0u;
  const Geometry& g = reps[r].g;
  const uint32_t soff = reps[r].space_off;
  const uint32_t s = w.elem_space[b] - soff;
  const double4 si = B.spos[b];
  const double t[3] = {si.x, si.y, si.z};
  const double rmax = g.sigma / (1.0 - xmax);
  uint32_t nb[3][5];
  int nn[3] = {1, 1, 1};
  for (int a = 0; a < D; ++a) {
    const int32_t sh = g.shape[a];
    const int32_t m = (int32_t)((s / g.stride[a]) % (uint32_t)sh);
    const int32_t reach = g.width[a] >= rmax * (1.0 + 0x1p-30) ? 1 : 2;
    int k = 0;
    for (int off = -reach; off <= reach; ++off) {
      const int32_t mm = ((m + off) % sh + sh) % sh;
      bool seen = false;
      for (int j = 0; j < k; ++j) seen |= nb[a][j] == (uint32_t)mm * g.stride[a];
      if (!seen) nb[a][k++] = (uint32_t)mm * g.stride[a];
    }
    nn[a] = k;
  }
  auto r2_to = [&](const double4& sj) {
    const double u[3] = {sj.x, sj.y, sj.z};
    double r2 = 0.0;
    for (int a = 0; a < D; ++a) {
      double d = t[a] - u[a];
      if (d > g.half[a])
        d -= g.L[a];
      else if (d < -g.half[a])
        d += g.L[a];
      r2 = r2 + d * d;
    }
    return r2;
  };
  double best = __builtin_inf();
  constexpr int NC = D == 3 ? 27 : 9;
  if (nn[0] == 3 && nn[1] == 3 && (D == 2 || nn[2] == 3)) {
    // 3^D distinct neighbour spaces: their ranges first, then rounds (the
    // k-th slot of every space), each with its loads issued before any test
    uint32_t lo[NC], hi[NC];
    uint32_t rounds = 0;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const uint32_t ns = soff + nb[0][c / (D == 3 ? 9 : 3)] + nb[1][(c / (D == 3 ? 3 : 1)) % 3] +
                          (D == 3 ? nb[2][c % 3] : 0u);
      lo[c] = w.start[ns];
      hi[c] = w.start[ns + 1];
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) rounds = max(rounds, hi[c] - lo[c]);
    constexpr int CH = 9;
    for (uint32_t k = 0; k < rounds; ++k) {
#pragma unroll
      for (int c0 = 0; c0 < NC; c0 += CH) {
        uint32_t jv[CH];
#pragma unroll
        for (int j = 0; j < CH; ++j) jv[j] = lo[c0 + j] + k < hi[c0 + j] ? w.tmp_body[lo[c0 + j] + k] : b;
        double4 sv[CH];
#pragma unroll
        for (int j = 0; j < CH; ++j) sv[j] = B.spos[jv[j]];
#pragma unroll
        for (int j = 0; j < CH; ++j)
          if (jv[j] != b) best = fmin(best, r2_to(sv[j]));
      }
    }
  } else {
    for (int i0 = 0; i0 < nn[0]; ++i0)
      for (int i1 = 0; i1 < nn[1]; ++i1)
        for (int i2 = 0; i2 < (D == 3 ? nn[2] : 1); ++i2) {
          const uint32_t ns = soff + nb[0][i0] + nb[1][i1] + (D == 3 ? nb[2][i2] : 0u);
          const uint32_t q1 = w.start[ns + 1];
          for (uint32_t q = w.start[ns]; q < q1; ++q) {
            const uint32_t j = w.tmp_body[q];
            if (j != b) best = fmin(best, r2_to(B.spos[j]));
          }
        }
  }
  if (!(best < __builtin_inf())) return;  // no other body nearby
  const double x = 1.0 - g.sigma / sqrt(best);
  if (x < 0.0) {
    atomicAdd(&ovl[r], 1u);
  } else if (x < xmax) {
    const double k = floor(x / dx);
    if (k < (double)nbins) atomicAdd(&hist[(size_t)r * nbins + (uint32_t)k], 1u);
  }
}

// Site positions (given, or wrap(body + 0)) and their ghost flags.
template <int D>
__global__ void k_init_sites(Geometry g, Bodies B, int have_sites) {
  uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.n_bodies) return;
  double4 b = B.bpos[e];
  double x[3] = {b.x, b.y, b.z};
  double s[3] = {0.0, 0.0, 0.0};
  if (have_sites) {
    double4 sp = B.spos[e];
    s[0] = sp.x;
    s[1] = sp.y;
    s[2] = sp.z;
  } else {
    for (int a = 0; a < D; ++a) s[a] = wrap1(x[a] + 0.0, g.L[a]);
  }
  B.spos[e] = pack_site(s, ghost_flags<D>(g, s));
}

// Microstate::clone_with_boundary with every body mapped (microstate.rs:1529-1565):
// Hypercuboid::map_point (cuboid.rs:553-567), p' = clamp(p (L'/L), -L'/2,
// next_down(L'/2)) per axis, then update_body_properties (microstate.rs:985-1034):
// body = wrap(p'), site = wrap(body + 0) with its ghost flags.  g is the new
// geometry; a body outside the old box (map_point's error, on which the
// reference panics) sets *outside.
template <int D>
__global__ void k_map_bodies(Geometry g, double3 lold, const double4* __restrict__ src, Bodies B,
                             uint32_t* outside) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.n_bodies) return;
  const double4 b = src[e];
  const double x[3] = {b.x, b.y, b.z};
  const double lo[3] = {lold.x, lold.y, lold.z};
  double p[3] = {0.0, 0.0, 0.0};
  bool inside = true;
  for (int a = 0; a < D; ++a) {
    inside = inside && -lo[a] / 2.0 <= x[a] && x[a] < lo[a] / 2.0;  // cuboid.rs:449-454
    const double scale = g.L[a] / lo[a];
    const double low = -g.L[a] / 2.0;
    const double high = __longlong_as_double(__double_as_longlong(g.L[a] / 2.0) - 1);  // next_down of a positive
    const double v = scale * x[a];
    p[a] = v < low ? low : (v > high ? high : v);  // f64::clamp
  }
  if (!inside) atomicOr(outside, 1u);
  double s[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < D; ++a) p[a] = wrap1(p[a], g.L[a]);
  for (int a = 0; a < D; ++a) s[a] = wrap1(p[a] + 0.0, g.L[a]);
  B.bpos[e] = make_double4(p[0], p[1], p[2], 0.0);
  B.spos[e] = pack_site(s, ghost_flags<D>(g, s));
}

// ------------------------------------------------------------- debug hooks --
__global__ void k_debug_math(int kind, const double* in, double* out, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r = 0.0;
  switch (kind) {
    case 0: r = hmc_cbrt(in[i]); break;
    case 1: r = glibc_log(in[i]); break;
    case 2: r = glibc_exp(in[i]); break;
    case 3: r = sqrt(in[i]); break;
    case 5: {
      double c;
      glibc_sincos(in[i], &r, &c);
      break;
    }
    case 6: {
      double sn;
      glibc_sincos(in[i], &sn, &r);
      break;
    }
    default: r = in[2 * i] / in[2 * i + 1]; break;
  }
  out[i] = r;
}

__global__ void k_debug_normals(uint64_t step, uint32_t substep, uint32_t seed, uint64_t index, double* out,
                                uint64_t n) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  Sfc64 r = counter_rng(step, substep, seed, index);
  for (uint64_t i = 0; i < n; ++i) out[i] = standard_normal(r, kZigX, kZigF);
}

}  // namespace hmc

// ======================================================================= host
using namespace hmc;

struct hmc_ctx {
  hmc_config cfg;
  hmc_config cfg_user;              // as given to hmc_create (vertices -> verts_host)
  std::vector<double> verts_host;   // convex: the vertex table, for hmc_clone_with_boundary
  Geometry g;
  int device;
  hipStream_t stream = nullptr;
  Bodies bodies{};
  Cells cells{};
  Work w{};
  ApplyState* st = nullptr;
  bool have_state = false;
  uint32_t cycles_guess = 1;
  bool fast = false;  // every axis: >= 3 spaces and reach 1
  bool tiled = false;  // LDS-staged pass (requires fast)
  TileCfg tc{};
  bool persist = false;  // one cooperative launch per apply (k_apply_tile)
  TileCfg pc{};
  uint32_t pgrid = 0, pthreads = 0;
  double* verts = nullptr;  // convex polyhedron vertices (device)
  Geometry* gdev = nullptr;  // device copy of g (g.self)
  uint32_t n_scan_blocks = 0;
  uint64_t applies = 0;
  std::string err;
  bool profiling = false;
  bool capturing = false;       // recording a graph: no event timing
  uint32_t profile_period = 1;  // time the launches of every k-th apply
  // one apply (with increment_step) captured as a graph, and what it depends on
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  double graph_kT = 0.0, graph_move_size = 0.0;
  uint32_t graph_move = 0, graph_launched = 0;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pass_events, build_events;
  double pass_ms = 0.0, build_ms = 0.0;
  uint64_t pass_n = 0, build_n = 0;
};

namespace {

int fail(hmc_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, expr)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) return fail((ctx), HMC_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

template <class T>
hipError_t dalloc(T** p, size_t n) {
  return hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
}

uint32_t grid_for(uint64_t n, uint32_t block) { return (uint32_t)((n + block - 1) / block); }

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
  if constexpr (PREC == HMC_MIXED) {
    if (c->persist) {
      hipEvent_t p0 = nullptr, p1 = nullptr;
      if (prof) {
        p0 = take_event(c);
        (void)hipEventRecord(p0, s);
      }
      ApplyState* stp = c->st;
      void* args[] = {(void*)&g, (void*)&stp, (void*)&c->bodies, (void*)&c->cells, (void*)&c->w, (void*)&c->pc};
      const hipError_t le = hipLaunchCooperativeKernel((const void*)k_apply_tile<D>, dim3(c->pgrid),
                                                       dim3(c->pthreads), args, c->pc.lds_bytes, s);
      if (le != hipSuccess) return fail(c, HMC_ERR_DEVICE, std::string("cooperative launch: ") + hipGetErrorString(le));
      if (prof) {
        p1 = take_event(c);
        (void)hipEventRecord(p1, s);
        c->pass_events.push_back({p0, p1});
      }
      c->applies += 1;
      return HMC_OK;
    }
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
      else if (convex)
        k_pass_convex<<<cv_chunk * 8, kBlock, 0, s>>>(g, c->st, c->bodies, c->cells, c->w, (uint32_t)col, k,
                                                      cv_chunk);
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

int read_state(hmc_ctx* c, ApplyState* out) {
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
  if (out->error & kErrBarrier) return fail(c, HMC_ERR_DEVICE, "a grid barrier of the persistent apply timed out");
  if (out->error) return fail(c, HMC_ERR_COVER, "a body lies outside the checkerboard");
  return HMC_OK;
}

}  // namespace

struct hmc_batch {
  int device = 0;
  uint32_t R = 0;
  int dim = 3;
  int precision = HMC_FP64;
  bool fast = true;
  double kT = 1.0;
  std::vector<BatchRep> hreps;
  std::vector<uint32_t> boff, coff;  // R + 1 running offsets of bodies and colour spaces
  uint32_t n_total = 0, s_total = 0, pc_total = 0;
  Geometry gtot{};  // totals for the replica-agnostic kernels
  BatchRep* reps = nullptr;
  ApplyState* sts = nullptr;  // R + 1 (the last one is scratch)
  uint32_t *dboff = nullptr, *dcoff = nullptr;
  uint64_t* dstep = nullptr;
  uint32_t* dseed = nullptr;
  double* dmove = nullptr;
  Bodies bodies{};
  Cells cells{};
  Work w{};
  hipStream_t stream = nullptr;
  uint32_t n_scan_blocks = 0;
  uint32_t cycles_guess = 1;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint32_t graph_cycles = 0;
  // LDS-tiled passes (k_bpass_tile): per-replica tile shapes, first tiles
  bool tiled = false;
  TileCfg* dtcs = nullptr;
  uint32_t* dtoff = nullptr;
  uint32_t* dtrep = nullptr;  // replica of each tile
  uint32_t dtrep_cap = 0;
  uint32_t ntiles = 0, tile_threads = 0, tile_lds = 0;
  std::string err;
};

// defined with the configuration checks below
extern "C" {
static bool contact_band(double E, double sigma, double sigma2, double edge2, int D, float* lo, float* hi);
static bool choose_tiles(const Geometry& g, int D, double occ, size_t vt, int cus, TileCfg* out);
}

namespace {

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
  for (uint32_t r = 0; r < b->R; ++r)
    if (st[r].error) return bfail(b, HMC_ERR_COVER, "a body lies outside the checkerboard (replica " + std::to_string(r) + ")");
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

// FP32 thresholds around contact for a per-axis error bound E on the FP32
// displacement: d2 < lo overlaps (and lies below the edge region), d2 >= hi
// does not.  The bound on d^2 near sigma^2 is 2 E sqrt(D) d + D E^2 +
// 4 (2^-24 + 2^-53) d^2 at d = 1.01 sigma (E must be far below sigma), with
// 10% margin; the float thresholds are rounded outward.
static bool contact_band(double E, double sigma, double sigma2, double edge2, int D, float* lo, float* hi) {
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
static bool choose_tiles(const Geometry& g, int D, double occ, size_t vt, int cus, TileCfg* out) {
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
static bool choose_regions(const Geometry& g, int D, double occ, int cus, TileCfg* out) {
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
        const size_t lds = ((16 * ((size_t)rows + 1) + 4 * (size_t)rows * (EL + 1) + 15) & ~(size_t)15) +
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

// Geometry of one system: validation, checkerboard dimensions, model
// constants and the FP32 contact band.  Shared by a context and by every
// replica of a batch.
static int make_geometry(const hmc_config* cfg, Geometry& g, bool* fast, double* sigma_out, double* lmax_out) {
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

int hmc_create(const hmc_config* cfg_in, hmc_ctx** out) {
  if (!cfg_in || !out) return HMC_ERR_INVALID;
  *out = nullptr;
  // LJ in mixed precision keeps the FP64 data layout and passes; only its
  // pair energies run in FP32 (Geometry::lj_mixed)
  hmc_config cfg_local = *cfg_in;
  const bool lj_mixed = (cfg_in->model == HMC_LJ || cfg_in->model == HMC_LJ_SHIFTED) &&
                        cfg_in->precision == HMC_MIXED;
  if (lj_mixed) cfg_local.precision = HMC_FP64;
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
  c->g.lj_fp32_error = (cfg_in->flags & HMC_FLAG_FP32_ERROR) && cfg_in->precision == HMC_MIXED ? 1 : 0;
  c->fast = fast0;
  Geometry& g = c->g;
  if (hipSetDevice(c->device) != hipSuccess) {
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
          const size_t fbytes = (size_t)pc.ntiles * kMailboxStride * sizeof(uint32_t);
          if (hipMalloc((void**)&c->w.mbox, fbytes) != hipSuccess || hipMemset(c->w.mbox, 0, fbytes) != hipSuccess) {
            delete c;
            return HMC_ERR_DEVICE;
          }
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
      kind > 2)
    return HMC_ERR_INVALID;
  if (hipSetDevice(device) != hipSuccess) return HMC_ERR_DEVICE;
  const size_t w_in = kind == 0 ? 7 : kind == 1 ? 14 : 5, w_out = kind == 2 ? 4 : 1;
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
      k_debug_convex<<<grid_for(n, 64), 64>>>(kind, dv, n_vertices, (R + R) * (R + R), di, dout, n);
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

int hmc_run_translate(hmc_ctx* c, uint64_t step0, uint32_t seed, double kT, double max_distance,
                      uint32_t n_steps, hmc_count* out) {
  if (!c) return HMC_ERR_INVALID;
  if (!c->have_state) return fail(c, HMC_ERR_STATE, "hmc_set_state was not called");
  if (!(max_distance > 0.0) || !(kT > 0.0)) return fail(c, HMC_ERR_INVALID, "max_distance and kT must be > 0");
  HIPCHK(c, hipSetDevice(c->device));
  ApplyState before;
  int rc = read_state(c, &before);
  if (rc) return rc;
  const bool lj = is_isotropic(c->g.model);
  // (a cooperative launch is not captured: the persistent apply launches directly)
  const bool graphs = !(c->cfg.flags & HMC_FLAG_NO_GRAPH) && !c->persist;
  for (uint32_t i = 0; i < n_steps; ++i) {
    const bool timed = c->profiling && c->applies % c->profile_period == 0;
    if (i == 0 || !graphs || timed) {
      rc = dispatch_apply(c, i == 0 ? 1 : 0, i == 0 ? 0 : 1, step0, 0, seed, kT, max_distance, 0u);
      if (rc) return rc;
      continue;
    }
    // later steps: replay one captured { apply; increment_step }
    const uint32_t launched = c->cycles_guess + (lj ? 1u : 0u);
    if (!c->graph_exec || c->graph_kT != kT || c->graph_move_size != max_distance || c->graph_move != 0u ||
        c->graph_launched != launched) {
      drop_graph(c);
      HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      c->capturing = true;
      rc = dispatch_apply(c, 0, 1, step0, 0, seed, kT, max_distance, 0u);
      c->capturing = false;
      c->applies -= 1;  // counted when replayed
      hipGraph_t gr = nullptr;
      const hipError_t ce = hipStreamEndCapture(c->stream, &gr);
      if (rc) return rc;
      HIPCHK(c, ce);
      c->graph = gr;
      HIPCHK(c, hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
      c->graph_kT = kT;
      c->graph_move_size = max_distance;
      c->graph_move = 0u;
      c->graph_launched = launched;
    }
    HIPCHK(c, hipGraphLaunch(c->graph_exec, c->stream));
    c->applies += 1;
  }
  ApplyState after;
  rc = read_state(c, &after);
  if (rc) return rc;
  if (n_steps) c->cycles_guess = std::max<uint32_t>(after.cycles, 1);
  uint64_t acc = after.accepted_total - before.accepted_total;
  uint64_t tri = after.trials_total - before.trials_total;
  if (out) {
    out->accepted = acc;
    out->rejected = tri - acc;
  }
  return HMC_OK;
}

// Arguments of the SDF for one system: hard spheres/disks, 0 < xmax < 1/2,
// dx > 0, a box wider than 2 r_max (one image within r_max) and spaces at
// least r_max / 2 wide (two neighbour spaces on each side suffice).
static const char* sdf_check(const Geometry& g, double xmax, double dx, uint32_t nbins) {
  if (g.model != HMC_HARD_SPHERE) return "the SDF is built for hard spheres and disks";
  if (!(xmax > 0.0 && xmax < 0.5) || !(dx > 0.0) || !std::isfinite(dx) || nbins == 0 || nbins > (1u << 24))
    return "SDF needs 0 < xmax < 1/2, dx > 0 and 0 < nbins <= 2^24";
  const double rmax = g.sigma / (1.0 - xmax);
  for (int a = 0; a < g.dim; ++a) {
    if (!(g.L[a] > 2.0 * rmax * (1.0 + 0x1p-30))) return "box edge below 2 sigma / (1 - xmax)";
    if (!(2.0 * g.width[a] >= rmax * (1.0 + 0x1p-30))) return "checkerboard spaces narrower than r_max / 2";
  }
  return nullptr;
}

// The SDF's own cells: the checkerboard with spaces at least r_max (1 + 2^-29)
// wide (never more spaces than the sweep's, whose buffers it reuses), so
// that one neighbour space on each side covers r_max.
static Geometry sdf_geometry(const Geometry& g, double xmax) {
  Geometry s = g;
  const double rmax = g.sigma / (1.0 - xmax);
  uint64_t spaces = 1;
  for (int a = 0; a < g.dim; ++a) {
    int64_t k = (int64_t)std::floor(g.L[a] / (rmax * (1.0 + 0x1p-29)));
    k = std::max<int64_t>(1, std::min<int64_t>(k, g.shape[a]));
    s.shape[a] = (int32_t)k;
    s.half_shape[a] = (uint32_t)std::max<int64_t>(1, k / 2);
    s.width[a] = g.L[a] / (double)k;
    spaces *= (uint64_t)k;
  }
  uint32_t str = 1;
  for (int a = g.dim - 1; a >= 0; --a) {
    s.stride[a] = str;
    str *= (uint32_t)s.shape[a];
  }
  s.n_spaces = (uint32_t)spaces;
  s.per_colour = s.n_spaces >> g.dim;
  return s;
}

int hmc_sdf(hmc_ctx* c, double xmax, double dx, uint32_t nbins, uint64_t* hist, uint64_t* overlaps) {
  if (!c || !hist) return HMC_ERR_INVALID;
  if (!c->have_state) return fail(c, HMC_ERR_STATE, "hmc_set_state was not called");
  if (const char* why = sdf_check(c->g, xmax, dx, nbins))
    return fail(c, c->g.model != HMC_HARD_SPHERE ? HMC_ERR_UNSUPPORTED : HMC_ERR_INVALID, why);
  const Geometry g = sdf_geometry(c->g, xmax);
  const uint32_t nsb = grid_for(g.n_spaces, kScanBlock);
  const uint32_t N = g.n_bodies;
  HIPCHK(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  // device: the system as a one-entry batch table, body offsets, counts
  const size_t bytes = sizeof(BatchRep) + 2 * sizeof(uint32_t) + ((size_t)nbins + 1) * sizeof(uint32_t);
  char* buf = nullptr;
  HIPCHK(c, hipMallocAsync((void**)&buf, bytes, s));
  BatchRep* drep = (BatchRep*)buf;
  uint32_t* dboff = (uint32_t*)(buf + sizeof(BatchRep));
  uint32_t* dh = dboff + 2;
  const BatchRep hrep{g, 0u, 0u, 0u, 0u};
  const uint32_t hboff[2] = {0u, N};
  hipError_t e = hipMemcpyAsync(drep, &hrep, sizeof hrep, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(dboff, hboff, sizeof hboff, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    k_fill_u32<<<grid_for(nbins + 1, 256), 256, 0, s>>>(dh, nbins + 1, 0u);
    if (N) {
      // cells of the current state on the SDF's checkerboard (origin drawn
      // from the current counter; the next apply rebuilds its own)
      k_setup<<<1, 64, 0, s>>>(g, c->st, 0, 0, 0, 1.0, 1.0, 0, 0, 0u, 0);
      if (g.dim == 2)
        k_bin<2><<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->st, c->bodies, c->w);
      else
        k_bin<3><<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->st, c->bodies, c->w);
      k_scan_blocks<<<nsb, 256, 0, s>>>(c->w.cnt, c->w.start, c->w.block_sums, c->w.block_ne, g.n_spaces);
      k_scan_top<<<1, 1024, 0, s>>>(g, c->st, c->w.block_sums, c->w.block_ne, nsb, c->w.start, 0);
      k_scan_add<<<grid_for(g.n_spaces, 256), 256, 0, s>>>(c->w.start, c->w.block_sums, c->w.cnt, g.n_spaces);
      k_scatter<<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->w);
      if (g.dim == 2)
        k_sdf<2><<<grid_for(N, kBlock), kBlock, 0, s>>>(drep, dboff, 1u, N, c->bodies, c->w, xmax, dx, nbins, dh,
                                                        dh + nbins);
      else
        k_sdf<3><<<grid_for(N, kBlock), kBlock, 0, s>>>(drep, dboff, 1u, N, c->bodies, c->w, xmax, dx, nbins, dh,
                                                        dh + nbins);
    }
    e = hipGetLastError();
  }
  std::vector<uint32_t> h(nbins + 1, 0u);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), dh, h.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFreeAsync(buf, s);
  if (e != hipSuccess) return fail(c, HMC_ERR_DEVICE, std::string("sdf: ") + hipGetErrorString(e));
  ApplyState st;  // a body outside the checkerboard is reported here
  int rc = read_state(c, &st);
  if (rc) return rc;
  for (uint32_t k = 0; k < nbins; ++k) hist[k] = h[k];
  if (overlaps) *overlaps = h[nbins];
  return HMC_OK;
}

int hmc_delta_energy_insert(hmc_ctx* c, const double* body_pos, uint64_t n, double* energy, double* bound) {
  if (!c || (n && (!body_pos || !energy))) return HMC_ERR_INVALID;
  if (!c->have_state) return fail(c, HMC_ERR_STATE, "hmc_set_state was not called");
  const Geometry& g = c->g;
  if (g.model == HMC_CONVEX_POLYHEDRON)
    return fail(c, HMC_ERR_UNSUPPORTED, "insertion energies of oriented bodies are not offloaded");
  if (n >= 0xffffffffull) return fail(c, HMC_ERR_INVALID, "too many candidates");
  if (n == 0) return HMC_OK;
  const int D = g.dim;
  const uint32_t N = g.n_bodies;
  if (N == 0) {  // no sites: every insertion energy is 0
    for (uint64_t k = 0; k < n; ++k) {
      energy[k] = 0.0;
      if (bound) bound[k] = 0.0;
    }
    return HMC_OK;
  }
  HIPCHK(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  double* dc = nullptr;
  double* dout = nullptr;
  uint32_t* flag = nullptr;
  hipError_t e = hipMallocAsync((void**)&dc, (size_t)n * D * sizeof(double), s);
  if (e == hipSuccess) e = hipMallocAsync((void**)&dout, 2 * (size_t)n * sizeof(double), s);
  if (e == hipSuccess) e = hipMallocAsync((void**)&flag, sizeof(uint32_t), s);
  if (e == hipSuccess) e = hipMemcpyAsync(dc, body_pos, (size_t)n * D * sizeof(double), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemsetAsync(flag, 0, sizeof(uint32_t), s);
  if (e == hipSuccess) {
    // cells of the current state, as for the total energy
    k_setup<<<1, 64, 0, s>>>(g, c->st, 0, 0, 0, 1.0, 1.0, 0, 0, 0u, 0);
    if (D == 2)
      k_bin<2><<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->st, c->bodies, c->w);
    else
      k_bin<3><<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->st, c->bodies, c->w);
    k_scan_blocks<<<c->n_scan_blocks, 256, 0, s>>>(c->w.cnt, c->w.start, c->w.block_sums, c->w.block_ne, g.n_spaces);
    k_scan_top<<<1, 1024, 0, s>>>(g, c->st, c->w.block_sums, c->w.block_ne, c->n_scan_blocks, c->w.start, 0);
    k_scan_add<<<grid_for(g.n_spaces, 256), 256, 0, s>>>(c->w.start, c->w.block_sums, c->w.cnt, g.n_spaces);
    k_scatter<<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->w);
    if (D == 2)
      k_insert_energy<2><<<grid_for(n, kEnergyBlock), kEnergyBlock, 0, s>>>(g, c->st, c->bodies, c->w, dc,
                                                                           (uint32_t)n, dout, flag);
    else
      k_insert_energy<3><<<grid_for(n, kEnergyBlock), kEnergyBlock, 0, s>>>(g, c->st, c->bodies, c->w, dc,
                                                                           (uint32_t)n, dout, flag);
    e = hipGetLastError();
  }
  std::vector<double> hout(2 * (size_t)n);
  uint32_t outside = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(hout.data(), dout, 2 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(&outside, flag, sizeof(uint32_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (dc) (void)hipFreeAsync(dc, s);
  if (dout) (void)hipFreeAsync(dout, s);
  if (flag) (void)hipFreeAsync(flag, s);
  if (e != hipSuccess) return fail(c, HMC_ERR_DEVICE, std::string("insertion energy: ") + hipGetErrorString(e));
  ApplyState st;
  int rc = read_state(c, &st);
  if (rc) return rc;
  if (outside) return fail(c, HMC_ERR_COVER, "a candidate lies outside the checkerboard");
  for (uint64_t k = 0; k < n; ++k) {
    energy[k] = hout[2 * k];
    if (bound) bound[k] = hout[2 * k + 1];
  }
  return HMC_OK;
}

int hmc_total_energy(hmc_ctx* c, double* energy, double* bound) {
  if (!c || !energy) return HMC_ERR_INVALID;
  if (!c->have_state) return fail(c, HMC_ERR_STATE, "hmc_set_state was not called");
  const Geometry& g = c->g;
  const uint32_t N = g.n_bodies;
  HIPCHK(c, hipSetDevice(c->device));
  if (N == 0) {
    *energy = 0.0;
    if (bound) *bound = 0.0;
    return HMC_OK;
  }
  hipStream_t s = c->stream;
  const uint32_t nblk = grid_for(N, kEnergyBlock);
  double* part = nullptr;
  HIPCHK(c, hipMallocAsync((void**)&part, (4 * (size_t)nblk + 4) * sizeof(double), s));
  // cells for the neighbour search: the checkerboard of the current counter
  k_setup<<<1, 64, 0, s>>>(g, c->st, 0, 0, 0, 1.0, 1.0, 0, 0, 0u, 0);
  if (g.dim == 2)
    k_bin<2><<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->st, c->bodies, c->w);
  else
    k_bin<3><<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->st, c->bodies, c->w);
  k_scan_blocks<<<c->n_scan_blocks, 256, 0, s>>>(c->w.cnt, c->w.start, c->w.block_sums, c->w.block_ne, g.n_spaces);
  k_scan_top<<<1, 1024, 0, s>>>(g, c->st, c->w.block_sums, c->w.block_ne, c->n_scan_blocks, c->w.start, 0);
  k_scan_add<<<grid_for(g.n_spaces, 256), 256, 0, s>>>(c->w.start, c->w.block_sums, c->w.cnt, g.n_spaces);
  k_scatter<<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->w);
  if (g.dim == 2)
    k_energy<2><<<nblk, kEnergyBlock, 0, s>>>(g, c->st, c->bodies, c->w, part);
  else
    k_energy<3><<<nblk, kEnergyBlock, 0, s>>>(g, c->st, c->bodies, c->w, part);
  k_energy_final<<<1, kEnergyFinal, 0, s>>>(part, nblk, part + 4 * (size_t)nblk);
  double out[4] = {0, 0, 0, 0};
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, part + 4 * (size_t)nblk, sizeof out, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFreeAsync(part, s);
  if (e != hipSuccess) return fail(c, HMC_ERR_DEVICE, std::string("total energy: ") + hipGetErrorString(e));
  // the next apply rebuilds the cells; the neighbour copies were not touched
  ApplyState st;
  int rc = read_state(c, &st);
  if (rc) return rc;
  // any two summation orders of n terms differ by at most 2 gamma_{n-1} sum|terms|
  const double n = std::max(out[2], 1.0);
  const double gam = n * 0x1p-53 / (1.0 - n * 0x1p-53);
  *energy = out[3] != 0.0 ? std::numeric_limits<double>::infinity() : out[0];
  if (bound) *bound = out[3] != 0.0 ? 0.0 : 2.0 * gam * out[1];
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
  int rc = read_state(c, &s);
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
    if (i == 0 || !use_graph) {
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
  out->tiled = b->tiled ? 1u : 0u;
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
                  (void*)b->w.block_ne, (void*)b->w.acc_slots, (void*)b->dtcs, (void*)b->dtoff, (void*)b->dtrep})
    if (p) (void)hipFree(p);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

}  // extern "C"
