// MI355X (gfx950) implementation of hoomd-mc's checkerboard-parallel local
// trial-move sweep for single-site point bodies in a periodic cuboid with a
// hard-sphere pair energy.  Exposed through the C ABI in include/hmc.h.
//
// Per apply (ParallelSweep::apply, hoomd-mc/src/parallel_sweep.rs:390-421):
//   setup    draw the checkerboard origin from Counter(step, substep, seed)
//            .index(u64::MAX) (parallel_sweep.rs:182, hypercuboid.rs:315-341)
//   build    counting sort of the bodies into checkerboard spaces: bin
//            (atomic histogram) -> exclusive scan -> scatter -> place.  place
//            orders each space by body index (the order ParallelSweep's
//            per-space lists have, parallel_sweep.rs:194-200) and moves the
//            body/site data into the new space-sorted slots, so every later
//            neighbour gather reads contiguous memory.
//   passes   for each colour cycle and colour: one lane per non-empty space
//            of that colour runs the whole trial (choose, propose, wrap, space
//            check, overlap scan of the 3^D neighbour spaces, accept, in-place
//            write).  Same-colour spaces never touch each other's neighbour
//            spaces, so in-place writes need no synchronisation.
//   guard    the number of colour cycles is ceil(N / non-empty spaces); the
//            host launches its best guess of pass kernels and one
//            single-workgroup guard kernel runs any cycles beyond it.
//
// Data layout in HBM: slot-ordered (space-sorted) structure of arrays,
// double-buffered between applies: body index (u32), body position (double4),
// site position + ghost flags (double4), site position in FP32 + flags
// (float4, mixed mode only).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hmc.h"
#include "dev_math.hpp"
#include "dev_rng.hpp"

namespace hmc {

constexpr int kBlock = 256;
constexpr int kScanBlock = 1024;  // elements per scan block (256 threads x 4)
constexpr int kGuardThreads = 1024;

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
  int32_t reach[3];  // neighbour spaces scanned on each side (1, or 2 when width ~ sigma)
  double sigma2;
  double edge2;     // sigma^2 (1 - 2^-30): at or above it the spatial-structure rule is applied
  double vc_width;  // VecCell cell width of the reference microstate
  double vc_k;      // VecCell stencil radius ceil(sigma / vc_width)
  float band_lo, band_hi;  // mixed mode: d2 < band_lo overlaps, d2 >= band_hi does not
};

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
  uint32_t colour_count[8];
  uint32_t error;
  uint32_t pad;
  unsigned long long accepted_total;
  unsigned long long trials_total;
  unsigned long long cycles_total;
  unsigned long long band_retests;
  unsigned long long exp_slow;
  unsigned long long guard_cycles;
  unsigned long long trace_n;
};

struct Slots {
  uint32_t* body;
  double4* bpos;
  double4* spos;
  float4* spos32;
};

struct Work {
  uint32_t* cnt;
  uint32_t* start;
  uint32_t* elem_space;
  uint32_t* elem_rank;
  uint32_t* tmp_body;
  uint32_t* tmp_src;
  uint32_t* colour_list;
  uint32_t* block_sums;
  hmc_trace_record* trace;
  uint64_t trace_cap;
  int tracing;
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
// 1 = image at x + L, 2 = image at x - L; all zero when the site lies outside
// [-L/2, L/2) on any axis (the reference then places no ghosts).
template <int D>
__device__ __forceinline__ uint32_t ghost_flags(const Geometry& g, const double* x) {
  uint32_t f = 0;
  for (int a = 0; a < D; ++a) {
    if (x[a] < -g.half[a] || x[a] >= g.half[a]) return 0;
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

template <int D>
__device__ __forceinline__ uint32_t colour_of(const Geometry& g, uint32_t s) {
  uint32_t c = 0;
  for (int a = 0; a < D; ++a) {
    uint32_t m = (s / g.stride[a]) % (uint32_t)g.shape[a];
    c = (c << 1) | (m & 1u);
  }
  return c;
}

__device__ __forceinline__ void pack_site(double4& out, const double* x, uint32_t flags) {
  out.x = x[0];
  out.y = x[1];
  out.z = x[2];
  out.w = __longlong_as_double((long long)flags);
}
__device__ __forceinline__ void pack_site32(float4& out, const double* x, uint32_t flags) {
  out.x = (float)x[0];
  out.y = (float)x[1];
  out.z = (float)x[2];
  out.w = __int_as_float((int)flags);
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
__global__ void k_setup(Geometry g, ApplyState* st, uint64_t step, uint32_t substep, uint32_t seed, double kT,
                        double max_move, int from_args, int advance) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
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
  for (int c = 0; c < 8; ++c) st->colour_count[c] = 0;
  st->cycles = 0;
  st->nonempty = 0;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <int D>
__global__ void __launch_bounds__(kBlock) k_bin(Geometry g, ApplyState* st, Slots cur, Work w) {
  uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.n_bodies) return;
  double4 p4 = cur.bpos[e];
  double x[3] = {p4.x, p4.y, p4.z};
  uint32_t s = space_of<D>(g, st->origin, x);
  if (s == UINT32_MAX) {
    atomicOr(&st->error, 1u);
    s = 0;
  }
  w.elem_space[e] = s;
  w.elem_rank[e] = atomicAdd(&w.cnt[s], 1u);
}

// Exclusive scan of cnt[0..n) into start[0..n], start[n] = total.
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

__global__ void __launch_bounds__(256) k_scan_blocks(const uint32_t* cnt, uint32_t* start, uint32_t* block_sums,
                                                    uint32_t n) {
  __shared__ uint32_t lds[256];
  uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * 4;
  uint32_t v[4];
  uint32_t sum = 0;
  for (int k = 0; k < 4; ++k) {
    v[k] = (base + k < n) ? cnt[base + k] : 0;
    sum += v[k];
  }
  uint32_t total;
  uint32_t off = block_exclusive_scan(sum, lds, &total);
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) start[base + k] = off;
    off += v[k];
  }
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// Single workgroup: exclusive scan of the block sums (up to 1024 * 16).
__global__ void __launch_bounds__(1024) k_scan_top(uint32_t* block_sums, uint32_t nblocks, uint32_t* start,
                                                  uint32_t n) {
  __shared__ uint32_t lds[1024];
  const int per = 16;
  uint32_t base = threadIdx.x * per;
  uint32_t v[per];
  uint32_t sum = 0;
  for (int k = 0; k < per; ++k) {
    v[k] = (base + k < nblocks) ? block_sums[base + k] : 0;
    sum += v[k];
  }
  uint32_t total;
  uint32_t off = block_exclusive_scan(sum, lds, &total);
  for (int k = 0; k < per; ++k) {
    if (base + k < nblocks) block_sums[base + k] = off;
    off += v[k];
  }
  if (threadIdx.x == 0) start[n] = total;
}

__global__ void k_scan_add(uint32_t* start, const uint32_t* block_sums, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) start[i] += block_sums[i / kScanBlock];
}

__global__ void k_scatter(Geometry g, Slots cur, Work w) {
  uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.n_bodies) return;
  uint32_t dst = w.start[w.elem_space[e]] + w.elem_rank[e];
  w.tmp_body[dst] = cur.body[e];
  w.tmp_src[dst] = e;
}

// Order each space by body index and move the data into the new slots.
template <int D, int PREC>
__global__ void __launch_bounds__(kBlock) k_place(Geometry g, ApplyState* st, Slots cur, Slots nxt, Work w) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.n_bodies) return;
  uint32_t e = w.tmp_src[p];
  uint32_t s = w.elem_space[e];
  uint32_t b = w.tmp_body[p];
  uint32_t s0 = w.start[s], s1 = w.start[s + 1];
  uint32_t rank = 0;
  for (uint32_t q = s0; q < s1; ++q) rank += w.tmp_body[q] < b ? 1u : 0u;
  uint32_t dst = s0 + rank;
  nxt.body[dst] = b;
  nxt.bpos[dst] = cur.bpos[e];
  nxt.spos[dst] = cur.spos[e];
  if (PREC == HMC_MIXED) nxt.spos32[dst] = cur.spos32[e];
  if (rank == 0) {
    uint32_t c = colour_of<D>(g, s);
    uint32_t idx = atomicAdd(&st->colour_count[c], 1u);
    w.colour_list[c * g.per_colour + idx] = s;
  }
}

__global__ void k_finish_build(Geometry g, ApplyState* st) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t ne = 0;
  for (int c = 0; c < (1 << g.dim); ++c) ne += st->colour_count[c];
  st->nonempty = ne;
  uint32_t cycles = ne ? (g.n_bodies + ne - 1) / ne : 0;
  st->cycles = cycles;
  st->cycles_total += cycles;
  st->trials_total += (unsigned long long)cycles * ne;
}

// Ball::sample (hoomd-vector/src/distribution.rs:39-70) added to the body
// position (Translate::propose, hoomd-mc/src/translate/cartesian.rs:39-52).
template <int D>
__device__ __forceinline__ void propose_translate(Sfc64& rng, double r, double* x, unsigned* exp_slow) {
  double v[3];
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
    double p0 = standard_normal(rng, exp_slow);
    double p1 = standard_normal(rng, exp_slow);
    double p2 = standard_normal(rng, exp_slow);
    double norm = sqrt(fma(p2, p2, fma(p1, p1, fma(p0, p0, 0.0))));
    double q = r / norm;
    double sc = q * hmc_cbrt(uniform01(rng));
    v[0] = p0 * sc;
    v[1] = p1 * sc;
    v[2] = p2 * sc;
  }
  for (int a = 0; a < D; ++a) x[a] += v[a];
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
__device__ __noinline__ bool overlap64_edge(const Geometry& g, const double* t, double4 sj) {
  uint32_t fl = (uint32_t)__double_as_longlong(sj.w);
  double xj[3] = {sj.x, sj.y, sj.z};
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
    if (!any) return false;
    d2 = d2 + best * best;
  }
  return d2 < g.sigma2;
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

struct TrialOut {
  unsigned accepted;
  unsigned retests;
  unsigned exp_slow;
};

// One trial in space s (parallel_sweep.rs:234-277).
template <int D, int PREC>
__device__ __forceinline__ void trial(const Geometry& g, const ApplyState* st, const Slots& S, const Work& w,
                                      uint32_t s, uint32_t colour, uint32_t substep, TrialOut& o) {
  const uint32_t s0 = w.start[s];
  const uint32_t n = w.start[s + 1] - s0;

  // neighbour spaces (periodic) within reach on each axis, without repeats
  // when the axis has fewer than 2 * reach + 1 spaces
  uint32_t nb[3][5];
  int nn[3] = {1, 1, 1};
  for (int a = 0; a < D; ++a) {
    const int32_t sh = g.shape[a];
    const int32_t m = (int32_t)((s / g.stride[a]) % (uint32_t)sh);
    const int32_t r = g.reach[a];
    if (r == 1 && sh >= 3) {
      nb[a][0] = (uint32_t)(m == 0 ? sh - 1 : m - 1) * g.stride[a];
      nb[a][1] = (uint32_t)m * g.stride[a];
      nb[a][2] = (uint32_t)(m == sh - 1 ? 0 : m + 1) * g.stride[a];
      nn[a] = 3;
    } else {
      int k = 0;
      for (int o = -r; o <= r; ++o) {
        int32_t mm = ((m + o) % sh + sh) % sh;
        bool seen = false;
        for (int j = 0; j < k; ++j) seen |= nb[a][j] == (uint32_t)mm * g.stride[a];
        if (!seen) nb[a][k++] = (uint32_t)mm * g.stride[a];
      }
      nn[a] = k;
    }
  }

  Sfc64 rng = counter_rng(st->step, substep, st->seed, s);
  const uint32_t slot = s0 + choose_index(rng, n);
  const double4 bp = S.bpos[slot];
  double x[3] = {bp.x, bp.y, bp.z};
  unsigned exp_slow = 0;
  propose_translate<D>(rng, st->max_move, x, &exp_slow);
  o.exp_slow = exp_slow;
  for (int a = 0; a < D; ++a) x[a] = wrap1(x[a], g.L[a]);  // generate_trial_moves: wrap(propose(..))

  bool accept = false;
  if (space_of<D>(g, st->origin, x) == s) {
    // trial site: wrap(body.transform(site)) (pairwise_cutoff.rs:436)
    double t[3];
    for (int a = 0; a < D; ++a) t[a] = wrap1(x[a] + 0.0, g.L[a]);
    float t32[3];
    for (int a = 0; a < D; ++a) t32[a] = (float)t[a];
    bool hit = false;
    unsigned retests = 0;
    for (int i0 = 0; i0 < nn[0] && !hit; ++i0) {
      for (int i1 = 0; i1 < nn[1] && !hit; ++i1) {
        for (int i2 = 0; i2 < (D == 3 ? nn[2] : 1) && !hit; ++i2) {
          uint32_t ns = nb[0][i0] + nb[1][i1] + (D == 3 ? nb[2][i2] : 0u);
          uint32_t q0 = w.start[ns], q1 = w.start[ns + 1];
          for (uint32_t q = q0; q < q1; ++q) {
            if (q == slot) continue;
            if (PREC == HMC_MIXED) {
              int r = overlap32<D>(g, t32, S.spos32[q]);
              if (r == 2) {
                ++retests;
                r = overlap64<D>(g, t, S.spos[q]) ? 1 : 0;
              }
              if (r) {
                hit = true;
                break;
              }
            } else if (overlap64<D>(g, t, S.spos[q])) {
              hit = true;
              break;
            }
          }
        }
      }
    }
    o.retests = retests;
    // Hard spheres: dH is 0 or +inf, so u is never drawn (parallel_sweep.rs:267-273).
    accept = !hit;
    if (accept) {
      // update_body_properties (microstate.rs:985-1034): body = wrap(props),
      // site = wrap(body + 0)
      double nbp[3] = {0.0, 0.0, 0.0}, nsp[3] = {0.0, 0.0, 0.0};
      for (int a = 0; a < D; ++a) nbp[a] = wrap1(x[a], g.L[a]);
      for (int a = 0; a < D; ++a) nsp[a] = wrap1(nbp[a] + 0.0, g.L[a]);
      uint32_t fl = ghost_flags<D>(g, nsp);
      S.bpos[slot] = make_double4(nbp[0], nbp[1], nbp[2], 0.0);
      double4 sp;
      pack_site(sp, nsp, fl);
      S.spos[slot] = sp;
      if (PREC == HMC_MIXED) {
        float4 s32;
        pack_site32(s32, nsp, fl);
        S.spos32[slot] = s32;
      }
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
      r.body = (int64_t)S.body[slot];
      r.status = accept ? 1 : 0;
      r.pad = 0;
      w.trace[idx] = r;
    }
  }
}

// One colour of one cycle: one lane per non-empty space of the colour.
template <int D, int PREC>
__global__ void __launch_bounds__(kBlock) k_pass(Geometry g, ApplyState* st, Slots S, Work w, uint32_t colour,
                                                 uint32_t cycle, uint32_t xcd_chunk) {
  if (cycle >= st->cycles) return;
  // XCD-aware block order: consecutive logical blocks (neighbouring spaces)
  // land on the same XCD and share its L2.
  uint32_t b = blockIdx.x;
  uint32_t lb = (b & 7u) * xcd_chunk + (b >> 3);
  uint32_t t = lb * kBlock + threadIdx.x;
  uint32_t count = st->colour_count[colour];
  if (t >= count) return;
  uint32_t s = w.colour_list[colour * g.per_colour + t];
  TrialOut o{0, 0, 0};
  trial<D, PREC>(g, st, S, w, s, colour, st->substep0 + cycle, o);
  wave_count(&st->accepted_total, o.accepted != 0);
  lane_add(&st->band_retests, o.retests);
  lane_add(&st->exp_slow, o.exp_slow);
}

// Runs the cycles beyond the ones the host launched, in one workgroup.
template <int D, int PREC>
__global__ void __launch_bounds__(kGuardThreads) k_guard(Geometry g, ApplyState* st, Slots S, Work w,
                                                         uint32_t first_cycle) {
  const uint32_t cycles = st->cycles;
  const int ncol = 1 << D;
  for (uint32_t k = first_cycle; k < cycles; ++k) {
    for (int c = 0; c < ncol; ++c) {
      const uint32_t count = st->colour_count[c];
      const uint32_t rounds = (count + kGuardThreads - 1) / kGuardThreads;
      for (uint32_t r = 0; r < rounds; ++r) {
        const uint32_t t = r * kGuardThreads + threadIdx.x;
        TrialOut o{0, 0, 0};
        if (t < count) {
          uint32_t s = w.colour_list[c * g.per_colour + t];
          trial<D, PREC>(g, st, S, w, s, (uint32_t)c, st->substep0 + k, o);
        }
        wave_count(&st->accepted_total, o.accepted != 0);
        lane_add(&st->band_retests, o.retests);
        lane_add(&st->exp_slow, o.exp_slow);
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) st->guard_cycles += 1;
  }
}

// Initial site positions and flags from body positions.
template <int D, int PREC>
__global__ void k_init_sites(Geometry g, Slots S, int have_sites) {
  uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.n_bodies) return;
  double4 b = S.bpos[e];
  double x[3] = {b.x, b.y, b.z};
  double s[3] = {0.0, 0.0, 0.0};
  if (have_sites) {
    double4 sp = S.spos[e];
    s[0] = sp.x;
    s[1] = sp.y;
    s[2] = sp.z;
  } else {
    for (int a = 0; a < D; ++a) s[a] = wrap1(x[a] + 0.0, g.L[a]);
  }
  uint32_t fl = ghost_flags<D>(g, s);
  double4 sp;
  pack_site(sp, s, fl);
  S.spos[e] = sp;
  float4 s32;
  pack_site32(s32, s, fl);
  S.spos32[e] = s32;
  S.body[e] = e;
}

// ------------------------------------------------------------- debug hooks --
__global__ void k_debug_math(int kind, const double* in, double* out, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r = 0.0;
  switch (kind) {
    case 0: r = hmc_cbrt(in[i]); break;
    case 1: r = hmc_log_cr(in[i]); break;
    case 2: r = hmc_exp_cr(in[i]); break;
    case 3: r = sqrt(in[i]); break;
    default: r = in[2 * i] / in[2 * i + 1]; break;
  }
  out[i] = r;
}

__global__ void k_debug_normals(uint64_t step, uint32_t substep, uint32_t seed, uint64_t index, double* out,
                                uint64_t n) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  Sfc64 r = counter_rng(step, substep, seed, index);
  unsigned slow = 0;
  for (uint64_t i = 0; i < n; ++i) out[i] = standard_normal(r, &slow);
}

}  // namespace hmc

// ======================================================================= host
using namespace hmc;

struct hmc_ctx {
  hmc_config cfg;
  Geometry g;
  int device;
  hipStream_t stream = nullptr;
  Slots a{}, b{};  // a: current state, b: next
  Work w{};
  ApplyState* st = nullptr;
  bool have_state = false;
  uint32_t cycles_guess = 1;
  uint32_t n_scan_blocks = 0;
  uint64_t applies = 0;
  std::string err;
  bool profiling = false;
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
                 double max_move) {
  const Geometry& g = c->g;
  hipStream_t s = c->stream;
  const uint32_t N = g.n_bodies;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (c->profiling) {
    e0 = take_event(c);
    hipEventRecord(e0, s);
  }
  k_setup<<<1, 64, 0, s>>>(g, c->st, step, substep, seed, kT, max_move, from_args, advance);
  k_fill_u32<<<grid_for(g.n_spaces, 256), 256, 0, s>>>(c->w.cnt, g.n_spaces, 0u);
  k_bin<D><<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->st, c->a, c->w);
  k_scan_blocks<<<c->n_scan_blocks, 256, 0, s>>>(c->w.cnt, c->w.start, c->w.block_sums, g.n_spaces);
  k_scan_top<<<1, 1024, 0, s>>>(c->w.block_sums, c->n_scan_blocks, c->w.start, g.n_spaces);
  k_scan_add<<<grid_for(g.n_spaces, 256), 256, 0, s>>>(c->w.start, c->w.block_sums, g.n_spaces);
  k_scatter<<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->a, c->w);
  k_place<D, PREC><<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->st, c->a, c->b, c->w);
  k_finish_build<<<1, 64, 0, s>>>(g, c->st);
  if (c->profiling) {
    e1 = take_event(c);
    hipEventRecord(e1, s);
    c->build_events.push_back({e0, e1});
  }
  // passes on the new slots
  const uint32_t blocks = grid_for(g.per_colour, kBlock);
  const uint32_t chunk = (blocks + 7) / 8;
  const uint32_t grid = chunk * 8;
  for (uint32_t k = 0; k < c->cycles_guess; ++k) {
    for (int col = 0; col < (1 << D); ++col) {
      hipEvent_t p0 = nullptr, p1 = nullptr;
      if (c->profiling) {
        p0 = take_event(c);
        hipEventRecord(p0, s);
      }
      k_pass<D, PREC><<<grid, kBlock, 0, s>>>(g, c->st, c->b, c->w, (uint32_t)col, k, chunk);
      if (c->profiling) {
        p1 = take_event(c);
        hipEventRecord(p1, s);
        c->pass_events.push_back({p0, p1});
      }
    }
  }
  k_guard<D, PREC><<<1, kGuardThreads, 0, s>>>(g, c->st, c->b, c->w, c->cycles_guess);
  std::swap(c->a, c->b);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(c, HMC_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
  c->applies += 1;
  return HMC_OK;
}

int dispatch_apply(hmc_ctx* c, int from_args, int advance, uint64_t step, uint32_t substep, uint32_t seed,
                   double kT, double max_move) {
  int d = c->g.dim, p = c->cfg.precision;
  if (d == 2 && p == HMC_FP64) return launch_apply<2, HMC_FP64>(c, from_args, advance, step, substep, seed, kT, max_move);
  if (d == 2) return launch_apply<2, HMC_MIXED>(c, from_args, advance, step, substep, seed, kT, max_move);
  if (p == HMC_FP64) return launch_apply<3, HMC_FP64>(c, from_args, advance, step, substep, seed, kT, max_move);
  return launch_apply<3, HMC_MIXED>(c, from_args, advance, step, substep, seed, kT, max_move);
}

int read_state(hmc_ctx* c, ApplyState* out) {
  HIPCHK(c, hipMemcpyAsync(out, c->st, sizeof(ApplyState), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->profiling) {
    for (auto& pr : c->pass_events) {
      float ms = 0.f;
      hipEventElapsedTime(&ms, pr.first, pr.second);
      c->pass_ms += ms;
      c->pass_n += 1;
      c->ev_pool.push_back(pr.first);
      c->ev_pool.push_back(pr.second);
    }
    for (auto& pr : c->build_events) {
      float ms = 0.f;
      hipEventElapsedTime(&ms, pr.first, pr.second);
      c->build_ms += ms;
      c->build_n += 1;
      c->ev_pool.push_back(pr.first);
      c->ev_pool.push_back(pr.second);
    }
    c->pass_events.clear();
    c->build_events.clear();
  }
  if (out->error) return fail(c, HMC_ERR_COVER, "a body lies outside the checkerboard");
  return HMC_OK;
}

}  // namespace

extern "C" {

int hmc_create(const hmc_config* cfg, hmc_ctx** out) {
  if (!cfg || !out) return HMC_ERR_INVALID;
  *out = nullptr;
  if (cfg->dim != 2 && cfg->dim != 3) return HMC_ERR_UNSUPPORTED;
  if (cfg->model != HMC_HARD_SPHERE) return HMC_ERR_UNSUPPORTED;
  if (cfg->precision != HMC_FP64 && cfg->precision != HMC_MIXED) return HMC_ERR_INVALID;
  if (cfg->n_bodies >= 0xffffffffull) return HMC_ERR_UNSUPPORTED;
  const int D = (int)cfg->dim;
  const double sigma = cfg->params[0];
  if (!(sigma > 0.0) || !(cfg->body_interaction_range > 0.0) || !(cfg->boundary_range >= 0.0) ||
      !(cfg->spatial_cell_width > 0.0))
    return HMC_ERR_INVALID;
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

  auto* c = new hmc_ctx;
  c->cfg = *cfg;
  c->device = cfg->device;
  Geometry& g = c->g;
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
      delete c;
      return HMC_ERR_INVALID;
    }
    if (inside % 2 != 0) inside -= 1;
    g.shape[a] = (int32_t)inside;
    g.width[a] = L / (double)inside;
    // a pair closer than sigma lies in neighbouring spaces unless the width is
    // within rounding of sigma; then the scan reaches two spaces out
    g.reach[a] = g.width[a] > sigma * (1.0 + 1e-9) ? 1 : 2;
    spaces *= (uint64_t)inside;
  }
  if (spaces >= 0xffffffffull || spaces > (uint64_t)kScanBlock * 1024 * 16) {
    delete c;
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
    g.stride[a] = 0;
    g.L[a] = 1.0;
    g.Lf[a] = 1.0f;
    g.width[a] = 1.0;
  }
  g.n_spaces = (uint32_t)spaces;
  g.per_colour = g.n_spaces >> D;
  g.sigma2 = sigma * sigma;
  g.edge2 = g.sigma2 * (1.0 - 0x1p-30);
  g.vc_width = cfg->spatial_cell_width;
  g.vc_k = std::ceil(sigma / cfg->spatial_cell_width);
  // Mixed mode band.  Per axis |d32 - d64| <= E with
  // E = 2^-24 (3.5 Lmax + 2 sigma) + 2^-53 (1.5 Lmax + 2 sigma), from the
  // roundings of t, x, L, x + T and t - x; the band on d^2 near
  // sigma^2 is 2 E sqrt(D) d + D E^2 + 4 (2^-24 + 2^-53) d^2 at d = 1.01 sigma.
  double lmax = 0.0;
  for (int a = 0; a < D; ++a) lmax = std::max(lmax, g.L[a]);
  double E = (0x1p-24 * (3.5 * lmax + 2.0 * sigma) + 0x1p-53 * (1.5 * lmax + 2.0 * sigma)) * 1.01;
  double dref = 1.01 * sigma;
  double B = (2.0 * E * std::sqrt((double)D) * dref + D * E * E + 4.0 * (0x1p-24 + 0x1p-53) * dref * dref) * 1.1;
  if (cfg->precision == HMC_MIXED && !(B < 0.01 * g.sigma2)) {
    delete c;
    return HMC_ERR_UNSUPPORTED;
  }
  // the float thresholds are rounded outward
  g.band_lo = (float)(g.edge2 - B);
  if ((double)g.band_lo > g.edge2 - B) g.band_lo = std::nextafter(g.band_lo, 0.0f);
  g.band_hi = (float)(g.sigma2 + B);
  if ((double)g.band_hi < g.sigma2 + B) g.band_hi = std::nextafter(g.band_hi, 1e30f);

  if (hipSetDevice(c->device) != hipSuccess) {
    delete c;
    return HMC_ERR_DEVICE;
  }
  const uint32_t N = g.n_bodies;
  c->n_scan_blocks = grid_for(g.n_spaces, kScanBlock);
  bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess;
  for (Slots* sl : {&c->a, &c->b}) {
    ok = ok && dalloc(&sl->body, N) == hipSuccess && dalloc(&sl->bpos, N) == hipSuccess &&
         dalloc(&sl->spos, N) == hipSuccess && dalloc(&sl->spos32, N) == hipSuccess;
  }
  ok = ok && dalloc(&c->w.cnt, g.n_spaces) == hipSuccess && dalloc(&c->w.start, g.n_spaces + 1) == hipSuccess &&
       dalloc(&c->w.elem_space, N) == hipSuccess && dalloc(&c->w.elem_rank, N) == hipSuccess &&
       dalloc(&c->w.tmp_body, N) == hipSuccess && dalloc(&c->w.tmp_src, N) == hipSuccess &&
       dalloc(&c->w.colour_list, g.n_spaces) == hipSuccess &&
       dalloc(&c->w.block_sums, c->n_scan_blocks) == hipSuccess && dalloc(&c->st, 1) == hipSuccess;
  if (!ok) {
    hmc_destroy(c);
    return HMC_ERR_DEVICE;
  }
  if (hipMemset(c->st, 0, sizeof(ApplyState)) != hipSuccess) {
    hmc_destroy(c);
    return HMC_ERR_DEVICE;
  }
  // first guess of the number of colour cycles per apply: every space occupied
  c->cycles_guess = (uint32_t)std::max<uint64_t>(1, (N + g.n_spaces - 1) / std::max<uint32_t>(g.n_spaces, 1));
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
  HIPCHK(c, hipMemcpyAsync(c->a.bpos, hb.data(), N * sizeof(double4), hipMemcpyHostToDevice, c->stream));
  if (site_pos)
    HIPCHK(c, hipMemcpyAsync(c->a.spos, hs.data(), N * sizeof(double4), hipMemcpyHostToDevice, c->stream));
  if (N) {
    if (D == 2)
      k_init_sites<2, 0><<<grid_for(N, 256), 256, 0, c->stream>>>(c->g, c->a, site_pos ? 1 : 0);
    else
      k_init_sites<3, 0><<<grid_for(N, 256), 256, 0, c->stream>>>(c->g, c->a, site_pos ? 1 : 0);
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->have_state = true;
  return HMC_OK;
}

int hmc_get_state(hmc_ctx* c, double* body_pos, double* site_pos) {
  if (!c || !c->have_state) return c ? fail(c, HMC_ERR_STATE, "no state") : HMC_ERR_INVALID;
  const uint32_t N = c->g.n_bodies;
  const int D = c->g.dim;
  std::vector<uint32_t> body(N);
  std::vector<double4> hb(N), hs(N);
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(body.data(), c->a.body, N * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(hb.data(), c->a.bpos, N * sizeof(double4), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(hs.data(), c->a.spos, N * sizeof(double4), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (uint32_t i = 0; i < N; ++i) {
    uint32_t b = body[i];
    if (b >= N) return fail(c, HMC_ERR_STATE, "corrupt body index");
    double v[3] = {hb[i].x, hb[i].y, hb[i].z};
    double u[3] = {hs[i].x, hs[i].y, hs[i].z};
    for (int a = 0; a < D; ++a) {
      if (body_pos) body_pos[(size_t)b * D + a] = v[a];
      if (site_pos) site_pos[(size_t)b * D + a] = u[a];
    }
  }
  return HMC_OK;
}

int hmc_apply_translate(hmc_ctx* c, uint64_t step, uint32_t substep, uint32_t seed, double kT,
                        double max_distance, hmc_count* out, uint32_t* substeps_used) {
  if (!c) return HMC_ERR_INVALID;
  if (!c->have_state) return fail(c, HMC_ERR_STATE, "hmc_set_state was not called");
  if (!(max_distance > 0.0) || !(kT > 0.0)) return fail(c, HMC_ERR_INVALID, "max_distance and kT must be > 0");
  HIPCHK(c, hipSetDevice(c->device));
  ApplyState before;
  int rc = read_state(c, &before);
  if (rc) return rc;
  rc = dispatch_apply(c, 1, 0, step, substep, seed, kT, max_distance);
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

int hmc_run_translate(hmc_ctx* c, uint64_t step0, uint32_t seed, double kT, double max_distance,
                      uint32_t n_steps, hmc_count* out) {
  if (!c) return HMC_ERR_INVALID;
  if (!c->have_state) return fail(c, HMC_ERR_STATE, "hmc_set_state was not called");
  if (!(max_distance > 0.0) || !(kT > 0.0)) return fail(c, HMC_ERR_INVALID, "max_distance and kT must be > 0");
  HIPCHK(c, hipSetDevice(c->device));
  ApplyState before;
  int rc = read_state(c, &before);
  if (rc) return rc;
  for (uint32_t i = 0; i < n_steps; ++i) {
    rc = dispatch_apply(c, i == 0 ? 1 : 0, i == 0 ? 0 : 1, step0, 0, seed, kT, max_distance);
    if (rc) return rc;
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

int hmc_trace_enable(hmc_ctx* c, int on, uint64_t capacity) {
  if (!c) return HMC_ERR_INVALID;
  HIPCHK(c, hipSetDevice(c->device));
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
  out->exp_slow_path = s.exp_slow;
  out->guard_cycles = s.guard_cycles;
  return HMC_OK;
}

int hmc_profile_enable(hmc_ctx* c, int on) {
  if (!c) return HMC_ERR_INVALID;
  c->profiling = on != 0;
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
    hipFree(din);
    return HMC_ERR_DEVICE;
  }
  int rc = HMC_OK;
  if (hipMemcpy(din, in, nin * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) rc = HMC_ERR_DEVICE;
  if (!rc) {
    k_debug_math<<<(uint32_t)((n + 255) / 256), 256>>>(kind, din, dout, n);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = HMC_ERR_DEVICE;
  }
  if (!rc && hipMemcpy(out, dout, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) rc = HMC_ERR_DEVICE;
  hipFree(din);
  hipFree(dout);
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
  hipFree(d);
  return rc;
}

const char* hmc_last_error(const hmc_ctx* c) { return c ? c->err.c_str() : "null context"; }

void hmc_destroy(hmc_ctx* c) {
  if (!c) return;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  for (Slots* sl : {&c->a, &c->b}) {
    hipFree(sl->body);
    hipFree(sl->bpos);
    hipFree(sl->spos);
    hipFree(sl->spos32);
  }
  hipFree(c->w.cnt);
  hipFree(c->w.start);
  hipFree(c->w.elem_space);
  hipFree(c->w.elem_rank);
  hipFree(c->w.tmp_body);
  hipFree(c->w.tmp_src);
  hipFree(c->w.colour_list);
  hipFree(c->w.block_sums);
  hipFree(c->w.trace);
  hipFree(c->st);
  for (auto& pr : c->pass_events) c->ev_pool.push_back(pr.first), c->ev_pool.push_back(pr.second);
  for (auto& pr : c->build_events) c->ev_pool.push_back(pr.first), c->ev_pool.push_back(pr.second);
  for (hipEvent_t e : c->ev_pool) hipEventDestroy(e);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

}  // extern "C"
