// Device-side types, constants and helpers shared by the sweep's kernels.
// One translation unit per kernel family: k_build (setup, counting sort,
// scans), k_hard (hard-shape passes of a context and of a batch), k_persist
// (the persistent apply), k_lj (isotropic pair energies), k_convex (convex
// polyhedra), k_measure (total energy, insertion energies, SDF), k_misc
// (initial sites, boundary maps, test hooks); the host side is api_*.hip.
//
// Data in HBM:
//   bodies (body-index order, authoritative, FP64): body position double4,
//       site position double4 with the ghost-image flags in .w
//   cells (slot order = space order, rebuilt each apply): body index u32 and
//       the neighbour copy of the site (float4 mixed / double4 FP64)
#pragma once
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
  int cv_mixed;            // convex: FP32 separating-axis decisions with an FP64 XenoCollide band
  // A copy of this struct in device memory.  Out-of-line helpers receive this
  // pointer instead of the address of a kernel argument, which would make
  // every lane copy the whole struct to scratch memory at kernel entry.
  const Geometry* self;
};

// Mixed-precision hard convex polyhedra: the polytope's separating axes in
// FP32, built on the host from its vertices and passed to the convex pass by
// value (so that the pass reads them through the scalar cache).  Face axes
// and edge directions are kept once per +-pair.
constexpr uint32_t kSatFaces = 16;  // face axes of a polytope that the FP32 tests take
constexpr uint32_t kSatEdges = 8;   // edge directions that the edge-pair axes take
struct SatShape {
  float face[kSatFaces][4];   // unit face normal n and hmax = max over the vertices of n.v
  float fmin[kSatFaces];      // hmin = min over the vertices of n.v
  float edge[kSatEdges][4];   // unit edge direction
  float hv[kMaxVertices][4];  // the scanned vertices: all of them, or one of each antipodal pair (sym)
  uint32_t nf, ne, nh, sym;   // counts; sym: the vertex set is closed under negation
  float tol;                  // margin of every FP32 decision: 2^-12 R, R the circumsphere radius
  uint32_t use;               // 0: no FP32 tests (every pair goes to XenoCollide in FP64)
  uint32_t use_edges;         // the edge-pair axes are complete (ne <= kSatEdges)
  uint32_t pad;
};

// The isotropic pair energies, which run one trial per wavefront with a
// Metropolis decision on dH (the others are hard overlap tests).
__host__ __device__ inline bool is_isotropic(int model) {
  return model == HMC_LJ || model == HMC_LJ_SHIFTED || model == HMC_OVERLAP_PENALTY;
}

constexpr int kMaxMoves = 4;  // moves per step of hmc_run

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
  // hmc_run with several moves per step: accepted moves and trials per move
  // (k_tally), and the totals at the last tally
  unsigned long long move_acc[kMaxMoves];
  unsigned long long move_trials[kMaxMoves];
  unsigned long long tally_acc, tally_trials;
  uint32_t wait_at;  // the first neighbour wait that timed out: (region << 12) | (pass + 1)
  uint32_t pad2;
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
// The moved-body stamp an accepted move stores in the body position's w: the
// epoch (applies since the context's creation) of the apply that last moved
// the body; 0 for a body not moved since its state was set (hmc_moved_since).
__device__ __forceinline__ double move_stamp(uint32_t epoch) { return __longlong_as_double((long long)epoch); }
__device__ __forceinline__ uint32_t stamp_of(double w) { return (uint32_t)__double_as_longlong(w); }

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
// A pre-drawn stream read from a given offset.  Past its end it reports
// `over` and returns a word with which every sampler returns at once (a
// uniform of 1/2, a ziggurat draw in layer 1 with u = 0, an index of 0).
struct StreamRng {
  const uint64_t* x;
  uint32_t pos, end;
  bool over;
  __device__ __forceinline__ uint64_t next() {
    if (pos < end) return x[pos++];
    over = true;
    return 0x8000000000000001ull;
  }
  __device__ __forceinline__ uint32_t next_u32() { return (uint32_t)next(); }
};

// Per-lane add of a rare count.
__device__ __forceinline__ void lane_add(unsigned long long* dst, unsigned v) {
  if (v) atomicAdd(dst, (unsigned long long)v);
}

constexpr uint32_t kLookbackBits = 27;  // count bits of the look-back scan's status words

// Ball::sample (hoomd-vector/src/distribution.rs:39-70): the displacement of
// Translate::propose (hoomd-mc/src/translate/cartesian.rs:39-52).
template <int D, class R>
__device__ __forceinline__ void ball_sample(R& rng, double r, const double* zx, const double* zf, double* v) {
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

// Copy of the ziggurat tables in LDS (the normal sampler's table reads are
// data dependent; LDS keeps them off the global-memory latency path).
__device__ __forceinline__ void stage_zig(double* zx, double* zf) {
  for (int i = threadIdx.x; i < 257; i += blockDim.x) {
    zx[i] = kZigX[i];
    zf[i] = kZigF[i];
  }
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
  float band_lo, band_hi;  // FP32 band of the box-relative test (isotropic: of the r_cut membership test)
  float perr;              // isotropic persistent apply: bound on the FP32 displacement error per axis
  float beta_max;          // isotropic persistent apply: bound on |r2_32 - r2_ref| of a pair below band_hi
  float r2_open;           // isotropic persistent apply: pairs with FP32 r2 below it are left to FP64
  uint32_t probe;          // persistent apply: 1 = only check that every region is resident (region_probe)
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

// Persistent apply: words between region mailboxes (256 B: separate channels)
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
//
// A store of more than 8 bytes reads its data registers some cycles after it
// issues, and a VALU write to those registers in that window corrupts the
// stored value (two wait states on gfx940 and later).  The compiler inserts
// those wait states after the stores it emits itself but cannot see into
// inline asm, and takes the asm's inputs as consumed once the asm is issued.
// So both stores sit in one asm statement that ends with an s_nop covering
// the window (5 wait states) before any later instruction can overwrite
// their data.
__device__ __forceinline__ void st_dev_d4(double4* p, double4 v) {
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  const unsigned long long a = (unsigned long long)__double_as_longlong(v.x),
                           b = (unsigned long long)__double_as_longlong(v.y),
                           c = (unsigned long long)__double_as_longlong(v.z),
                           d = (unsigned long long)__double_as_longlong(v.w);
  const u32x4 lo = {(unsigned)a, (unsigned)(a >> 32), (unsigned)b, (unsigned)(b >> 32)};
  const u32x4 hi = {(unsigned)c, (unsigned)(c >> 32), (unsigned)d, (unsigned)(d >> 32)};
  asm volatile(
      "global_store_dwordx4 %0, %1, off sc1\n\t"
      "global_store_dwordx4 %0, %2, off offset:16 sc1\n\t"
      "s_nop 4" ::"v"(p),
      "v"(lo), "v"(hi)
      : "memory");
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
// The same from an FP32 r^2 (the conversion of r^2 it skips is within the
// bound above).
__device__ __forceinline__ float lj_energy32f(float sig2, float eps4, float shift, float r2, float* a) {
  const float s2 = sig2 * __builtin_amdgcn_rcpf(r2);
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

__device__ __forceinline__ void stage_verts(const Geometry& g, double* vx) {
  for (uint32_t i = threadIdx.x; i < 3 * g.nverts; i += blockDim.x) vx[i] = g.verts[i];
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

constexpr uint32_t kCvTrials = 4;  // convex pass: trials per wavefront
constexpr uint32_t kCvLanes = 64 / kCvTrials;

constexpr int kEnergyBlock = 256;
constexpr int kEnergyFinal = 256;

}  // namespace hmc
