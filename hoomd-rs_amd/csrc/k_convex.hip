// Hard convex polyhedra: translate and rotate trials with XenoCollide.
#include "sweep_device.hpp"

namespace hmc {

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
      B.bpos[body] = make_double4(nb3[0], nb3[1], nb3[2], move_stamp(st->epoch));
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

// ------------------------------------------------ FP32 separating axes --
// Mixed precision for hard convex polyhedra.  Shape B (the neighbour) is
// placed in the frame of shape A (the trial) as XenoCollide places it
// (convex.rs:154-170): centre v = conj(q_i)(r_j - r_i), rotation R =
// rotmat(conj(q_i) q_j), both formed in FP32.  Every FP32 quantity then lies
// within 500 u R of its exact value (u = 2^-24, R the circumsphere radius,
// |v| < 2R for a listed pair, unit axes), and every decision below keeps a
// margin tol = 2^-12 R = 4096 u R:
// - apart: the extents of A and B along some axis w are separated by more
//   than tol (any such w is a separating direction, so A and B are disjoint by
//   more than XenoCollide's 2e-12 tolerance and its FP64 rounding);
// - overlap: a vertex of one shape lies inside every face slab of the other
//   by more than tol, or every separating-axis candidate (the face normals of
//   A and B and the cross products of their edge directions, which include
//   every facet normal of the Minkowski difference) shows an overlap depth
//   above tol; the origin is then inside the Minkowski difference by a margin
//   that no FP64 rounding in XenoCollide reaches, so XenoCollide reports an
//   overlap;
// - anything else (near contact, near-parallel edges) is left to FP64
//   XenoCollide.
// So every decision equals the FP64 one.

// Extent [*lo, *hi] of the polytope along w (in its own frame); with IN,
// bit k of the result is set where off + w.v_k lies in (a, b), bit 32 + k
// where off - w.v_k does (the antipode of a scanned vertex of a symmetric
// polytope).
template <int NH, bool SYM, bool IN>
__device__ __forceinline__ uint64_t sat_scan(const SatShape& S, float wx, float wy, float wz, float off, float a,
                                             float b, float* lo, float* hi) {
  float mx = -INFINITY, mn = INFINITY;
  uint64_t keep = 0;
  if constexpr (NH > 0) {
#pragma unroll
    for (int k = 0; k < NH; ++k) {
      const float x = fmaf(S.hv[k][2], wz, fmaf(S.hv[k][1], wy, S.hv[k][0] * wx));
      mx = fmaxf(mx, x);
      mn = fminf(mn, x);
      if (IN) {
        const float p = off + x, q = off - x;
        keep |= (uint64_t)(p > a && p < b) << k;
        if (SYM) keep |= (uint64_t)(q > a && q < b) << (32 + k);
      }
    }
  } else {
    for (uint32_t k = 0; k < S.nh; ++k) {
      const float x = fmaf(S.hv[k][2], wz, fmaf(S.hv[k][1], wy, S.hv[k][0] * wx));
      mx = fmaxf(mx, x);
      mn = fminf(mn, x);
      if (IN) {
        const float p = off + x, q = off - x;
        keep |= (uint64_t)(p > a && p < b) << k;
        keep |= (uint64_t)(q > a && q < b) << (32 + k);
      }
    }
  }
  const bool sym = NH > 0 ? SYM : S.sym != 0;
  if (sym) {
    mx = fmaxf(mx, -mn);
    mn = -mx;
  }
  *lo = mn;
  *hi = mx;
  return keep;
}

// 0: apart, 1: overlap, 2: undecided (XenoCollide in FP64).  Lanes with act
// false take part in the wave-uniform stage skips only.  NH, NF, NE > 0: the
// counts of S fixed at compile time (0: read from S).
template <int NH, int NF, int NE, bool SYM>
__device__ __forceinline__ uint32_t sat_classify(const SatShape& S, bool act, const float* v, const float* R) {
  const uint32_t nf = NF > 0 ? (uint32_t)NF : S.nf, ne = NE > 0 ? (uint32_t)NE : S.ne;
  const bool sym = NH > 0 ? SYM : S.sym != 0;
  const uint32_t nh = NH > 0 ? (uint32_t)NH : S.nh;
  const float tol = S.tol;
  bool sep = false, deep = true;
  uint64_t bin = ~0ull, ain = ~0ull;  // B's vertices inside A, A's vertices inside B
  for (uint32_t f = 0; f < nf; ++f) {
    const float nx = S.face[f][0], ny = S.face[f][1], nz = S.face[f][2];
    const float hmax = S.face[f][3], hmin = S.fmin[f];
    // A's face axis n: B's extent along it from R^T n
    {
      const float s = fmaf(nz, v[2], fmaf(ny, v[1], nx * v[0]));
      const float bx = fmaf(R[6], nz, fmaf(R[3], ny, R[0] * nx));
      const float by = fmaf(R[7], nz, fmaf(R[4], ny, R[1] * nx));
      const float bz = fmaf(R[8], nz, fmaf(R[5], ny, R[2] * nx));
      float lo, hi;
      bin &= sat_scan<NH, SYM, true>(S, bx, by, bz, s, hmin + tol, hmax - tol, &lo, &hi);
      const float gap = fmaxf((s + lo) - hmax, hmin - (s + hi));
      sep |= gap > tol;
      deep &= -gap > tol;
    }
    // B's face axis w = R n: A's extent along it
    {
      const float wx = fmaf(R[2], nz, fmaf(R[1], ny, R[0] * nx));
      const float wy = fmaf(R[5], nz, fmaf(R[4], ny, R[3] * nx));
      const float wz = fmaf(R[8], nz, fmaf(R[7], ny, R[6] * nx));
      const float s = fmaf(wz, v[2], fmaf(wy, v[1], wx * v[0]));
      float lo, hi;
      ain &= sat_scan<NH, SYM, true>(S, wx, wy, wz, 0.0f, s + hmin + tol, s + hmax - tol, &lo, &hi);
      const float gap = fmaxf((s + hmin) - hi, lo - (s + hmax));
      sep |= gap > tol;
      deep &= -gap > tol;
    }
  }
  const uint64_t valid = (nh >= 32 ? 0xffffffffull : ((1ull << nh) - 1ull)) * (sym ? 0x100000001ull : 1ull);
  uint32_t r = sep ? 0u : (((bin | ain) & valid) ? 1u : 2u);
  if (nf == 0) r = 2u;
  if (!S.use_edges || __ballot(act && r == 2u) == 0ull) return r;
  // the edge-pair axes w = e_a x R e_b; B's extent along it from
  // R^T w = (R^T e_a) x e_b
  for (uint32_t ia = 0; ia < ne; ++ia) {
    const float ax = S.edge[ia][0], ay = S.edge[ia][1], az = S.edge[ia][2];
    const float tx = fmaf(R[6], az, fmaf(R[3], ay, R[0] * ax));
    const float ty = fmaf(R[7], az, fmaf(R[4], ay, R[1] * ax));
    const float tz = fmaf(R[8], az, fmaf(R[5], ay, R[2] * ax));
    for (uint32_t ib = 0; ib < ne; ++ib) {
      const float ex = S.edge[ib][0], ey = S.edge[ib][1], ez = S.edge[ib][2];
      const float px = fmaf(R[2], ez, fmaf(R[1], ey, R[0] * ex));
      const float py = fmaf(R[5], ez, fmaf(R[4], ey, R[3] * ex));
      const float pz = fmaf(R[8], ez, fmaf(R[7], ey, R[6] * ex));
      const float wx = ay * pz - az * py, wy = az * px - ax * pz, wz = ax * py - ay * px;
      const float ux = ty * ez - tz * ey, uy = tz * ex - tx * ez, uz = tx * ey - ty * ex;
      const float s = fmaf(wz, v[2], fmaf(wy, v[1], wx * v[0]));
      float alo, ahi, blo, bhi;
      (void)sat_scan<NH, SYM, false>(S, wx, wy, wz, 0.0f, 0.0f, 0.0f, &alo, &ahi);
      (void)sat_scan<NH, SYM, false>(S, ux, uy, uz, 0.0f, 0.0f, 0.0f, &blo, &bhi);
      const float gap = fmaxf((s + blo) - ahi, alo - (s + bhi));
      sep |= gap > tol;
      deep &= -gap > tol;
    }
  }
  if (r == 2u) r = sep ? 0u : (deep ? 1u : 2u);
  return r;
}

// B's centre and rotation in A's frame, in FP32, from the FP64 site data.
__device__ __forceinline__ void sat_frame(const double* ti, const double* qi, const v3& rj, const double4& qj,
                                          float* v, float* R) {
  const float dx = (float)(rj.x - ti[0]), dy = (float)(rj.y - ti[1]), dz = (float)(rj.z - ti[2]);
  const float a = (float)qi[0], b = (float)qi[1], c = (float)qi[2], d = (float)qi[3];
  // rows of rotmat(q_i) (quaternion.rs:665-690); v = rotmat(q_i)^T (r_j - r_i)
  const float m0 = a * a + b * b - c * c - d * d, m1 = 2.0f * (b * c - a * d), m2 = 2.0f * (b * d + a * c);
  const float m3 = 2.0f * (b * c + a * d), m4 = a * a - b * b + c * c - d * d, m5 = 2.0f * (c * d - a * b);
  const float m6 = 2.0f * (b * d - a * c), m7 = 2.0f * (c * d + a * b), m8 = a * a - b * b - c * c + d * d;
  v[0] = fmaf(m6, dz, fmaf(m3, dy, m0 * dx));
  v[1] = fmaf(m7, dz, fmaf(m4, dy, m1 * dx));
  v[2] = fmaf(m8, dz, fmaf(m5, dy, m2 * dx));
  // o = conj(q_i) q_j (quaternion.rs:405-418 with the conjugate)
  const float js = (float)qj.x, jx = (float)qj.y, jy = (float)qj.z, jz = (float)qj.w;
  const float os = a * js + b * jx + c * jy + d * jz;
  const float ox = a * jx - js * b + (-c * jz + d * jy);
  const float oy = a * jy - js * c + (-d * jx + b * jz);
  const float oz = a * jz - js * d + (-b * jy + c * jx);
  R[0] = os * os + ox * ox - oy * oy - oz * oz;
  R[1] = 2.0f * (ox * oy - os * oz);
  R[2] = 2.0f * (ox * oz + os * oy);
  R[3] = 2.0f * (ox * oy + os * oz);
  R[4] = os * os - ox * ox + oy * oy - oz * oz;
  R[5] = 2.0f * (oy * oz - os * ox);
  R[6] = 2.0f * (ox * oz - os * oy);
  R[7] = 2.0f * (oy * oz + os * ox);
  R[8] = os * os - ox * ox - oy * oy + oz * oz;
}

__device__ __forceinline__ uint32_t sat_dispatch(const SatShape& S, bool act, const float* v, const float* R) {
  // unit cubes and other centrally symmetric hexahedra: 4 scanned vertices,
  // 3 face axes, 3 edge directions
  if (S.sym && S.nh == 4 && S.nf == 3 && S.ne == 3) return sat_classify<4, 3, 3, true>(S, act, v, R);
  return sat_classify<0, 0, 0, false>(S, act, v, R);
}

// Convex pass layout: kCvTrials trials per wavefront, one per group of
// kCvLanes lanes (the group draws its trial's proposal and lists its
// neighbour spaces); the candidate sites of all the wavefront's trials are
// tested against the circumsphere cutoff 64 at a time (kCvBatch rounds with
// their loads in flight together), the pairs within it are gathered in a
// per-wavefront list, and XenoCollide runs over that list 64 pairs at a time.
// Any overlap rejects a hard-shape trial, so the order of the pair tests does
// not change a decision (parallel_sweep.rs:267-273).
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

// Mixed precision: the FP32 separating-axis tests over the np listed pairs.
// Certain overlaps set their trial's bit in *hits; the undecided pairs are
// moved to the front of the list (in order) for XenoCollide in FP64, and
// their number is returned.
__device__ __forceinline__ uint32_t cv_sat_filter(const Geometry& g, const Cells& C, CvWave& W, const SatShape& S,
                                                  uint32_t np, uint32_t* hits) {
  const uint32_t lane = __lane_id();
  uint32_t nb = 0;
  for (uint32_t p0 = 0; p0 < np; p0 += 64) {
    const uint32_t p = p0 + lane;
    bool act = false;
    uint32_t kp = 0, qs = 0;
    float v[3] = {0.0f, 0.0f, 0.0f}, R[9] = {1.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 1.0f};
    if (p < np) {
      kp = W.tk[p];
      if (!((*hits >> kp) & 1u)) {
        act = true;
        qs = W.qs[p];
        const double4 sj = C.pos64[qs];
        const double4 qj4 = C.quat64[qs];
        const double tq[3] = {W.t[kp][0], W.t[kp][1], W.t[kp][2]};
        v3 rj;
        (void)convex_image(g, tq, sj, &rj);  // the image found when the pair was listed
        sat_frame(tq, W.qt[kp], rj, qj4, v, R);
      }
    }
    const uint32_t r = sat_dispatch(S, act, v, R);
#pragma unroll
    for (uint32_t i = 0; i < kCvTrials; ++i)
      if (__ballot(act && r == 1u && kp == i)) *hits |= 1u << i;
    const bool band = act && r == 2u;
    const unsigned long long mk = __ballot(band);
    if (band) {
      const uint32_t o = nb + (uint32_t)__popcll(mk & ((1ull << lane) - 1ull));
      W.qs[o] = qs;  // o <= p: entries not yet read are never overwritten
      W.tk[o] = kp;
    }
    nb += (uint32_t)__popcll(mk);
  }
  return nb;
}

// One colour of one cycle for convex polyhedra: kCvTrials spaces per
// wavefront.  Each trial is the one of trial_convex (parallel_sweep.rs:234-277
// with Translate or Rotate<Versor>).
// (three wavefronts per SIMD: 22% faster on C3 than the two its free
// register use allows)
// MIXED: the listed pairs go through the FP32 separating-axis tests first
// (cv_sat_filter); only the undecided ones run XenoCollide in FP64.
template <bool MIXED>
__global__ void __launch_bounds__(kBlock, 3) k_pass_convex(Geometry g, ApplyState* st, Bodies B, Cells C, Work w,
                                                        uint32_t colour, uint32_t cycle, uint32_t xcd_chunk,
                                                        SatShape sat) {
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
  uint32_t nband = 0; // MIXED: pairs left to FP64 XenoCollide (uniform)
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
      if (MIXED && sat.use) {
        np = cv_sat_filter(g, C, W, sat, np, &hits);
        nband += np;
        wave_lds_order();
      }
      // (the pairs left by the FP32 tests run the inlined XenoCollide too: a
      // pass lasts as long as its slowest wavefront, and the out-of-line one
      // made those wavefronts slower)
      if (np == 0) {
      } else if (g.nverts == 8)
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
  if (MIXED && lane == 0 && nband) atomicAdd(&st->band_retests, (unsigned long long)nband);
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
      B.bpos[body] = make_double4(nb3[0], nb3[1], nb3[2], move_stamp(st->epoch));
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

// tune_local_trial (hoomd-mc/src/tune_local.rs:21-89) for hard convex
// polyhedra with Translate (move 0) or Rotate<Versor> (move 1): one sample
// from every offset p of the pre-drawn stream xs[0, M), as k_tune_eval does
// for point bodies.  The trial site is tested against the sites of every
// other body within the circumsphere cutoff, from the cells of the current
// state around the trial site's space; hard shapes accept without a draw.
// next[p]: the offset after the sample (UINT32_MAX: the stream ran out);
// flags[p]: bit 0 accepted, bit 2 a site outside the checkerboard.
__global__ void __launch_bounds__(256) k_tune_eval_convex(Geometry g, const ApplyState* st, Bodies B, Work w,
                                                          const uint64_t* __restrict__ xs, uint32_t M, uint32_t move,
                                                          double size, uint32_t* __restrict__ next,
                                                          uint8_t* __restrict__ flags) {
  __shared__ double zx[257], zf[257];
  __shared__ double vx[3 * kMaxVertices];
  stage_zig(zx, zf);
  stage_verts(g, vx);
  __syncthreads();
  const uint32_t p = blockIdx.x * 256 + threadIdx.x;
  if (p >= M) return;
  StreamRng r{xs, p, M, false};
  const uint32_t body = choose_index(r, g.n_bodies);
  const double4 bp = B.bpos[body];
  const double4 bq = B.bquat[body];
  double x[3] = {bp.x, bp.y, bp.z};
  quat q = {bq.x, {bq.y, bq.z, bq.w}};
  if (move == 0) {
    double v[3] = {0.0, 0.0, 0.0};
    ball_sample<3>(r, size, zx, zf, v);
    for (int a = 0; a < 3; ++a) x[a] += v[a];
  } else {
    q = qmul(versor_displacement(r, size, zx, zf), q);  // rotate/versor.rs:111-122
  }
  for (int a = 0; a < 3; ++a) x[a] = wrap1(x[a], g.L[a]);  // wrap(propose(..))
  // the trial site: wrap(x + rotate(q, 0)), orientation q * identity
  const v3 r0 = qrotate(q, v3{0.0, 0.0, 0.0});
  const double t[3] = {wrap1(x[0] + r0.x, g.L[0]), wrap1(x[1] + r0.y, g.L[1]), wrap1(x[2] + r0.z, g.L[2])};
  const quat qt = qmul(q, quat{1.0, {0.0, 0.0, 0.0}});
  const v3 tv = {t[0], t[1], t[2]};
  bool hit = false, outside = false;
  const uint32_t s = space_of<3>(g, st->origin, t);
  if (s == UINT32_MAX) {
    outside = true;
  } else {
    int32_t m[3];
    for (int a = 0; a < 3; ++a) m[a] = (int32_t)((s / g.stride[a]) % (uint32_t)g.shape[a]);
    uint32_t nb[3][5];
    int nn[3] = {1, 1, 1};
    for (int a = 0; a < 3; ++a) {
      const int32_t sh = g.shape[a], rr = g.reach[a];
      int k = 0;
      for (int off = -rr; off <= rr; ++off) {
        const int32_t mm = ((m[a] + off) % sh + sh) % sh;
        bool seen = false;
        for (int j = 0; j < k; ++j) seen |= nb[a][j] == (uint32_t)mm * g.stride[a];
        if (!seen) nb[a][k++] = (uint32_t)mm * g.stride[a];
      }
      nn[a] = k;
    }
    for (int i0 = 0; i0 < nn[0] && !hit; ++i0)
      for (int i1 = 0; i1 < nn[1] && !hit; ++i1)
        for (int i2 = 0; i2 < nn[2] && !hit; ++i2) {
          const uint32_t ns = nb[0][i0] + nb[1][i1] + nb[2][i2];
          const uint32_t q0 = w.start[ns], q1 = w.start[ns + 1];
          for (uint32_t qq = q0; qq < q1 && !hit; ++qq) {
            const uint32_t bj = w.tmp_body[qq];
            if (bj == body) continue;
            v3 rj;
            const double d2 = convex_image(g, t, B.spos[bj], &rj);
            if (d2 < g.cv_rcut2) {
              const double4 qj4 = B.squat[bj];
              hit = convex_overlap(vx, g.nverts, g.cv_rcut2, d2, tv, qt, rj, quat{qj4.x, {qj4.y, qj4.z, qj4.w}});
            }
          }
        }
  }
  next[p] = r.over ? 0xFFFFFFFFu : r.pos;
  flags[p] = (uint8_t)((hit ? 0u : 1u) | (outside ? 4u : 0u));
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
                               double* out, uint64_t n, SatShape sat) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (kind == 3) {
    // the FP32 separating-axis class of a pair (in as for kind 1): 0 apart,
    // 1 overlap, 2 undecided; every lane of the wavefront takes part
    const bool act = i < n;
    float v[3] = {0.0f, 0.0f, 0.0f}, R[9] = {1.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 1.0f};
    if (act) {
      const double* p = in + 14 * i;
      sat_frame(p, p + 3, v3{p[7], p[8], p[9]}, make_double4(p[10], p[11], p[12], p[13]), v, R);
    }
    const uint32_t r = sat.use ? sat_dispatch(sat, act, v, R) : 2u;
    if (act) out[i] = (double)r;
    return;
  }
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

// explicit instantiations (launched from the host translation units)
template __global__ void k_pass_convex<false>(Geometry, ApplyState*, Bodies, Cells, Work, uint32_t, uint32_t, uint32_t,
                                              SatShape);
template __global__ void k_pass_convex<true>(Geometry, ApplyState*, Bodies, Cells, Work, uint32_t, uint32_t, uint32_t,
                                             SatShape);

}  // namespace hmc
