// The persistent apply for isotropic pair energies (LJ and shifted LJ, mixed
// precision): one cooperative launch runs every cycle and colour of an apply,
// as k_apply_tile does for hard shapes.  A trial runs on one half-wavefront
// (32 lanes); half-wave h of G takes trials h, h + G, ... of its region.
// Workgroups have 1, 2 or 4 wavefronts, so that the wavefronts of the
// workgroups sharing a CU spread evenly over its four SIMDs.
//
// A trial's candidates are the staged sites of its 3^D neighbour cells in
// LDS (box-relative FP32, body index in w).  For each candidate the final
// and initial FP32 displacements give r^2; membership r^2 < r_cut^2 is
// decided in FP32 only outside a band around r_cut^2 that bounds the FP32
// error of r^2 against the reference's FP64 value, and the pair energies are
// FP32 (lj_energy32f).  The decision stands when the FP32 dH, with a
// rigorous bound on its distance to the reference's FP64 dH, decides the
// Metropolis test; otherwise (a candidate in the membership band, a site
// without images, or a bound that straddles the decision) the trial is
// re-taken exactly in FP64 from the FP64 sites with the reference's image
// rules (lj_r2), with the same random draws.  Decisions therefore equal the
// per-colour FP64 path's bit for bit.
//
// Bound of one FP32 pair term against the reference's FP64 term, for the
// per-axis displacement error E (tc.perr: the roundings of the FP64 sites
// to box-relative FP32 on both sides, and the reference's own roundings of
// the displacement and image shift), with u = 2^-24:
//   |r2_32 - r2_ref| <= beta = E (2 sum|d_a| + D E) + (5 u + 2^-50) r2_32
//     (the FP32 difference, squares and sums; the reference's r^2),
//   |e_32 - e(r2_32)| <= 48 u A   (lj_energy32, A = 4 eps (s12 + s6) + |shift|),
//   |e(r2_32) - e(r2_ref)| <= |de/dr2| beta, |de/dr2| = 4 eps |6 s12 - 3 s6| / r2
//     taken at r2_32 with a relative margin (2^-16 of 4 eps (6 s12 + 3 s6) / r2)
//     and a factor 1.02 for the interval (beta <= r2 / 1024, else the pair is
//     left open).
// The sums over a half-wave add gamma_(rounds + 5) (A_f + A_i); the
// reference's FP64 terms and sum differ from exact by less than
// 2^-40 (A_f + A_i).  A trial the FP32 values leave open is re-taken in FP64
// over the candidates whose FP32 r^2 (final or initial) lies below the top
// of the membership band -- every term the reference sums is among them.
#include "persist.hpp"

namespace hmc {

constexpr int kLjRegionThreads = 256;  // upper bound of the workgroup size: 8 half-waves
constexpr int kLjMaxTrials = 12;       // trials of a region per pass

// The proposal of a half-wave's next trial, kept in LDS across the wait for
// the neighbour regions (in registers it would push the kernel past its
// register budget).  Written by lane 0 of the half-wave.
struct LjProp {
  double v[3];   // displacement (Ball::sample)
  double bp[3];  // the chosen body's position
  Sfc64 rng;     // the trial's stream after the proposal (Metropolis draws)
  uint32_t s, e0, pick, body, go, pad;
};

static_assert(sizeof(LjProp) == 104, "proposal size counted in the host's LDS budget");

struct Lj32 {
  float e, a, b;  // sum of FP32 energies, of their magnitudes A, of the position-error bounds
};

// One candidate q (box-relative FP32 entry) against the FP32 site p: adds
// its term to acc when within r_cut; sets *open when the membership or the
// bound cannot be decided in FP32.
template <int D>
__device__ __forceinline__ void lj_pair32(const float* p, float4 q, float lo, float hi, float perr, float sig2,
                                          float eps4, float shift, Lj32& acc, bool& open) {
  const float dx = q.x - p[0], dy = q.y - p[1];
  float r2, sabs;
  if (D == 3) {
    const float dz = q.z - p[2];
    r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    sabs = fabsf(dx) + fabsf(dy) + fabsf(dz);
  } else {
    r2 = fmaf(dx, dx, dy * dy);
    sabs = fabsf(dx) + fabsf(dy);
  }
  const bool in = r2 < lo, out = r2 >= hi;
  if (!(in || out)) open = true;  // also NaN: a site without images
  if (in) {
    const float beta = perr * (2.0f * sabs + (float)D * perr) + r2 * (5.0f * 0x1p-24f + 0x1p-50f);
    if (beta * 1024.0f > r2) open = true;
    const float rr = __builtin_amdgcn_rcpf(r2);
    const float s2 = sig2 * rr;
    const float s4 = s2 * s2, s8 = s4 * s4;
    const float s12 = s4 * s8, s6 = s2 * s4;
    acc.e += eps4 * (s12 - s6) - shift;
    acc.a += eps4 * (s12 + s6) + fabsf(shift);
    const float dd = eps4 * (fabsf(6.0f * s12 - 3.0f * s6) + 0x1p-16f * (6.0f * s12 + 3.0f * s6)) * rr;
    acc.b += 1.02f * dd * beta;
  }
}

__device__ __forceinline__ float half_sum(float v) {
#pragma unroll
  for (int d = 16; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}
__device__ __forceinline__ double half_sum(double v) {
#pragma unroll
  for (int d = 16; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}
__device__ __forceinline__ uint32_t half_sum(uint32_t v) {
#pragma unroll
  for (int d = 16; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}
__device__ __forceinline__ bool half_any(bool v) {
  const unsigned long long m = __ballot(v);
  return ((m >> (__lane_id() & 32u)) & 0xffffffffull) != 0;
}

// FP64 sums of one trial over candidate bodies, exactly as the per-colour
// FP64 path forms each term (lj_r2: the reference's image and cell-edge
// rules), from the FP64 sites read with device-scope loads.  Lane hl of the
// half-wave takes candidates hl, hl + 32, ...; body(c) gives candidate c's
// body (or UINT32_MAX to skip it).
struct Lj64 {
  double ef, af, ei, ai;
  uint32_t nf, ni;
};
template <int D, class BodyOf>
__device__ __forceinline__ Lj64 lj_sums64(const Geometry& g, const double4* spos, const double* t, const double* xi,
                                          uint32_t total, BodyOf body_of) {
  const uint32_t hl = __lane_id() & 31u;
  Lj64 s{0.0, 0.0, 0.0, 0.0, 0u, 0u};
  for (uint32_t c = hl; c < total; c += 32u) {
    const uint32_t bd = body_of(c);
    if (bd == UINT32_MAX) continue;
    const double4 sj = ld_dev_d4(&spos[bd]);
    const double r2f = lj_r2<D>(g, t, sj);
    if (r2f < g.lj_rc2) {
      const double e = lj_energy(g, r2f);
      s.ef += e;
      s.af += fabs(e);
      s.nf += 1;
    }
    const double r2i = lj_r2<D>(g, xi, sj);
    if (r2i < g.lj_rc2) {
      const double e = lj_energy(g, r2i);
      s.ei += e;
      s.ai += fabs(e);
      s.ni += 1;
    }
  }
  s.ef = half_sum(s.ef);
  s.af = half_sum(s.af);
  s.ei = half_sum(s.ei);
  s.ai = half_sum(s.ai);
  s.nf = half_sum(s.nf);
  s.ni = half_sum(s.ni);
  return s;
}

// The FP64 decision of the per-colour path (trial_lj): infinite final energy
// rejects, infinite initial energy accepts, else Metropolis with the
// summation-order bound (*amb when it cannot settle the decision).
__device__ __forceinline__ bool lj_decide64(Sfc64& rng, const Lj64& s, double kT, bool* amb) {
  *amb = false;
  if (s.ef == INFINITY) return false;
  if (s.ei == INFINITY) return true;
  const double dH = s.ef - s.ei;
  const double nmax = (double)max(max(s.nf, s.ni), 1u);
  const double gam = nmax * 0x1p-53 / (1.0 - nmax * 0x1p-53);
  const double delta = (2.0 * gam * (s.af + s.ai) + 0x1p-52 * fabs(dH)) * (1.0 + 0x1p-20) + 0x1p-1070;
  return lj_metropolis(rng, dH, delta, kT, amb) != 0;
}

// The halo reload of region_reload with the entries, not the cells, shared
// out over the workgroup (an isotropic region's cells hold ~15 sites each):
// wave 0 lists the halo cells of `colour` and the prefix of their entry
// counts (at most kLjHalo cells, (T + 1)^D), then every thread reloads
// entries tid, tid + NT, ... from the FP64 sites with device-scope loads, all
// in flight together.  Ends with a barrier.
constexpr uint32_t kLjHalo = 128;
template <int D>
__device__ __forceinline__ void region_reload_flat(const Geometry& g, const TileCfg& tc, const Region& rg,
                                                   const Bodies& B, const RegionLds& l, uint32_t colour,
                                                   uint32_t* hlo, uint32_t* hpre) {
  constexpr int AL = D - 1;
  const uint32_t CW = tc.E[AL] + 1;
  const uint32_t tid = threadIdx.x, NT = blockDim.x;
  if (tid < 64) {
    uint32_t nc = 1, cnt[3] = {1, 1, 1};
    for (int a = 0; a < D; ++a) {
      cnt[a] = tc.T[a] + 1;
      nc *= cnt[a];
    }
    uint32_t base = 0;
    for (uint32_t c0 = 0; c0 < nc; c0 += 64) {
      const uint32_t ci = c0 + tid;
      uint32_t n = 0, ea = 0;
      if (ci < nc) {
        uint32_t r = ci, bc[3] = {0, 0, 0};
        bool own = true;
        for (int a = D - 1; a >= 0; --a) {
          const uint32_t j = r % cnt[a];
          r /= cnt[a];
          bc[a] = 2 * j + (((colour >> (D - 1 - a)) & 1u) ? 0u : 1u);
          const int32_t lc = (int32_t)bc[a] - 1;
          if (lc < 0 || lc >= rg.rl[a]) own = false;
        }
        const uint32_t row = D == 3 ? bc[0] * tc.E[1] + bc[1] : bc[0];
        ea = l.coff[row * CW + bc[AL]];
        n = own ? 0u : l.coff[row * CW + bc[AL] + 1] - ea;
      }
      const uint32_t incl = wave_incl_scan(n);
      if (ci < nc) {
        hlo[ci] = ea;
        hpre[ci] = base + incl - n;
      }
      base += __shfl(incl, 63, 64);
    }
    if (tid == 0) hpre[nc] = base;
  }
  __syncthreads();
  uint32_t nc = 1;
  for (int a = 0; a < D; ++a) nc *= tc.T[a] + 1;
  const uint32_t total = hpre[nc];
  for (uint32_t e0 = 0; e0 < total; e0 += 2 * NT) {
    uint32_t dst[2] = {0, 0}, bd[2] = {0, 0};
    double4 v[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint32_t e = e0 + (uint32_t)h * NT + tid;
      if (e < total) {
        uint32_t lo = 0, hi = nc;  // the cell: the last k with hpre[k] <= e
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (hpre[mid] <= e)
            lo = mid;
          else
            hi = mid;
        }
        dst[h] = hlo[lo] + (e - hpre[lo]);
        bd[h] = (uint32_t)__float_as_int(l.lb[dst[h]].w);
        v[h] = ld_dev_d4(&B.spos[bd[h]]);
      }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h)
      if (e0 + (uint32_t)h * NT + tid < total) l.lb[dst[h]] = stage_entry64<D>(g, rg.c64, v[h], bd[h]);
  }
  __syncthreads();
}

template <int D>
__global__ void __launch_bounds__(kLjRegionThreads, 2) k_apply_lj(Geometry g, ApplyState* st, Bodies B, Cells C, Work w,
                                                             TileCfg tc) {
  constexpr int AL = D - 1;  // the row axis (stride 1)
  constexpr uint32_t NCOL = 1u << D;
  constexpr int NR = D == 3 ? 9 : 3;  // rows of three cells around a trial's cell
  const uint32_t b = blockIdx.x;
  const uint32_t tile = (b & 7u) * tc.xcd_chunk + (b >> 3);  // XCD-aware region order
  if (tile >= tc.ntiles) return;  // padding workgroups (grid rounded up to 8) take part in nothing
  if (tc.probe) {
    region_probe(st, w.mbox, tc.ntiles);
    return;
  }
  extern __shared__ __align__(16) unsigned char smem[];
  const RegionLds l = region_lds<D>(smem, tc);
  const uint32_t CW = tc.E[AL] + 1;
  const uint32_t* coff = l.coff;
  float4* lb = l.lb;
  __shared__ double zx[257], zf[257];
  __shared__ unsigned blk_acc;
  __shared__ uint32_t hlo[kLjHalo], hpre[kLjHalo + 1];
  __shared__ LjProp prop[kLjMaxTrials];
  __shared__ uint32_t radj[kLjRegionThreads / 32][NR];
  const uint32_t tid = threadIdx.x;
  const uint32_t wid = tid >> 6, lane = tid & 63u;
  const uint32_t hw = tid >> 5, hl = tid & 31u, G = blockDim.x >> 5;  // trials hw, hw + G, ...
  const uint32_t cycles = st->cycles, substep0 = st->substep0;
  const uint64_t step = st->step;
  const uint32_t seed = st->seed;
  const double max_move = st->max_move, kT = st->kT;
  const double stamp = move_stamp(st->epoch);
  double origin[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < D; ++a) origin[a] = st->origin[a];
  const Region rg = region_of<D>(g, tc, origin, tile);
  const float sig2 = (float)(g.lj_sigma * g.lj_sigma), eps4 = (float)g.lj_eps4, shift = (float)g.lj_shift;
  const float blo = tc.band_lo, bhi = tc.band_hi, perr = tc.perr;
  const uint32_t ntr = tc.trials;
  const uint32_t nnb = D == 3 ? 26u : 8u;
  const uint32_t nbt = wid == 0 && lane < nnb ? region_neighbour<D>(tc, rg, lane) : tile;

  if (D == 3) stage_zig(zx, zf);
  if (tid == 0) blk_acc = 0;
  const bool overflow = region_stage<D>(g, tc, rg, B, C, w, l);
  if (tid == 0) w.mbox[tile * kMailboxStride + 3] = 1u;  // progress: staged

  unsigned retests = 0, ambiguous = 0;
  // The proposals of pass i, drawn while the neighbour regions finish pass
  // i - 1 (they read only static cell data and this region's own bodies);
  // trial k's waits in prop[k].  Trial k of colour c is the space with digits
  // 2 h + c_a of the region (h the digits of k in T).
  auto trial_space = [&](uint32_t colour, uint32_t k, uint32_t* cb, bool* active) {
    uint32_t s = 0, hh[3] = {0, 0, 0};
    for (int a = D - 1; a >= 0; --a) {
      hh[a] = k % tc.T[a];
      k /= tc.T[a];
    }
    *active = true;
    for (int a = 0; a < D; ++a) {
      const uint32_t lc = 2 * hh[a] + ((colour >> (D - 1 - a)) & 1u);
      if ((int32_t)lc >= rg.rl[a]) *active = false;
      s += (uint32_t)(rg.org[a] + (int32_t)lc) * g.stride[a];
      cb[a] = lc + 1;
    }
    return s;
  };
  auto propose = [&](uint32_t i, uint32_t k) {
    const uint32_t colour = i & (NCOL - 1), substep = substep0 + (i >> D);
    uint32_t cb[3] = {0, 0, 0};
    bool active = false;
    const uint32_t s = trial_space(colour, k, cb, &active);
    uint32_t go = 0, e0 = 0, pick = 0, body = 0;
    double v[3] = {0.0, 0.0, 0.0};
    double4 bp = make_double4(0.0, 0.0, 0.0, 0.0);
    Sfc64 rng{0, 0, 0, 0};
    if (active && !overflow) {
      const uint32_t crow = D == 3 ? cb[0] * tc.E[1] + cb[1] : cb[0];
      e0 = coff[crow * CW + cb[AL]];
      const uint32_t n = coff[crow * CW + cb[AL] + 1] - e0;
      if (n) {
        go = 1;
        rng = counter_rng(step, substep, seed, s);
        pick = choose_index(rng, n);
        body = (uint32_t)__float_as_int(lb[e0 + pick].w);
        bp = B.bpos[body];
        ball_sample<D>(rng, max_move, zx, zf, v);
      }
    }
    if (hl == 0) {
      LjProp& pr = prop[k];
      pr.s = s;
      pr.go = go;
      pr.e0 = e0;
      pr.pick = pick;
      pr.body = body;
      pr.rng = rng;
      for (int a = 0; a < 3; ++a) pr.v[a] = v[a];
      pr.bp[0] = bp.x;
      pr.bp[1] = bp.y;
      pr.bp[2] = bp.z;
    }
  };

  const uint32_t npass = cycles * NCOL;
  if (npass)
    for (uint32_t k = hw; k < ntr; k += G) propose(0, k);
  wave_lds_order();
  unsigned nacc = 0;  // accepted trials (lane 0 of each half-wave)
  for (uint32_t i = 0; i < npass; ++i) {
    const uint32_t colour = i & (NCOL - 1), substep = substep0 + (i >> D);
    for (uint32_t k = hw; k < ntr; k += G) {
      const LjProp& pr = prop[k];
      bool accepted = false, traced = false;
      uint32_t tbody = 0;
      if (overflow) {
        // the region's box exceeds its LDS capacity (rare): the trial from the
        // cell lists in global memory, exactly in FP64
        bool active = true;
        uint32_t s = 0;
        int32_t m[3] = {0, 0, 0};
        {
          uint32_t cb[3] = {0, 0, 0};
          s = trial_space(colour, k, cb, &active);
          for (int a = 0; a < D; ++a) m[a] = rg.org[a] + (int32_t)cb[a] - 1;
        }
        const uint32_t s0 = active ? w.start[s] : 0u, n = active ? w.start[s + 1] - s0 : 0u;
        if (active && n) {
          Sfc64 rng = counter_rng(step, substep, seed, s);
          const uint32_t pick = choose_index(rng, n);
          const uint32_t slot = s0 + pick, body = C.body[slot];
          double v[3] = {0.0, 0.0, 0.0};
          ball_sample<D>(rng, max_move, zx, zf, v);
          const double4 bp = B.bpos[body];
          double x[3] = {bp.x, bp.y, bp.z};
          for (int a = 0; a < D; ++a) x[a] = wrap1(x[a] + v[a], g.L[a]);
          traced = true;
          tbody = body;
          if (space_of<D>(g, origin, x) == s) {
            double t[3] = {0.0, 0.0, 0.0};
            for (int a = 0; a < D; ++a) t[a] = wrap1(x[a], g.L[a]);
            const double4 sp = ld_dev_d4(&B.spos[body]);
            const double xi[3] = {sp.x, sp.y, sp.z};
            // candidates: the slots of the 3^D neighbour spaces in order, their
            // first slots and prefix in this half-wave's share of the LDS site
            // buffer (unused when the region overflows)
            constexpr uint32_t NS = D == 3 ? 27u : 9u;
            uint32_t* seg0 = (uint32_t*)lb + hw * 64u;
            uint32_t* segp = seg0 + 32u;
            if (hl < NS) {
              uint32_t q = hl, ns = 0;
              for (int a = D - 1; a >= 0; --a) {
                const int32_t sh = g.shape[a];
                int32_t mm = m[a] + (int32_t)(q % 3u) - 1;
                q /= 3u;
                mm = mm < 0 ? mm + sh : (mm >= sh ? mm - sh : mm);
                ns += (uint32_t)mm * g.stride[a];
              }
              seg0[hl] = w.start[ns];
              segp[hl + 1] = w.start[ns + 1] - w.start[ns];
            }
            if (hl == 0) segp[0] = 0;
            wave_lds_order();
            if (hl == 0)
              for (uint32_t k = 1; k <= NS; ++k) segp[k] += segp[k - 1];
            wave_lds_order();
            const uint32_t total = segp[NS];
            const Lj64 s64 = lj_sums64<D>(g, B.spos, t, xi, total, [&](uint32_t c) -> uint32_t {
              uint32_t lo = 0, hi = NS;  // the last k with segp[k] <= c
              while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (segp[mid] <= c)
                  lo = mid;
                else
                  hi = mid;
              }
              const uint32_t q = seg0[lo] + (c - segp[lo]);
              return q == slot ? UINT32_MAX : C.body[q];
            });
            bool amb = false;
            accepted = lj_decide64(rng, s64, kT, &amb);
            ambiguous += (amb && hl == 0) ? 1u : 0u;
            if (accepted && hl == 0) {
              double nsp[3] = {0.0, 0.0, 0.0};
              for (int a = 0; a < D; ++a) nsp[a] = wrap1(t[a] + 0.0, g.L[a]);
              B.bpos[body] = make_double4(t[0], t[1], t[2], stamp);
              st_dev_d4(&B.spos[body], pack_site(nsp, ghost_flags<D>(g, nsp)));
            }
          }
        }
      } else if (pr.go) {
        // ---- the trial of this pass from LDS
        const uint32_t ps = pr.s, self = pr.e0 + pr.pick, body = pr.body;
        uint32_t pcb[3] = {0, 0, 0};
        bool act = false;
        (void)trial_space(colour, k, pcb, &act);
        traced = true;
        tbody = body;
        double x[3] = {pr.bp[0], pr.bp[1], pr.bp[2]};
        for (int a = 0; a < D; ++a) x[a] += pr.v[a];
        for (int a = 0; a < D; ++a) x[a] = wrap1(x[a], g.L[a]);  // generate_trial_moves: wrap(propose(..))
        if (space_of<D>(g, origin, x) == ps) {
          double t[3] = {0.0, 0.0, 0.0};  // wrap(x + 0) == wrap(x)
          for (int a = 0; a < D; ++a) t[a] = wrap1(x[a], g.L[a]);
          float tr[3] = {0.0f, 0.0f, 0.0f}, xr[3] = {0.0f, 0.0f, 0.0f};
          {
            const float4 te = stage_entry64<D>(g, rg.c64, make_double4(t[0], t[1], D == 3 ? t[2] : 0.0, 0.0), 0u);
            tr[0] = te.x;
            tr[1] = te.y;
            tr[2] = te.z;
            const float4 se = lb[self];
            xr[0] = se.x;
            xr[1] = se.y;
            xr[2] = se.z;
          }
          // rows of three cells around the trial's cell: the prefix of their
          // lengths in registers, and per row the offset from a candidate's
          // index to its LDS entry in this half-wave's table (row r* of
          // candidate c is the number of row starts at or below c; indexing a
          // register array with it would put the array in scratch memory)
          uint32_t pre[NR + 1];
          pre[0] = 0;
          uint32_t* adj = radj[hw];
  #pragma unroll
          for (int r = 0; r < NR; ++r) {
            const int i0 = D == 3 ? r / 3 : r;
            const int i1 = D == 3 ? r % 3 : 0;
            const uint32_t row = D == 3 ? (pcb[0] + i0 - 1) * tc.E[1] + (pcb[1] + i1 - 1) : pcb[0] + i0 - 1;
            const uint32_t lo_r = coff[row * CW + pcb[AL] - 1];
            pre[r + 1] = pre[r] + (coff[row * CW + pcb[AL] + 2] - lo_r);
            if (hl == (uint32_t)r) adj[r] = lo_r - pre[r];
          }
          wave_lds_order();
          const uint32_t total = pre[NR];
          auto slot_of = [&](uint32_t c) -> uint32_t {
            uint32_t r = 0;
  #pragma unroll
            for (int k = 1; k < NR; ++k) r += c >= pre[k] ? 1u : 0u;
            return c + adj[r];
          };
          Lj32 af{0.0f, 0.0f, 0.0f}, ai{0.0f, 0.0f, 0.0f};
          bool open = false;
          for (uint32_t c = hl; c < total; c += 32u) {
            const uint32_t q = slot_of(c);
            if (q == self) continue;
            const float4 e = lb[q];
            lj_pair32<D>(tr, e, blo, bhi, perr, sig2, eps4, shift, af, open);
            lj_pair32<D>(xr, e, blo, bhi, perr, sig2, eps4, shift, ai, open);
          }
          open = half_any(open);
          bool decided = false;
          Sfc64 rng = pr.rng;
          if (!open) {
            const double ef = (double)half_sum(af.e), ei = (double)half_sum(ai.e);
            const double sa = (double)half_sum(af.a) + (double)half_sum(ai.a);
            const double sb = (double)half_sum(af.b) + (double)half_sum(ai.b);
            const double rounds = (double)((total + 31u) / 32u);
            const double dH32 = ef - ei;
            const double delta = (((48.0 + rounds + 8.0) * 0x1p-24 + 0x1p-40) * sa + sb) * (1.0 + 0x1p-10) + 0x1p-1070;
            if (isfinite(dH32) && isfinite(delta)) {
              bool amb32 = false;
              accepted = lj_metropolis(rng, dH32, delta, kT, &amb32) != 0;
              decided = !amb32;
            }
          }
          if (!decided) {
            // exact FP64 re-take with the same draws
            rng = pr.rng;
            retests += hl == 0 ? 1u : 0u;
            const double4 sp = ld_dev_d4(&B.spos[body]);
            const double xi[3] = {sp.x, sp.y, sp.z};
            const Lj64 s64 = lj_sums64<D>(g, B.spos, t, xi, total, [&](uint32_t c) -> uint32_t {
              const uint32_t q = slot_of(c);
              if (q == self) return UINT32_MAX;
              const float4 e = lb[q];
              float dfx = e.x - tr[0], dfy = e.y - tr[1], dix = e.x - xr[0], diy = e.y - xr[1];
              float r2f = fmaf(dfx, dfx, dfy * dfy), r2i = fmaf(dix, dix, diy * diy);
              if (D == 3) {
                const float dfz = e.z - tr[2], diz = e.z - xr[2];
                r2f = fmaf(dfz, dfz, r2f);
                r2i = fmaf(diz, diz, r2i);
              }
              // both far beyond r_cut in FP32: no term of the reference's sums
              return (r2f >= bhi && r2i >= bhi) ? UINT32_MAX : (uint32_t)__float_as_int(e.w);
            });
            bool amb = false;
            accepted = lj_decide64(rng, s64, kT, &amb);
            ambiguous += (amb && hl == 0) ? 1u : 0u;
          }
          if (accepted && hl == 0) {
            // update_body_properties (microstate.rs:985-1034)
            double nsp[3] = {0.0, 0.0, 0.0};
            for (int a = 0; a < D; ++a) nsp[a] = wrap1(t[a] + 0.0, g.L[a]);
            const double4 sp = pack_site(nsp, ghost_flags<D>(g, nsp));
            B.bpos[body] = make_double4(t[0], t[1], t[2], stamp);
            st_dev_d4(&B.spos[body], sp);
            lb[self] = stage_entry64<D>(g, rg.c64, sp, body);
          }
        }
      }
      if (traced && hl == 0 && w.tracing) {
        unsigned long long ti = atomicAdd((unsigned long long*)&st->trace_n, 1ull);
        if (ti < w.trace_cap) {
          uint32_t cbt[3] = {0, 0, 0};
          bool act = false;
          hmc_trace_record rec;
          rec.substep = substep;
          rec.color = colour;
          rec.space = trial_space(colour, k, cbt, &act);
          rec.body = (int64_t)tbody;
          rec.status = accepted ? 1 : 0;
          rec.pad = 0;
          w.trace[ti] = rec;
        }
      }
      nacc += (accepted && hl == 0) ? 1u : 0u;
    }
    if (i + 1 == npass) break;  // no later pass reads the halo
    region_publish(w.mbox, nbt, nnb, i & 1u);
    if (tid == 0) w.mbox[tile * kMailboxStride + 4] = i + 1;  // progress: passes published
    for (uint32_t k = hw; k < ntr; k += G) propose(i + 1, k);
    region_wait(st, w.mbox, tile, nnb, i);
    if (tid == 0) w.mbox[tile * kMailboxStride + 5] = i + 1;  // progress: waits passed
    if (!overflow)
      region_reload_flat<D>(g, tc, rg, B, l, colour, hlo, hpre);
    else
      __syncthreads();
  }
  lane_add(&st->band_retests, retests);
  lane_add(&st->ambiguous, ambiguous);
  if (tid == 0 && overflow) atomicAdd(&st->tile_overflows, 1ull);
  if (nacc) atomicAdd(&blk_acc, nacc);
  __syncthreads();
  if (tid == 0 && blk_acc) atomicAdd(&w.acc_slots[b & (kAccSlots - 1)], (unsigned long long)blk_acc);
}

template __global__ void k_apply_lj<2>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, TileCfg tc);
template __global__ void k_apply_lj<3>(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, TileCfg tc);

}  // namespace hmc
