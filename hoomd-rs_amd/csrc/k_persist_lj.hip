// The persistent apply for isotropic pair energies (LJ and shifted LJ, mixed
// precision): one cooperative launch runs every cycle and colour of an apply,
// as k_apply_tile does for hard shapes.  Trials 0 .. G - 1 of a region run
// on one half-wavefront (32 lanes) each, G the half-waves of the workgroup;
// a trial beyond them runs on the whole workgroup, its half-waves' sums
// combined in LDS.  Workgroups have 1, 2 or 4 wavefronts, so that the
// wavefronts of the workgroups sharing a CU spread evenly over its four
// SIMDs.
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
//     (the FP32 difference, squares and sums; the reference's r^2), and for
//     every pair counted (r2_32 < band_hi) beta <= tc.beta_max;
//   |e_32 - e(r2_32)| <= 48 u A   (A = 4 eps (s12 + s6) + |shift|; the FP32
//     powers s12 and s6 are summed separately and combined in FP64),
//   |e(r2_32) - e(r2_ref)| <= |de/dr2| beta <= 4 eps (6 s12 + 3 s6) / r2 x
//     beta_max x 1.02 (de/dr2 varies by less than 2% over the interval when
//     beta <= r2 / 1024, i.e. r2_32 >= tc.r2_open; closer pairs are left
//     open).  W = (6 s12 + 3 s6) / r2 is summed in FP32 and the bound taken
//     as 1.03 (1 + 2^-16) 4 eps beta_max sum W, the extra factors covering
//     W's own roundings and its sum.
// The energy sums over a half-wave add gamma_(rounds + 5) (A_f + A_i); the
// reference's FP64 terms and sum differ from exact by less than
// 2^-40 (A_f + A_i).  A trial the FP32 values leave open is re-taken in FP64
// over the candidates whose FP32 r^2 (final or initial) lies below the top
// of the membership band -- every term the reference sums is among them.
#include "persist.hpp"

namespace hmc {

constexpr int kLjRegionThreads = 256;  // upper bound of the workgroup size: 8 half-waves
constexpr int kLjMaxTrials = 12;       // trials of a region per pass

// The proposal of a trial of the next pass, kept in LDS across the wait for
// the neighbour regions.  Lane k of the workgroup draws trial k's, and with
// it everything of the trial that depends on the proposal alone: the trial
// position, its space test, the FP32 images of the trial and initial sites,
// and the site and LDS entry an acceptance writes.
struct LjProp {
  Sfc64 rng;     // the trial's stream after the proposal (Metropolis draws)
  double t[3];   // trial body position wrap(wrap(x + v)) (the site before wrap(.. + 0))
  double4 sp;    // the accepted site: wrap(t + 0) with its ghost flags
  float4 ent;    // its LDS entry
  float tr[3];   // FP32 box-relative trial site
  float xr[3];   // FP32 box-relative initial site (the body's LDS entry)
  uint32_t s, self, body, go, in_space, pad;
};

static_assert(sizeof(LjProp) == 160, "proposal size counted in the host's LDS budget");

struct Lj32 {
  // over the pairs within r_cut: the sums of s^12 and s^6 (s = sigma / r), of
  // s^12 / r^2 and s^6 / r^2, and their count
  float s12, s6, t12, t6, n;
};

// One candidate q (box-relative FP32 entry) against the FP32 site p: adds
// its powers to acc when within r_cut; sets *open when the membership cannot
// be decided in FP32 or the pair is closer than the error bound allows.
// The energy sum is 4 eps (S12 - S6) - n shift, its magnitude sum
// 4 eps (S12 + S6) + n |shift|, and W = 6 T12 + 3 T6.
template <int D>
__device__ __forceinline__ void lj_pair32(const float* p, float4 q, float lo, float hi, float r2_open, float sig2,
                                          Lj32& acc, bool& open) {
  const float dx = q.x - p[0], dy = q.y - p[1];
  float r2;
  if (D == 3) {
    const float dz = q.z - p[2];
    r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
  } else {
    r2 = fmaf(dx, dx, dy * dy);
  }
  const bool in = r2 < lo, out = r2 >= hi;
  if (!(in || out) || r2 < r2_open) open = true;  // also NaN: a site without images
  if (in) {
    const float rr = __builtin_amdgcn_rcpf(r2);
    const float s2 = sig2 * rr;
    const float s4 = s2 * s2, s8 = s4 * s4;
    const float s12 = s4 * s8, s6 = s2 * s4;
    acc.s12 += s12;
    acc.s6 += s6;
    acc.t12 = fmaf(s12, rr, acc.t12);
    acc.t6 = fmaf(s6, rr, acc.t6);
    acc.n += 1.0f;
  }
}

// A DPP move of v (quad_perm 0x00-0xff, row_mirror 0x140, row_half_mirror
// 0x141, row_bcast15 0x142); rows outside ROW_MASK read 0.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ float dpp_f32(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false));
}
// Sum over the 32 lanes of each half-wave, in every lane of it: within
// quads, half-rows and rows by DPP moves (no LDS round trips), then row 1
// (3) adds row 0's (2's) total and lane 31 (63) is read back.
__device__ __forceinline__ float half_sum(float v) {
  v += dpp_f32<0xb1>(v);  // quad_perm [1, 0, 3, 2]
  v += dpp_f32<0x4e>(v);  // quad_perm [2, 3, 0, 1]
  v += dpp_f32<0x141>(v);
  v += dpp_f32<0x140>(v);
  v += dpp_f32<0x142, 0xa>(v);
  const float lo = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 31));
  const float hi = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
  return (__lane_id() & 32u) ? hi : lo;
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
  const uint16_t* coff = l.coff;
  float4* lb = l.lb;
  __shared__ double zx[257], zf[257];
  __shared__ unsigned blk_acc;
  __shared__ uint32_t hlo[kLjHalo], hpre[kLjHalo + 1];
  __shared__ LjProp prop[kLjMaxTrials];
  __shared__ float shr[2][kLjRegionThreads / 32][11];  // shared trials: per half-wave partial sums, two buffers
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
  const float blo = tc.band_lo, bhi = tc.band_hi, r2_open = tc.r2_open;
  const uint32_t ntr = tc.trials;
  const uint32_t nnb = D == 3 ? 26u : 8u;
  const uint32_t nbt = wid == 0 && lane < nnb ? region_neighbour<D>(tc, rg, lane) : tile;

  if (D == 3) stage_zig(zx, zf);
  if (tid == 0) blk_acc = 0;
  const bool overflow = region_stage<D>(g, tc, rg, B, C, w, l);
  if (tid == 0) w.mbox[tile * kMailboxStride + 3] = 1u;  // progress: staged

  unsigned retests = 0, ambiguous = 0;
  // The proposals of pass i, drawn while the neighbour regions finish pass
  // i - 1 (they read only static cell data and this region's own bodies),
  // one lane per trial; trial k's waits in prop[k].  Trial k of colour c is
  // the space with digits 2 h + c_a of the region (h the digits of k in T).
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
    LjProp& pr = prop[k];
    pr.s = s;
    pr.go = 0;
    if (active && !overflow) {
      const uint32_t crow = D == 3 ? cb[0] * tc.E[1] + cb[1] : cb[0];
      const uint32_t e0 = coff[crow * CW + cb[AL]];
      const uint32_t n = coff[crow * CW + cb[AL] + 1] - e0;
      if (n) {
        Sfc64 rng = counter_rng(step, substep, seed, s);
        const uint32_t self = e0 + choose_index(rng, n);
        const float4 se = lb[self];
        const uint32_t body = (uint32_t)__float_as_int(se.w);
        const double4 bp = B.bpos[body];
        double v[3] = {0.0, 0.0, 0.0};
        ball_sample<D>(rng, max_move, zx, zf, v);
        double x[3] = {bp.x, bp.y, bp.z};
        for (int a = 0; a < D; ++a) x[a] += v[a];
        for (int a = 0; a < D; ++a) x[a] = wrap1(x[a], g.L[a]);  // generate_trial_moves: wrap(propose(..))
        double t[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < D; ++a) t[a] = wrap1(x[a], g.L[a]);  // wrap(x + 0) == wrap(x)
        // update_body_properties (microstate.rs:985-1034): site = wrap(body + 0)
        double nsp[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < D; ++a) nsp[a] = wrap1(t[a] + 0.0, g.L[a]);
        const double4 sp = pack_site(nsp, ghost_flags<D>(g, nsp));
        const float4 te = stage_entry64<D>(g, rg.c64, make_double4(t[0], t[1], t[2], 0.0), 0u);
        pr.rng = rng;
        for (int a = 0; a < 3; ++a) pr.t[a] = t[a];
        pr.sp = sp;
        pr.ent = stage_entry64<D>(g, rg.c64, sp, body);
        pr.tr[0] = te.x;
        pr.tr[1] = te.y;
        pr.tr[2] = te.z;
        pr.xr[0] = se.x;
        pr.xr[1] = se.y;
        pr.xr[2] = se.z;
        pr.self = self;
        pr.body = body;
        pr.in_space = space_of<D>(g, origin, x) == s ? 1u : 0u;
        pr.go = 1;
      }
    }
  };

  // The LDS trial of proposal pr (colour `colour`, trial k): the trial site
  // and its FP32 image, and the rows of three cells around its cell -- the
  // prefix of their lengths and the offset from a candidate's index to its
  // LDS entry (that of row 0, then its change at each later row), in
  // registers and the same in every lane.  A candidate's offset is summed
  // from masks rather than selected, so that the compiler keeps the table
  // in registers.
  struct LdsTrial {
    bool in_space;
    uint32_t self, body, total;
    double t[3];
    float tr[3], xr[3];
    uint32_t pre[NR + 1], dadj[NR];
  };
  auto lds_trial = [&](const LjProp& pr, uint32_t colour, uint32_t k) {
    LdsTrial lt;
    lt.self = pr.self;
    lt.body = pr.body;
    lt.in_space = pr.in_space != 0;
    uint32_t pcb[3] = {0, 0, 0};
    bool act = false;
    (void)trial_space(colour, k, pcb, &act);
    for (int a = 0; a < 3; ++a) {
      lt.t[a] = pr.t[a];
      lt.tr[a] = pr.tr[a];
      lt.xr[a] = pr.xr[a];
    }
    lt.pre[0] = 0;
    uint32_t adj_prev = 0;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int i0 = D == 3 ? r / 3 : r;
      const int i1 = D == 3 ? r % 3 : 0;
      const uint32_t row = D == 3 ? (pcb[0] + i0 - 1) * tc.E[1] + (pcb[1] + i1 - 1) : pcb[0] + i0 - 1;
      const uint32_t lo_r = coff[row * CW + pcb[AL] - 1];
      lt.pre[r + 1] = lt.pre[r] + (coff[row * CW + pcb[AL] + 2] - lo_r);
      const uint32_t adj_r = lo_r - lt.pre[r];
      lt.dadj[r] = adj_r - adj_prev;  // dadj[0] = the offset of row 0
      adj_prev = adj_r;
    }
    lt.total = lt.pre[NR];
    return lt;
  };
  auto slot_of = [&](const LdsTrial& lt, uint32_t c) -> uint32_t {
    uint32_t off = lt.dadj[0];
#pragma unroll
    for (int k = 1; k < NR; ++k) off += lt.dadj[k] & (0u - (uint32_t)(c >= lt.pre[k]));
    return c + off;
  };
  // FP32 sums over candidates gl, gl + GS, ... of the trial; the next
  // candidate's entry is loaded while this one is summed (past the end, and
  // for the trial's own body, the entry read is the body's own and skipped)
  auto sums32 = [&](const LdsTrial& lt, uint32_t gl, uint32_t GS, Lj32& af, Lj32& ai, bool& open) {
    uint32_t q = gl < lt.total ? slot_of(lt, gl) : lt.self;
    float4 e = lb[q];
    for (uint32_t c = gl; c < lt.total; c += GS) {
      const uint32_t qn = c + GS < lt.total ? slot_of(lt, c + GS) : lt.self;
      const float4 en = lb[qn];
      if (q != lt.self) {
        lj_pair32<D>(lt.tr, e, blo, bhi, r2_open, sig2, af, open);
        lj_pair32<D>(lt.xr, e, blo, bhi, r2_open, sig2, ai, open);
      }
      q = qn;
      e = en;
    }
  };
  // The decision from the FP32 sums of both sides (combined in FP64 into
  // energies, magnitudes and W) when their bound settles it; per-lane rounds
  // of summation `rounds`
  struct Sums {
    double s12, s6, t12, t6, n;
  };
  auto decide32 = [&](Sfc64& rng, const Sums& f, const Sums& in, double rounds, bool* decided) {
    const double e4 = (double)eps4, sh = (double)shift;
    const double ef = e4 * (f.s12 - f.s6) - f.n * sh, ei = e4 * (in.s12 - in.s6) - in.n * sh;
    const double sa = e4 * (f.s12 + f.s6 + in.s12 + in.s6) + (f.n + in.n) * fabs(sh);
    const double sw = 6.0 * (f.t12 + in.t12) + 3.0 * (f.t6 + in.t6);
    const double sb = 1.03 * (1.0 + 0x1p-16) * e4 * (double)tc.beta_max * sw;
    const double dH32 = ef - ei;
    const double delta = (((48.0 + rounds + 8.0) * 0x1p-24 + 0x1p-40) * sa + sb) * (1.0 + 0x1p-10) + 0x1p-1070;
    bool accepted = false;
    *decided = false;
    if (isfinite(dH32) && isfinite(delta)) {
      bool amb32 = false;
      accepted = lj_metropolis(rng, dH32, delta, kT, &amb32) != 0;
      *decided = !amb32;
    }
    return accepted;
  };
  // The exact FP64 re-take on one half-wave, with the same draws
  auto retake64 = [&](const LdsTrial& lt, Sfc64& rng, bool* amb) {
    const double4 sp = ld_dev_d4(&B.spos[lt.body]);
    const double xi[3] = {sp.x, sp.y, sp.z};
    const Lj64 s64 = lj_sums64<D>(g, B.spos, lt.t, xi, lt.total, [&](uint32_t c) -> uint32_t {
      const uint32_t q = slot_of(lt, c);
      if (q == lt.self) return UINT32_MAX;
      const float4 e = lb[q];
      float dfx = e.x - lt.tr[0], dfy = e.y - lt.tr[1], dix = e.x - lt.xr[0], diy = e.y - lt.xr[1];
      float r2f = fmaf(dfx, dfx, dfy * dfy), r2i = fmaf(dix, dix, diy * diy);
      if (D == 3) {
        const float dfz = e.z - lt.tr[2], diz = e.z - lt.xr[2];
        r2f = fmaf(dfz, dfz, r2f);
        r2i = fmaf(diz, diz, r2i);
      }
      // both far beyond r_cut in FP32: no term of the reference's sums
      return (r2f >= bhi && r2i >= bhi) ? UINT32_MAX : (uint32_t)__float_as_int(e.w);
    });
    return lj_decide64(rng, s64, kT, amb);
  };
  // update_body_properties (microstate.rs:985-1034) of an accepted trial
  auto accept_move = [&](const LjProp& pr) {
    B.bpos[pr.body] = make_double4(pr.t[0], pr.t[1], pr.t[2], stamp);
    st_dev_d4(&B.spos[pr.body], pr.sp);
    lb[pr.self] = pr.ent;
  };
  auto trace_trial = [&](uint32_t colour, uint32_t substep, uint32_t k, uint32_t body, bool accepted) {
    unsigned long long ti = atomicAdd((unsigned long long*)&st->trace_n, 1ull);
    if (ti < w.trace_cap) {
      uint32_t cbt[3] = {0, 0, 0};
      bool act = false;
      hmc_trace_record rec;
      rec.substep = substep;
      rec.color = colour;
      rec.space = trial_space(colour, k, cbt, &act);
      rec.body = (int64_t)body;
      rec.status = accepted ? 1 : 0;
      rec.pad = 0;
      w.trace[ti] = rec;
    }
  };

  const uint32_t npass = cycles * NCOL;
  if (npass && tid < ntr) propose(0, tid);
  __syncthreads();
  unsigned nacc = 0;  // accepted trials (lane 0 of each half-wave)
  for (uint32_t i = 0; i < npass; ++i) {
    const uint32_t colour = i & (NCOL - 1), substep = substep0 + (i >> D);
    // trials 0 .. G - 1: one half-wave each (in a region that overflows its
    // LDS, every trial: half-wave h takes h, h + G, ...)
    const uint32_t own_end = overflow ? ntr : min(ntr, G);
    for (uint32_t k = hw; k < own_end; k += G) {
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
        const LdsTrial lt = lds_trial(pr, colour, k);
        traced = true;
        tbody = lt.body;
        if (lt.in_space) {
          Lj32 af{0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, ai{0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
          bool open = false;
          sums32(lt, hl, 32u, af, ai, open);
          open = half_any(open);
          bool decided = false;
          Sfc64 rng = pr.rng;
          if (!open) {
            const Sums f{half_sum(af.s12), half_sum(af.s6), half_sum(af.t12), half_sum(af.t6), half_sum(af.n)};
            const Sums in{half_sum(ai.s12), half_sum(ai.s6), half_sum(ai.t12), half_sum(ai.t6), half_sum(ai.n)};
            accepted = decide32(rng, f, in, (double)((lt.total + 31u) / 32u), &decided);
          }
          if (!decided) {
            rng = pr.rng;
            retests += hl == 0 ? 1u : 0u;
            bool amb = false;
            accepted = retake64(lt, rng, &amb);
            ambiguous += (amb && hl == 0) ? 1u : 0u;
          }
          if (accepted && hl == 0) accept_move(pr);
        }
      }
      if (traced && hl == 0 && w.tracing) trace_trial(colour, substep, k, tbody, accepted);
      nacc += (accepted && hl == 0) ? 1u : 0u;
    }
    // trials G .. ntr - 1: the whole workgroup on each, its half-waves' sums
    // combined in LDS (in FP64, in half-wave order) and the decision taken
    // by half-wave 0 (the FP64 re-take too)
    for (uint32_t k = G; k < ntr && !overflow; ++k) {
      const LjProp& pr = prop[k];
      bool accepted = false, run = false;
      LdsTrial lt{};
      if (pr.go) {  // uniform over the workgroup
        lt = lds_trial(pr, colour, k);
        run = lt.in_space;
      }
      if (run) {
        Lj32 af{0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, ai{0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        bool open = false;
        sums32(lt, tid, blockDim.x, af, ai, open);
        float* red = shr[k & 1u][hw];
        const bool hopen = half_any(open);
        const float part[11] = {half_sum(af.s12), half_sum(af.s6), half_sum(af.t12), half_sum(af.t6), half_sum(af.n),
                                half_sum(ai.s12), half_sum(ai.s6), half_sum(ai.t12), half_sum(ai.t6), half_sum(ai.n),
                                hopen ? 1.0f : 0.0f};
        if (hl == 0)
          for (int q = 0; q < 11; ++q) red[q] = part[q];
        __syncthreads();
        if (hw == 0) {
          Sums f{0.0, 0.0, 0.0, 0.0, 0.0}, in{0.0, 0.0, 0.0, 0.0, 0.0};
          bool any_open = false;
          for (uint32_t h = 0; h < G; ++h) {
            const float* r = shr[k & 1u][h];
            f.s12 += (double)r[0];
            f.s6 += (double)r[1];
            f.t12 += (double)r[2];
            f.t6 += (double)r[3];
            f.n += (double)r[4];
            in.s12 += (double)r[5];
            in.s6 += (double)r[6];
            in.t12 += (double)r[7];
            in.t6 += (double)r[8];
            in.n += (double)r[9];
            any_open |= r[10] != 0.0f;
          }
          bool decided = false;
          Sfc64 rng = pr.rng;
          const double rounds = (double)((lt.total + blockDim.x - 1) / blockDim.x);
          if (!any_open) accepted = decide32(rng, f, in, rounds, &decided);
          if (!decided) {
            rng = pr.rng;
            retests += hl == 0 ? 1u : 0u;
            bool amb = false;
            accepted = retake64(lt, rng, &amb);
            ambiguous += (amb && hl == 0) ? 1u : 0u;
          }
          if (accepted && hl == 0) accept_move(pr);
        }
      }
      if (tid == 0) {
        if (pr.go && w.tracing) trace_trial(colour, substep, k, lt.body, accepted);
        nacc += accepted ? 1u : 0u;
      }
    }
    if (i + 1 == npass) break;  // no later pass reads the halo
    region_publish(w.mbox, nbt, nnb, i & 1u);
    if (tid == 0) w.mbox[tile * kMailboxStride + 4] = i + 1;  // progress: passes published
    if (tid < ntr) propose(i + 1, tid);
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
