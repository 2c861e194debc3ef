// Move-size tuning and grand-canonical energies: trial moves of random
// bodies evaluated against the unchanged state.
#include "sweep_device.hpp"

namespace hmc {


// Energy of site t with every site of another body than `own` (UINT32_MAX:
// none excluded), over the neighbour spaces of t's space in the cells of the
// current state: the filtered site energy of pairwise_cutoff.rs:355-387.
// INFINITY on an overlap.  *bound: 2 gamma_n sum |e|, the reach of another
// summation order (the reference sums in VecCell order).
template <int D>
__device__ double site_energy_excl(const Geometry& g, const ApplyState* st, const Bodies& B, const Work& w,
                                   const double* t, uint32_t own, double* bound, bool* outside) {
  *bound = 0.0;
  const uint32_t s = space_of<D>(g, st->origin, t);
  if (s == UINT32_MAX) {
    *outside = true;
    return 0.0;
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
  for (int i0 = 0; i0 < nn[0]; ++i0)
    for (int i1 = 0; i1 < nn[1]; ++i1)
      for (int i2 = 0; i2 < (D == 3 ? nn[2] : 1); ++i2) {
        const uint32_t ns = nb[0][i0] + nb[1][i1] + (D == 3 ? nb[2][i2] : 0u);
        const uint32_t q0 = w.start[ns], q1 = w.start[ns + 1];
        for (uint32_t q = q0; q < q1; ++q) {
          const uint32_t bj = w.tmp_body[q];
          if (bj == own) continue;
          const double4 sj = B.spos[bj];
          if (g.model == HMC_HARD_SPHERE) {
            if (overlap64<D>(g, t, sj)) return INFINITY;
          } else {
            const double r2 = lj_r2<D>(g, t, sj);
            if (r2 < g.lj_rc2) {
              const double one = lj_energy(g, r2);
              if (one == INFINITY) return INFINITY;
              e += one;
              a += fabs(one);
              cnt += 1.0;
            }
          }
        }
      }
  const double nt = fmax(cnt, 1.0);
  const double gam = nt * 0x1p-53 / (1.0 - nt * 0x1p-53);
  *bound = 2.0 * gam * a;
  return e;
}

// The comparison u < exp(-dH / kT) of a Metropolis test for dH known to
// within +-delta of the reference's value, u given: decided by the fast exp
// where u is clear of exp(q) for every q of the interval, else by glibc's
// exp at the GPU's dH with *amb set (delta > 0).
__device__ __forceinline__ bool metropolis_u(double u, double dH, double delta, double kT, bool* amb) {
  const double hi = dH + delta, lo = dH - delta;
  const double qa = -hi / kT;
  const double ea = exp(qa);
  if (u < ea * (1.0 - 0x1p-45 * (1.0 + fabs(qa)))) return true;
  if (lo > 0.0) {
    const double qb = -lo / kT;
    const double eb = exp(qb);
    if (u >= eb * (1.0 + 0x1p-45 * (1.0 + fabs(qb)))) return false;
  }
  *amb |= delta > 0.0;
  return u < glibc_exp(-dH / kT);
}

// tune_local_trial (hoomd-mc/src/tune_local.rs:21-89) with Translate, one
// sample read from every offset p of the pre-drawn stream xs[0, M): body =
// random_range(0..N), trial = wrap(propose(body)), dH = delta_energy_one
// against the unchanged state, and the acceptance draw when 0 < dH < inf.
// next[p]: the offset after the sample (UINT32_MAX: the stream ran out);
// flags[p]: bit 0 accepted, bit 1 the bound on the summation order could not
// fix whether the reference draws or accepts, bit 2 a site outside the
// checkerboard (an error).  The samples of a tuning step
// are the chain 0, next[0], next[next[0]], ... (followed on the host).
template <int D>
__global__ void __launch_bounds__(kEnergyBlock) k_tune_eval(Geometry g, const ApplyState* st, Bodies B, Work w,
                                                            const uint64_t* __restrict__ xs, uint32_t M,
                                                            double max_move, double kT, uint32_t* __restrict__ next,
                                                            uint8_t* __restrict__ flags) {
  __shared__ double zx[257], zf[257];
  stage_zig(zx, zf);
  __syncthreads();
  const uint32_t p = blockIdx.x * kEnergyBlock + threadIdx.x;
  if (p >= M) return;
  StreamRng r{xs, p, M, false};
  const uint32_t body = choose_index(r, g.n_bodies);
  const double4 bp = B.bpos[body];
  double x[3] = {bp.x, bp.y, D == 3 ? bp.z : 0.0};
  double v[3] = {0.0, 0.0, 0.0};
  ball_sample<D>(r, max_move, zx, zf, v);
  for (int a = 0; a < D; ++a) x[a] = wrap1(x[a] + v[a], g.L[a]);  // wrap(propose(..))
  double t[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < D; ++a) t[a] = wrap1(x[a] + 0.0, g.L[a]);  // the trial's site
  bool outside = false;
  double bf = 0.0, bi = 0.0;
  const double ef = site_energy_excl<D>(g, st, B, w, t, body, &bf, &outside);
  double dh = ef, delta = bf;
  if (ef != INFINITY && g.model != HMC_HARD_SPHERE) {  // pairwise_cutoff.rs:919-941
    const double4 sp = B.spos[body];
    const double ti[3] = {sp.x, sp.y, D == 3 ? sp.z : 0.0};
    const double ei = site_energy_excl<D>(g, st, B, w, ti, body, &bi, &outside);
    dh = ef - ei;
    delta = bf + bi;
  }
  bool acc = false, amb = false;
  if (dh == INFINITY) {
    acc = false;
  } else if (dh + delta <= 0.0) {
    acc = true;  // dH <= 0 for the reference too: no draw
  } else {
    amb |= dh - delta <= 0.0;  // whether the reference draws is open
    if (dh <= 0.0)
      acc = true;
    else
      acc = metropolis_u(uniform01(r), dh, delta, kT, &amb);
  }
  next[p] = r.over ? 0xFFFFFFFFu : r.pos;
  flags[p] = (uint8_t)((acc ? 1u : 0u) | (amb ? 2u : 0u) | (outside ? 4u : 0u));
}

// DeltaEnergyRemove (pairwise_cutoff.rs:1097-1109) of body i: minus the
// initial energy of its site with the other bodies' sites (0 for hard
// spheres, whose initial energies the reference takes as 0).  out = {dE, bound}.
template <int D>
__global__ void k_remove_energy(Geometry g, const ApplyState* st, Bodies B, Work w, uint32_t body, double* out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  if (g.model == HMC_HARD_SPHERE) {
    out[0] = -0.0;
    out[1] = 0.0;
    return;
  }
  const double4 sp = B.spos[body];
  const double ti[3] = {sp.x, sp.y, D == 3 ? sp.z : 0.0};
  bool outside = false;
  double b = 0.0;
  const double e = site_energy_excl<D>(g, st, B, w, ti, body, &b, &outside);
  out[0] = -e;
  out[1] = outside ? -1.0 : b;
}

template __global__ void k_tune_eval<2>(Geometry, const ApplyState*, Bodies, Work, const uint64_t* __restrict__,
                                        uint32_t, double, double, uint32_t* __restrict__, uint8_t* __restrict__);
template __global__ void k_tune_eval<3>(Geometry, const ApplyState*, Bodies, Work, const uint64_t* __restrict__,
                                        uint32_t, double, double, uint32_t* __restrict__, uint8_t* __restrict__);
template __global__ void k_remove_energy<2>(Geometry, const ApplyState*, Bodies, Work, uint32_t, double*);
template __global__ void k_remove_energy<3>(Geometry, const ApplyState*, Bodies, Work, uint32_t, double*);

}  // namespace hmc
