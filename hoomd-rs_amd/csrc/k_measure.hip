// Measurements over the current state: total energy, insertion energies of
// candidate bodies, and the scale distribution function.
#include "sweep_device.hpp"

namespace hmc {

// ------------------------------------------------------------- total energy --
// PairwiseCutoff::total_energy (pairwise_cutoff.rs:720-745): the sum over
// sites i and neighbours j with site_tag_i < site_tag_j (body index = site
// tag for single-site bodies) of the pair energy, infinite on any hard
// overlap.  Per block: the partial sum, the sum of |terms|, the number of
// non-zero terms and an overlap flag, reduced in a fixed order.

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

// explicit instantiations (launched from the host translation units)
template __global__ void k_energy<2>(Geometry g, const ApplyState* st, Bodies B, Work w, double* part);
template __global__ void k_energy<3>(Geometry g, const ApplyState* st, Bodies B, Work w, double* part);
template __global__ void k_insert_energy<2>(Geometry g, const ApplyState* st, Bodies B, Work w, const double* __restrict__ cand, uint32_t n, double* __restrict__ out, uint32_t* outside);
template __global__ void k_insert_energy<3>(Geometry g, const ApplyState* st, Bodies B, Work w, const double* __restrict__ cand, uint32_t n, double* __restrict__ out, uint32_t* outside);
template __global__ void k_sdf<2>(const BatchRep* reps, const uint32_t* boff, uint32_t R, uint32_t n_total, Bodies B, Work w, double xmax, double dx, uint32_t nbins, uint32_t* hist, uint32_t* ovl);
template __global__ void k_sdf<3>(const BatchRep* reps, const uint32_t* boff, uint32_t R, uint32_t n_total, Bodies B, Work w, double xmax, double dx, uint32_t nbins, uint32_t* hist, uint32_t* ovl);

}  // namespace hmc
