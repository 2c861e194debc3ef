// Oracle (test infrastructure only): the scale distribution function (SDF) of
// hard spheres and disks.  The reference has no pressure compute; the
// definition is HOOMD-blue's hpmc compute.SDF.  For each body, x is the
// smallest compression 1 - sigma / r over the other bodies (r the
// minimum-image distance between sites) at which the isotropically scaled
// system would first overlap.  Bodies with 0 <= x < xmax are counted in bin
// floor(x / dx) when that bin is below nbins, and bodies with x < 0 (overlaps)
// are counted apart.  x is monotone in r^2, so the minimum is taken over
// r^2 = (d_x^2 + d_y^2) + d_z^2.  Two neighbour searches give the same minima
// whenever x < xmax: all pairs, and a cell grid of width >= r_max (1 + 2^-30),
// r_max = sigma / (1 - xmax).  Built with -ffp-contract=off.
#pragma once
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

namespace oracle {

// Minimum-image squared distance: per axis d = a - b, moved by one box edge
// when |d| > L / 2 (positions lie in [-L/2, L/2)).
template <int D>
inline double sdf_r2(const double* a, const double* b, const double* L) {
  double r2 = 0.0;
  for (int k = 0; k < D; ++k) {
    double d = a[k] - b[k];
    const double h = L[k] / 2.0;
    if (d > h)
      d -= L[k];
    else if (d < -h)
      d += L[k];
    r2 = r2 + d * d;
  }
  return r2;
}

template <int D>
void sdf_min_r2_all(const double* pos, uint64_t n, const double* L, double* best) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    double m = std::numeric_limits<double>::infinity();
    for (uint64_t j = 0; j < n; ++j)
      if (j != (uint64_t)i) m = std::fmin(m, sdf_r2<D>(pos + i * D, pos + j * D, L));
    best[i] = m;
  }
}

// Cell grid of its own: nc[k] = floor(L / (r_max (1 + 2^-30))) cells per axis
// (all pairs when an axis has fewer than 3), each body visiting the 3^D
// cells around its own.
template <int D>
void sdf_min_r2_cells(const double* pos, uint64_t n, const double* L, double rmax, double* best) {
  int nc[3] = {1, 1, 1};
  for (int k = 0; k < D; ++k) {
    nc[k] = (int)std::floor(L[k] / (rmax * (1.0 + 0x1p-30)));
    if (nc[k] < 3) {
      sdf_min_r2_all<D>(pos, n, L, best);
      return;
    }
  }
  auto cell_of = [&](const double* p, int* c) {
    for (int k = 0; k < D; ++k) {
      int v = (int)std::floor((p[k] + L[k] / 2.0) / (L[k] / nc[k]));
      c[k] = v < 0 ? 0 : v >= nc[k] ? nc[k] - 1 : v;
    }
  };
  const int ncell = nc[0] * nc[1] * (D == 3 ? nc[2] : 1);
  std::vector<uint32_t> start(ncell + 1, 0), order(n);
  std::vector<int> cid(n);
  for (uint64_t i = 0; i < n; ++i) {
    int c[3] = {0, 0, 0};
    cell_of(pos + i * D, c);
    cid[i] = c[0] + nc[0] * (c[1] + nc[1] * (D == 3 ? c[2] : 0));
    start[cid[i] + 1] += 1;
  }
  for (int c = 0; c < ncell; ++c) start[c + 1] += start[c];
  std::vector<uint32_t> fill(start.begin(), start.end() - 1);
  for (uint64_t i = 0; i < n; ++i) order[fill[cid[i]]++] = (uint32_t)i;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    int c[3] = {0, 0, 0};
    cell_of(pos + i * D, c);
    double m = std::numeric_limits<double>::infinity();
    for (int o0 = -1; o0 <= 1; ++o0)
      for (int o1 = -1; o1 <= 1; ++o1)
        for (int o2 = (D == 3 ? -1 : 0); o2 <= (D == 3 ? 1 : 0); ++o2) {
          const int c0 = (c[0] + o0 + nc[0]) % nc[0], c1 = (c[1] + o1 + nc[1]) % nc[1];
          const int c2 = D == 3 ? (c[2] + o2 + nc[2]) % nc[2] : 0;
          const int cc = c0 + nc[0] * (c1 + nc[1] * c2);
          for (uint32_t q = start[cc]; q < start[cc + 1]; ++q) {
            const uint32_t j = order[q];
            if (j != (uint64_t)i) m = std::fmin(m, sdf_r2<D>(pos + i * D, pos + (uint64_t)j * D, L));
          }
        }
    best[i] = m;
  }
}

template <int D>
void sdf_histogram(const double* pos, uint64_t n, const double* L, double sigma, double xmax, double dx,
                   uint32_t nbins, bool all_pairs, uint64_t* hist, uint64_t* overlaps) {
  std::vector<double> best(n);
  if (all_pairs)
    sdf_min_r2_all<D>(pos, n, L, best.data());
  else
    sdf_min_r2_cells<D>(pos, n, L, sigma / (1.0 - xmax), best.data());
  for (uint32_t k = 0; k < nbins; ++k) hist[k] = 0;
  uint64_t ov = 0;
  for (uint64_t i = 0; i < n; ++i) {
    if (!(best[i] < std::numeric_limits<double>::infinity())) continue;
    const double x = 1.0 - sigma / std::sqrt(best[i]);
    if (x < 0.0) {
      ov += 1;
    } else if (x < xmax) {
      const double k = std::floor(x / dx);
      if (k < (double)nbins) hist[(uint32_t)k] += 1;
    }
  }
  *overlaps = ov;
}

}  // namespace oracle
