// Oracle (test infrastructure only): exact-FP64 restatement of the vector
// arithmetic, periodic cuboid boundary and hypercuboid checkerboard that
// hoomd-mc's local sweep uses.  Built with -ffp-contract=off; std::fma appears
// exactly where the Rust source calls mul_add.
#pragma once
#include <array>
#include <cmath>
#include <cstdint>
#include <vector>

#include "rng.hpp"

namespace oracle {

template <int D>
using Vec = std::array<double, D>;

// hoomd-vector/src/cartesian.rs:293-296: dot = fold(0, |p,(x,y)| x.mul_add(y, p)).
template <int D>
inline double dot(const Vec<D>& a, const Vec<D>& b) {
  double p = 0.0;
  for (int i = 0; i < D; ++i) p = std::fma(a[i], b[i], p);
  return p;
}

// hoomd-vector/src/cartesian.rs:339-342: unfused sum of squared differences.
template <int D>
inline double distance_squared(const Vec<D>& a, const Vec<D>& b) {
  double p = 0.0;
  for (int i = 0; i < D; ++i) {
    double d = a[i] - b[i];
    p = p + d * d;
  }
  return p;
}

// Periodic<Hypercuboid<D>>: hoomd-microstate/src/boundary/periodic.rs:96-110 and
// boundary/periodic/cuboid.rs:47-161.
template <int D>
struct Periodic {
  std::array<double, D> L{};
  double range = 0.0;  // maximum_interaction_range (ghost criterion)

  // cuboid.rs:47-56
  double maximum_allowable_interaction_range() const {
    double m = L[0];
    for (int i = 1; i < D; ++i) m = std::fmin(m, L[i]);
    return m / 2.0;
  }
  // periodic.rs:96-110: Err when range exceeds L_min/2.
  bool valid() const { return !(range > maximum_allowable_interaction_range()); }

  // cuboid.rs:90-103.  Not idempotent in floating point.
  Vec<D> wrap(Vec<D> r) const {
    for (int i = 0; i < D; ++i) {
      double lambda = r[i] / L[i];
      lambda = lambda - std::round(lambda);
      if (lambda == 0.5) lambda = -0.5;
      r[i] = lambda * L[i];
    }
    return r;
  }

  // cuboid.rs:108-161.  Returns the number of ghosts written to out (<= 2^D-1),
  // in the reference's descending-subset order.
  int ghosts(const Vec<D>& r, Vec<D>* out) const {
    uint32_t mask = 0;
    double t[D] = {};
    for (int i = 0; i < D; ++i) {
      double half = L[i] / 2.0;
      double x = r[i];
      if (x < -half || x >= half) return 0;
      if (x > half - range) {
        mask |= 1u << i;
        t[i] = -L[i];
      } else if (x < -half + range) {
        mask |= 1u << i;
        t[i] = L[i];
      }
    }
    int n = 0;
    uint32_t subset = mask;
    while (subset != 0) {
      Vec<D> g = r;
      for (int i = 0; i < D; ++i)
        if (subset & (1u << i)) g[i] += t[i];
      out[n++] = g;
      subset = (subset - 1) & mask;
    }
    return n;
  }
};

// HypercuboidCheckerboard<D> for a periodic cuboid: hoomd-mc/src/hypercuboid.rs:48-388.
template <int D>
struct Checkerboard {
  Vec<D> origin{};
  std::array<double, D> width{};
  std::array<int64_t, D> shape{};
  std::vector<std::vector<uint64_t>> by_color;

  // hypercuboid.rs:163-173 (last axis fastest)
  static uint64_t multi_to_index(const std::array<int64_t, D>& m, const std::array<int64_t, D>& shape) {
    uint64_t index = 0, w = 1;
    for (int i = D - 1; i >= 0; --i) {
      index += (uint64_t)m[i] * w;
      w *= (uint64_t)shape[i];
    }
    return index;
  }
  // hypercuboid.rs:177-184
  static std::array<int64_t, D> index_to_multi(uint64_t index, const std::array<int64_t, D>& shape) {
    std::array<int64_t, D> m{};
    for (int i = D - 1; i >= 0; --i) {
      m[i] = (int64_t)(index % (uint64_t)shape[i]);
      index /= (uint64_t)shape[i];
    }
    return m;
  }

  // hypercuboid.rs:188-228.  Returns false where the reference asserts.
  static bool compute_dimensions(const std::array<double, D>& edges, double range, bool periodic,
                                 std::array<double, D>& width, std::array<int64_t, D>& shape) {
    std::array<int64_t, D> inside{};
    for (int i = 0; i < D; ++i) inside[i] = (int64_t)std::floor(edges[i] / range);
    for (int i = 0; i < D; ++i)
      if (inside[i] < 2) return false;
    for (int i = 0; i < D; ++i) {
      if (periodic && inside[i] % 2 != 0) inside[i] -= 1;
      if (!periodic && inside[i] % 2 == 0) inside[i] -= 1;
    }
    for (int i = 0; i < D; ++i) {
      width[i] = edges[i] / (double)inside[i];
      shape[i] = periodic ? inside[i] : inside[i] + 1;
    }
    return true;
  }

  // hypercuboid.rs:231-255
  static std::vector<std::vector<uint64_t>> construct_by_color(const std::array<int64_t, D>& shape) {
    std::array<int64_t, D> half{};
    uint64_t total = 1;
    for (int i = 0; i < D; ++i) {
      half[i] = shape[i] / 2;
      total *= (uint64_t)half[i];
    }
    std::vector<uint64_t> base(total);
    for (uint64_t idx = 0; idx < total; ++idx) {
      auto hm = index_to_multi(idx, half);
      std::array<int64_t, D> m{};
      for (int i = 0; i < D; ++i) m[i] = 2 * hm[i];
      base[idx] = multi_to_index(m, shape);
    }
    std::array<int64_t, D> two{};
    for (int i = 0; i < D; ++i) two[i] = 2;
    std::vector<std::vector<uint64_t>> out(1u << D);
    for (uint64_t color = 0; color < (1u << D); ++color) {
      uint64_t delta = multi_to_index(index_to_multi(color, two), shape);
      out[color].resize(total);
      for (uint64_t k = 0; k < total; ++k) out[color][k] = base[k] + delta;
    }
    return out;
  }

  // hypercuboid.rs:315-341 via Cover for Periodic<Hypercuboid> (:367-388).
  // The offsets are drawn in axis order with rand_f64.
  bool update(Sfc64& rng, double range, const std::array<double, D>& edges) {
    std::array<double, D> w{};
    std::array<int64_t, D> s{};
    if (!compute_dimensions(edges, range, true, w, s)) return false;
    double offset[D];
    for (int i = 0; i < D; ++i) offset[i] = rand_f64(rng);
    for (int i = 0; i < D; ++i) origin[i] = -edges[i] / 2.0 - offset[i] * w[i];
    if (s != shape || by_color.empty()) {
      by_color = construct_by_color(s);
      shape = s;
    }
    width = w;
    return true;
  }

  // Fixed origin at -L/2 (hypercuboid.rs:259-274, test helper).
  bool with_fixed_origin(double range, const std::array<double, D>& edges, bool periodic) {
    if (!compute_dimensions(edges, range, periodic, width, shape)) return false;
    for (int i = 0; i < D; ++i) origin[i] = -edges[i] / 2.0;
    by_color = construct_by_color(shape);
    periodic_ = periodic;
    return true;
  }

  uint64_t num_spaces() const {
    uint64_t n = 1;
    for (int i = 0; i < D; ++i) n *= (uint64_t)shape[i];
    return n;
  }

  // hypercuboid.rs:86-127.  Returns -1 for None.
  int64_t point_to_space_index(const Vec<D>& point) const {
    std::array<int64_t, D> m{};
    for (int i = 0; i < D; ++i) {
      double p = point[i] - origin[i];
      double q = std::floor(p / width[i]);
      m[i] = to_i64(q);
    }
    for (int i = 0; i < D; ++i) {
      if (m[i] < 0) return -1;
      if (periodic_) {
        if (m[i] == shape[i]) m[i] = 0;
        if (m[i] > shape[i]) return -1;
      } else if (m[i] >= shape[i]) {
        return -1;
      }
    }
    return (int64_t)multi_to_index(m, shape);
  }

  bool periodic_ = true;

 private:
  // Rust `as i64`: saturating, NaN -> 0.
  static int64_t to_i64(double q) {
    if (std::isnan(q)) return 0;
    if (q >= 9.2233720368547758e18) return INT64_MAX;
    if (q <= -9.2233720368547758e18) return INT64_MIN;
    return (int64_t)q;
  }
};

}  // namespace oracle
