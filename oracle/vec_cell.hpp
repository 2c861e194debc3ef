// Oracle (test infrastructure only): restatement of hoomd-spatial's VecCell
// (hoomd-spatial/src/vec_cell.rs:93-733).  Iteration order (stencil order x
// in-cell insertion order with swap_remove) is kept because it fixes the
// floating-point summation order of pair energies.
//
// One deliberate difference in cost, not in results: the key -> cell map is a
// dense array indexed by site/ghost tag instead of an FxHashMap.  That makes
// this CPU baseline slightly faster than the reference, never slower.
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <vector>

namespace oracle {

// SiteKey (hoomd-microstate/src/microstate.rs:28-34) packed as tag*2 + is_ghost.
inline uint64_t key_primary(uint64_t tag) { return tag << 1; }
inline uint64_t key_ghost(uint64_t tag) { return (tag << 1) | 1; }

template <int D>
struct VecCell {
  using CellIndex = std::array<int64_t, D>;
  double w = 1.0;
  std::vector<std::vector<uint64_t>> keys_map;
  uint32_t half_extent = 1;
  std::vector<std::vector<CellIndex>> stencils;
  // dense replacement of cell_index: FxHashMap<K, CellIndex<D>>
  std::vector<CellIndex> cell_of[2];
  std::vector<uint8_t> has_cell[2];

  // vec_cell.rs:117-164 + 167-177: row-major cube, stable sort by the
  // lexicographic order of the squared components (Iterator::cmp).
  static std::vector<CellIndex> generate_stencil(uint32_t radius) {
    std::vector<CellIndex> result;
    CellIndex c;
    for (int i = 0; i < D; ++i) c[i] = -(int64_t)radius;
    for (;;) {
      result.push_back(c);
      int i = D - 1;
      c[i] += 1;
      while (i >= 0 && c[i] > (int64_t)radius) {
        if (i == 0) goto done;
        c[i] = -(int64_t)radius;
        c[i - 1] += 1;
        --i;
      }
    }
  done:
    std::stable_sort(result.begin(), result.end(), [](const CellIndex& a, const CellIndex& b) {
      for (int i = 0; i < D; ++i) {
        int64_t ra = a[i] * a[i], rb = b[i] * b[i];
        if (ra != rb) return ra < rb;
      }
      return false;
    });
    return result;
  }

  // vec_cell.rs:286-299 with nominal = maximum = radius (WithSearchRadius).
  explicit VecCell(double radius = 1.0) : w(radius) {
    uint32_t max_stencil = (uint32_t)std::ceil(radius / radius);
    if (max_stencil < 1) max_stencil = 1;
    { size_t t = 1; for (int i = 0; i < D; ++i) t *= 3; keys_map.assign(t, {}); }
    for (uint32_t r = 0; r < max_stencil; ++r) stencils.push_back(generate_stencil(r + 1));
  }

  // vec_cell.rs:356-358
  CellIndex cell_index_from_position(const std::array<double, D>& p) const {
    CellIndex c;
    for (int j = 0; j < D; ++j) c[j] = (int64_t)std::floor(p[j] / w);
    return c;
  }

  // vec_cell.rs:364-383.  Returns -1 for None.
  static int64_t map_index_from_cell(uint32_t h, const CellIndex& c) {
    int64_t index = 0, width = 1;
    for (int i = D - 1; i >= 0; --i) {
      if ((uint64_t)std::llabs(c[i]) > h) return -1;
      index += (c[i] + (int64_t)h) * width;
      width *= (int64_t)(h * 2 + 1);
    }
    return index;
  }

  // vec_cell.rs:447-475
  void expand_to(uint32_t target) {
    if (half_extent >= target) return;
    uint32_t nh = std::min(half_extent, 1u) * 2;
    while (nh < target) nh *= 2;
    size_t side = nh * 2 + 1, total = 1;
    for (int i = 0; i < D; ++i) total *= side;
    std::vector<std::vector<uint64_t>> nk(total);
    // walk the old cube
    CellIndex c;
    for (int i = 0; i < D; ++i) c[i] = -(int64_t)half_extent;
    for (;;) {
      nk[map_index_from_cell(nh, c)] = std::move(keys_map[map_index_from_cell(half_extent, c)]);
      int i = D - 1;
      c[i] += 1;
      bool end = false;
      while (c[i] > (int64_t)half_extent) {
        if (i == 0) { end = true; break; }
        c[i] = -(int64_t)half_extent;
        c[i - 1] += 1;
        --i;
      }
      if (end) break;
    }
    half_extent = nh;
    keys_map = std::move(nk);
  }

  void ensure(uint64_t key) {
    int g = (int)(key & 1);
    uint64_t tag = key >> 1;
    if (cell_of[g].size() <= tag) {
      cell_of[g].resize(tag + 1);
      has_cell[g].resize(tag + 1, 0);
    }
  }

  // vec_cell.rs:492-527
  void insert(uint64_t key, const std::array<double, D>& pos) {
    CellIndex c = cell_index_from_position(pos);
    ensure(key);
    int g = (int)(key & 1);
    uint64_t tag = key >> 1;
    bool had = has_cell[g][tag];
    CellIndex old = cell_of[g][tag];
    cell_of[g][tag] = c;
    has_cell[g][tag] = 1;
    int64_t mi = map_index_from_cell(half_extent, c);
    if (mi < 0) {
      uint64_t m = 0;
      for (int i = 0; i < D; ++i) m = std::max<uint64_t>(m, (uint64_t)std::llabs(c[i]));
      expand_to((uint32_t)m);
      mi = map_index_from_cell(half_extent, c);
    }
    if (!(had && old == c)) {
      keys_map[mi].push_back(key);
      if (had) {
        auto& ok = keys_map[map_index_from_cell(half_extent, old)];
        auto it = std::find(ok.begin(), ok.end(), key);
        if (it != ok.end()) {
          *it = ok.back();
          ok.pop_back();
        }
      }
    }
  }

  // vec_cell.rs:541-552
  void remove(uint64_t key) {
    int g = (int)(key & 1);
    uint64_t tag = key >> 1;
    if (tag >= has_cell[g].size() || !has_cell[g][tag]) return;
    has_cell[g][tag] = 0;
    int64_t mi = map_index_from_cell(half_extent, cell_of[g][tag]);
    if (mi >= 0) {
      auto& k = keys_map[mi];
      auto it = std::find(k.begin(), k.end(), key);
      if (it != k.end()) {
        *it = k.back();
        k.pop_back();
      }
    }
  }

  // vec_cell.rs:650-673 + 710-733: visit every key in the stencil around pos,
  // in order.  f returns false to stop early.
  template <class F>
  void points_near_ball(const std::array<double, D>& pos, double radius, F&& f) const {
    size_t si = (size_t)std::ceil(radius / w) - 1;
    if (si >= stencils.size()) std::abort();
    CellIndex center = cell_index_from_position(pos);
    for (const CellIndex& off : stencils[si]) {
      CellIndex c;
      for (int i = 0; i < D; ++i) c[i] = center[i] + off[i];
      int64_t mi = map_index_from_cell(half_extent, c);
      if (mi < 0) continue;
      for (uint64_t k : keys_map[mi])
        if (!f(k)) return;
    }
  }
};

}  // namespace oracle
