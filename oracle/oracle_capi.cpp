// Oracle (test infrastructure only): C entry points over the restatement, for
// the pytest suite and the CPU-baseline leg of bench.py (loaded with ctypes).
// Never linked into, or called by, the product library.
#include <omp.h>

#include <chrono>
#include <cstdint>
#include <cstring>
#include <memory>

#include "convex.hpp"
#include "geometry.hpp"
#include "mc.hpp"
#include "rng.hpp"
#include "sdf.hpp"
#include "vec_cell.hpp"

using namespace oracle;

namespace {

struct Handle {
  int dim;
  std::unique_ptr<System<2>> s2;
  std::unique_ptr<System<3>> s3;
};

template <int D>
void fill_model(System<D>& s, int model, const double* params) {
  s.model.kind = model;
  if (model == 0) {
    s.model.diameter = params[0];
  } else if (model == 4) {
    s.model.k = params[0];
    s.model.max_overlap = params[1];
    s.model.eps_shoulder = params[2];
    s.model.delta = params[3];
    s.model.r_cut = params[3];
  } else {
    s.model.epsilon = params[0];
    s.model.sigma = params[1];
    s.model.r_cut = params[2];
    s.model.r_shift = params[3];
  }
}

template <int D>
System<D>* make(const double* L, double boundary_range, double body_range, int model, const double* params,
                uint64_t step, uint32_t seed, const double* pos, uint64_t n) {
  std::array<double, D> edges;
  for (int i = 0; i < D; ++i) edges[i] = L[i];
  PairModel pm;
  pm.kind = model;
  double radius = model == 0 ? params[0] : model == 4 ? params[3] : params[2];
  auto* s = new System<D>(edges, boundary_range, radius);
  fill_model(*s, model, params);
  s->body_range = body_range;
  s->step = step;
  s->seed = seed;
  if (!s->boundary.valid()) {
    delete s;
    return nullptr;
  }
  for (uint64_t b = 0; b < n; ++b) {
    Vec<D> p;
    for (int i = 0; i < D; ++i) p[i] = pos[b * D + i];
    s->add_body(p);
  }
  return s;
}

}  // namespace

extern "C" {

// ---- host C library ---------------------------------------------------------
// The library functions the reference calls, evaluated in bulk on the host:
// kind 0 cbrt, 1 log, 2 exp, 5 and 6 the sine and cosine of one sincos call.
void oracle_libm_eval(int kind, const double* in, double* out, uint64_t n) {
  for (uint64_t i = 0; i < n; ++i) {
    double s, c;
    switch (kind) {
      case 0: out[i] = std::cbrt(in[i]); break;
      case 1: out[i] = std::log(in[i]); break;
      case 2: out[i] = std::exp(in[i]); break;
      case 5: lib_sincos(in[i], &s, &c); out[i] = s; break;
      case 6: lib_sincos(in[i], &s, &c); out[i] = c; break;
      default: out[i] = NAN; break;
    }
  }
}

// ---- RNG and adapter probes -------------------------------------------------
void oracle_sfc64_stream(uint64_t a, uint64_t b, uint64_t c, uint64_t ctr, int warm, uint64_t* out, uint64_t n) {
  Sfc64 r;
  if (warm) {
    r = Sfc64::initialize(a, b, c, ctr);
  } else {
    r.a = a; r.b = b; r.c = c; r.ctr = ctr;
  }
  for (uint64_t i = 0; i < n; ++i) out[i] = r.next_u64();
}
void oracle_counter_stream(uint64_t step, uint32_t substep, uint32_t seed, uint64_t index, uint64_t* out, uint64_t n) {
  Sfc64 r = counter_rng(step, substep, seed, index);
  for (uint64_t i = 0; i < n; ++i) out[i] = r.next_u64();
}
// kind: 0 rand_f64, 1 open01, 2 standard_normal, 3 rand_index(len=param)
void oracle_sample(uint64_t step, uint32_t substep, uint32_t seed, uint64_t index, int kind, uint32_t param,
                   double* out, uint64_t n) {
  Sfc64 r = counter_rng(step, substep, seed, index);
  for (uint64_t i = 0; i < n; ++i) {
    switch (kind) {
      case 0: out[i] = rand_f64(r); break;
      case 1: out[i] = rand_open01(r); break;
      case 2: out[i] = standard_normal(r); break;
      default: out[i] = (double)rand_index(r, param); break;
    }
  }
}
void oracle_zig_tables(double* x, double* f) {
  std::memcpy(x, zig().x, sizeof(double) * 257);
  std::memcpy(f, zig().f, sizeof(double) * 257);
}
// Translate proposal for one (step, substep, seed, index) stream, as used by a
// ParallelSweep trial after the choose draw.
void oracle_translate(int dim, uint64_t step, uint32_t substep, uint32_t seed, uint64_t index, uint32_t choose_len,
                      const double* pos, double r, double* out, uint32_t* chosen) {
  Sfc64 rng = counter_rng(step, substep, seed, index);
  *chosen = rand_index(rng, choose_len);
  if (dim == 2) {
    Vec<2> p{pos[0], pos[1]};
    auto q = System<2>::propose_translate(rng, p, r);
    out[0] = q[0]; out[1] = q[1];
  } else {
    Vec<3> p{pos[0], pos[1], pos[2]};
    auto q = System<3>::propose_translate(rng, p, r);
    out[0] = q[0]; out[1] = q[1]; out[2] = q[2];
  }
}

// ---- boundary / checkerboard / VecCell probes -------------------------------
void oracle_wrap(int dim, const double* L, const double* in, double* out, uint64_t n) {
  for (uint64_t k = 0; k < n; ++k) {
    if (dim == 2) {
      Periodic<2> p; p.L = {L[0], L[1]};
      auto r = p.wrap({in[2 * k], in[2 * k + 1]});
      out[2 * k] = r[0]; out[2 * k + 1] = r[1];
    } else if (dim == 3) {
      Periodic<3> p; p.L = {L[0], L[1], L[2]};
      auto r = p.wrap({in[3 * k], in[3 * k + 1], in[3 * k + 2]});
      for (int i = 0; i < 3; ++i) out[3 * k + i] = r[i];
    } else {
      Periodic<1> p; p.L = {L[0]};
      out[k] = p.wrap({in[k]})[0];
    }
  }
}
int oracle_ghosts(int dim, const double* L, double range, const double* in, double* out) {
  if (dim == 1) {
    Periodic<1> p; p.L = {L[0]}; p.range = range;
    Vec<1> g[1];
    int n = p.ghosts({in[0]}, g);
    for (int k = 0; k < n; ++k) out[k] = g[k][0];
    return n;
  } else if (dim == 2) {
    Periodic<2> p; p.L = {L[0], L[1]}; p.range = range;
    Vec<2> g[3];
    int n = p.ghosts({in[0], in[1]}, g);
    for (int k = 0; k < n; ++k) { out[2 * k] = g[k][0]; out[2 * k + 1] = g[k][1]; }
    return n;
  }
  Periodic<3> p; p.L = {L[0], L[1], L[2]}; p.range = range;
  Vec<3> g[7];
  int n = p.ghosts({in[0], in[1], in[2]}, g);
  for (int k = 0; k < n; ++k)
    for (int i = 0; i < 3; ++i) out[3 * k + i] = g[k][i];
  return n;
}
// Returns 0 on success, -1 where the reference asserts.
int oracle_compute_dimensions(int dim, const double* edges, double range, const int* periodic, double* width,
                              int64_t* shape) {
  if (dim == 2) {
    std::array<double, 2> w; std::array<int64_t, 2> s;
    std::array<double, 2> e{edges[0], edges[1]};
    if (!Checkerboard<2>::compute_dimensions(e, range, periodic[0] != 0, w, s)) return -1;
    for (int i = 0; i < 2; ++i) { width[i] = w[i]; shape[i] = s[i]; }
    return 0;
  }
  std::array<double, 3> w; std::array<int64_t, 3> s;
  std::array<double, 3> e{edges[0], edges[1], edges[2]};
  if (!Checkerboard<3>::compute_dimensions(e, range, periodic[0] != 0, w, s)) return -1;
  for (int i = 0; i < 3; ++i) { width[i] = w[i]; shape[i] = s[i]; }
  return 0;
}
uint64_t oracle_multi_to_index(int dim, const int64_t* m, const int64_t* shape) {
  uint64_t index = 0, w = 1;
  for (int i = dim - 1; i >= 0; --i) {
    index += (uint64_t)m[i] * w;
    w *= (uint64_t)shape[i];
  }
  return index;
}
// Colour sets of construct_space_indices_by_color, flattened colour-major.
void oracle_by_color(int dim, const int64_t* shape, uint64_t* out) {
  size_t k = 0;
  if (dim == 2) {
    auto v = Checkerboard<2>::construct_by_color({shape[0], shape[1]});
    for (auto& c : v) for (auto x : c) out[k++] = x;
  } else if (dim == 3) {
    auto v = Checkerboard<3>::construct_by_color({shape[0], shape[1], shape[2]});
    for (auto& c : v) for (auto x : c) out[k++] = x;
  } else {
    auto v = Checkerboard<4>::construct_by_color({shape[0], shape[1], shape[2], shape[3]});
    for (auto& c : v) for (auto x : c) out[k++] = x;
  }
}
// point_to_space_index with the origin fixed at -L/2 (test helper of the reference).
void oracle_point_to_space_fixed(int dim, double range, const double* edges, int periodic, const double* pts,
                                 int64_t* out, uint64_t n) {
  if (dim == 2) {
    Checkerboard<2> cb;
    cb.with_fixed_origin(range, {edges[0], edges[1]}, periodic != 0);
    for (uint64_t k = 0; k < n; ++k) out[k] = cb.point_to_space_index({pts[2 * k], pts[2 * k + 1]});
  } else {
    Checkerboard<3> cb;
    cb.with_fixed_origin(range, {edges[0], edges[1], edges[2]}, periodic != 0);
    for (uint64_t k = 0; k < n; ++k)
      out[k] = cb.point_to_space_index({pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]});
  }
}
// Checkerboard drawn exactly as ParallelSweep::update_checkerboard does.
void oracle_checkerboard_origin(int dim, uint64_t step, uint32_t substep, uint32_t seed, double range,
                                const double* edges, double* origin, double* width, int64_t* shape) {
  Sfc64 rng = counter_rng(step, substep, seed, UINT64_MAX);
  if (dim == 2) {
    Checkerboard<2> cb;
    cb.update(rng, range, {edges[0], edges[1]});
    for (int i = 0; i < 2; ++i) { origin[i] = cb.origin[i]; width[i] = cb.width[i]; shape[i] = cb.shape[i]; }
  } else {
    Checkerboard<3> cb;
    cb.update(rng, range, {edges[0], edges[1], edges[2]});
    for (int i = 0; i < 3; ++i) { origin[i] = cb.origin[i]; width[i] = cb.width[i]; shape[i] = cb.shape[i]; }
  }
}
// VecCell stencil of the given radius, flattened.
uint64_t oracle_stencil(int dim, uint32_t radius, int64_t* out) {
  size_t k = 0;
  if (dim == 2) {
    for (auto& c : VecCell<2>::generate_stencil(radius)) for (auto x : c) out[k++] = x;
  } else {
    for (auto& c : VecCell<3>::generate_stencil(radius)) for (auto x : c) out[k++] = x;
  }
  return k / dim;
}
void oracle_cell_index(int dim, double w, const double* p, int64_t* out) {
  if (dim == 2) {
    VecCell<2> vc(w);
    auto c = vc.cell_index_from_position({p[0], p[1]});
    out[0] = c[0]; out[1] = c[1];
  } else {
    VecCell<3> vc(w);
    auto c = vc.cell_index_from_position({p[0], p[1], p[2]});
    for (int i = 0; i < 3; ++i) out[i] = c[i];
  }
}

// ---- full systems ------------------------------------------------------------
// model: 0 HardSphere(params[0]=diameter), 1 LJ, 2 shifted LJ (params eps, sigma, r_cut, r_shift),
// 4 overlap penalty (params k, maximum_allowed_overlap, epsilon_shoulder, delta = r_cut)
void* oracle_system_create(int dim, const double* L, double boundary_range, double body_range, int model,
                           const double* params, uint64_t step, uint32_t seed, const double* pos, uint64_t n) {
  auto* h = new Handle;
  h->dim = dim;
  if (dim == 2)
    h->s2.reset(make<2>(L, boundary_range, body_range, model, params, step, seed, pos, n));
  else
    h->s3.reset(make<3>(L, boundary_range, body_range, model, params, step, seed, pos, n));
  if (!h->s2 && !h->s3) {
    delete h;
    return nullptr;
  }
  return h;
}
void oracle_system_destroy(void* p) { delete (Handle*)p; }
void oracle_set_threads(int n) { omp_set_num_threads(n); }
int oracle_max_threads() { return omp_get_max_threads(); }

#define DISPATCH(h, expr)         \
  do {                            \
    auto* H = (Handle*)(h);       \
    if (H->dim == 2) {            \
      auto& S = *H->s2;           \
      expr;                       \
    } else {                      \
      auto& S = *H->s3;           \
      expr;                       \
    }                             \
  } while (0)

int oracle_parallel_apply(void* h, double kt, double max_move, uint64_t* acc, uint64_t* rej) {
  int ok = 0;
  DISPATCH(h, ok = S.parallel_apply(kt, max_move, *acc, *rej) ? 0 : -1);
  return ok;
}
void oracle_sweep_apply(void* h, double kt, double max_move, uint64_t* acc, uint64_t* rej) {
  DISPATCH(h, S.sweep_apply(kt, max_move, *acc, *rej));
}
void oracle_increment_step(void* h) { DISPATCH(h, S.increment_step()); }

// Adds External(Harmonic) with spring constant k about centre c (dim values).
}  // extern "C"
template <int D>
static void set_external(System<D>& S, double k, const double* c) {
  S.ext_k = k;
  for (int i = 0; i < D; ++i) S.ext_c[i] = c[i];
}
extern "C" {
void oracle_set_external_harmonic(void* h, double k, const double* c) { DISPATCH(h, set_external(S, k, c)); }

// n steps of {Sweep (parallel = 0) or ParallelSweep apply; accumulate body 0's
// position and the external energy summed over sites; increment_step}, as the
// reference's harmonic-oscillator test does (hoomd-mc/src/sweep.rs:420-461).
}  // extern "C"
template <int D>
static void run_accumulate(System<D>& S, int parallel, double kt, double max_move, uint64_t n, double* pos_sum,
                           double* energy_sum) {
  for (int i = 0; i < D; ++i) pos_sum[i] = 0.0;
  *energy_sum = 0.0;
  for (uint64_t k = 0; k < n; ++k) {
    uint64_t a = 0, r = 0;
    if (parallel)
      S.parallel_apply(kt, max_move, a, r);
    else
      S.sweep_apply(kt, max_move, a, r);
    for (int i = 0; i < D; ++i) pos_sum[i] += S.bodies[0][i];
    double e = 0.0;
    for (const auto& st : S.sites.items) e += S.external(st.pos);
    *energy_sum += e;
    S.increment_step();
  }
}
extern "C" {
void oracle_run_accumulate(void* h, int parallel, double kt, double max_move, uint64_t n, double* pos_sum,
                           double* energy_sum) {
  DISPATCH(h, run_accumulate(S, parallel, kt, max_move, n, pos_sum, energy_sum));
}
void oracle_get_counter(void* h, uint64_t* step, uint32_t* substep) {
  DISPATCH(h, { *step = S.step; *substep = S.substep; });
}
void oracle_set_counter(void* h, uint64_t step, uint32_t substep) {
  DISPATCH(h, { S.step = step; S.substep = substep; });
}
uint64_t oracle_n_ghosts(void* h) {
  uint64_t n = 0;
  DISPATCH(h, n = S.ghosts.items.size());
  return n;
}
// Body positions (index order) and site positions (by body tag).
void oracle_get_positions(void* h, double* body, double* site) {
  auto* H = (Handle*)h;
  int D = H->dim;
  DISPATCH(h, {
    for (size_t b = 0; b < S.bodies.size(); ++b) {
      for (int i = 0; i < D; ++i) body[b * D + i] = S.bodies[b][i];
      const auto& st = S.sites.items[S.sites.indices[b]];
      for (int i = 0; i < D; ++i) site[b * D + i] = st.pos[i];
    }
  });
}
double oracle_total_energy(void* h) {
  double e = 0;
  DISPATCH(h, e = S.total_energy());
  return e;
}
// qi = {target, allowed_overlaps, inserted, complete}, updated in place
void oracle_quick_insert_apply(void* h, uint64_t* qi, uint64_t* acc, uint64_t* rej) {
  DISPATCH(h, S.quick_insert_apply(qi, *acc, *rej));
}
// tune_local with Translate: counts = 2 per step; *max_move updated
void oracle_tune_translate(void* h, double kt, double target, uint64_t samples, uint32_t steps, double* max_move,
                           uint64_t* counts) {
  DISPATCH(h, S.tune_translate(kt, target, samples, steps, *max_move, counts));
}
// GrandCanonical::apply; cnt = {insert acc, insert rej, remove acc, remove rej}, incremented
void oracle_grand_canonical_apply(void* h, double kt, double fugacity, uint64_t* cnt) {
  DISPATCH(h, S.grand_canonical_apply(kt, fugacity, cnt));
}
uint64_t oracle_n_bodies(void* h) {
  uint64_t n = 0;
  DISPATCH(h, n = S.bodies.size());
  return n;
}
double oracle_delta_energy_insert(void* h, const double* body) {
  auto* H = (Handle*)h;
  if (H->dim == 2) return H->s2->delta_energy_insert({body[0], body[1]});
  return H->s3->delta_energy_insert({body[0], body[1], body[2]});
}
double oracle_delta_energy_one(void* h, uint64_t body, const double* trial) {
  double e = 0;
  auto* H = (Handle*)h;
  if (H->dim == 2) {
    e = H->s2->delta_energy_one(body, {trial[0], trial[1]});
  } else {
    e = H->s3->delta_energy_one(body, {trial[0], trial[1], trial[2]});
  }
  return e;
}
void oracle_trace_enable(void* h, int on) {
  DISPATCH(h, { S.tracing = on != 0; S.trace.clear(); });
}
uint64_t oracle_trace_size(void* h) {
  uint64_t n = 0;
  DISPATCH(h, n = S.trace.size());
  return n;
}
void oracle_trace_read(void* h, TraceRecord* out) {
  DISPATCH(h, { std::memcpy(out, S.trace.data(), S.trace.size() * sizeof(TraceRecord)); S.trace.clear(); });
}
// Timed loop of ParallelSweep (parallel=1) or Sweep (parallel=0) applies,
// each followed by increment_step, for at least `seconds`.  Returns trial
// moves performed; *elapsed receives wall seconds.
uint64_t oracle_time_applies(void* h, int parallel, double kt, double max_move, double seconds, double* elapsed,
                             uint64_t* applies) {
  using clk = std::chrono::steady_clock;
  auto t0 = clk::now();
  uint64_t total = 0, n = 0;
  for (;;) {
    uint64_t a = 0, r = 0;
    if (parallel)
      oracle_parallel_apply(h, kt, max_move, &a, &r);
    else
      oracle_sweep_apply(h, kt, max_move, &a, &r);
    oracle_increment_step(h);
    total += a + r;
    ++n;
    double dt = std::chrono::duration<double>(clk::now() - t0).count();
    if (dt >= seconds) {
      *elapsed = dt;
      break;
    }
  }
  *applies = n;
  return total;
}


// ---- hard convex polyhedra (oracle/convex.hpp) -------------------------------
static std::vector<V3> read_verts(const double* v, uint32_t nv) {
  std::vector<V3> out(nv);
  for (uint32_t i = 0; i < nv; ++i) out[i] = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
  return out;
}
static Quat read_quat(const double* q) { return Quat{q[0], {q[1], q[2], q[3]}}; }

// collide3d on a polytope with itself at separation vij and relative
// orientation qij = (s, x, y, z)
int oracle_convex_collide(const double* verts, uint32_t nv, const double* vij, const double* qij) {
  Polytope p;
  p.set(read_verts(verts, nv));
  return collide3d(p, p, V3{vij[0], vij[1], vij[2]}, read_quat(qij)) ? 1 : 0;
}
int oracle_convex_intersects_global(const double* verts, uint32_t nv, const double* ri, const double* qi,
                                    const double* rj, const double* qj) {
  Polytope p;
  p.set(read_verts(verts, nv));
  return intersects_at_global(p, V3{ri[0], ri[1], ri[2]}, read_quat(qi), V3{rj[0], rj[1], rj[2]}, read_quat(qj))
             ? 1
             : 0;
}
// quaternion probes: op 0 product a*b, 1 rotate b[1..3] by a, 2 rotation-matrix rotate
void oracle_quat_op(int op, const double* a, const double* b, double* out) {
  Quat qa = read_quat(a);
  if (op == 0) {
    Quat r = qmul(qa, read_quat(b));
    out[0] = r.s;
    out[1] = r.v[0];
    out[2] = r.v[1];
    out[3] = r.v[2];
  } else {
    V3 x{b[1], b[2], b[3]};
    V3 r = op == 1 ? qrotate(qa, x) : mrotate(rotmat(qa), x);
    out[0] = 0.0;
    out[1] = r[0];
    out[2] = r[1];
    out[3] = r[2];
  }
}
void oracle_versor_displacement(uint64_t step, uint32_t substep, uint32_t seed, uint64_t index, double sigma,
                                double* out, uint64_t n) {
  Sfc64 rng = counter_rng(step, substep, seed, index);
  for (uint64_t i = 0; i < n; ++i) {
    Quat q = versor_displacement(rng, sigma);
    out[4 * i] = q.s;
    out[4 * i + 1] = q.v[0];
    out[4 * i + 2] = q.v[1];
    out[4 * i + 3] = q.v[2];
  }
}

void* oracle_convex_create(const double* L, double boundary_range, double body_range, const double* verts,
                           uint32_t nv, uint64_t step, uint32_t seed, const double* pos, const double* quat,
                           uint64_t n) {
  auto* s = new ConvexSystem({L[0], L[1], L[2]}, boundary_range, read_verts(verts, nv));
  s->body_range = body_range;
  s->step = step;
  s->seed = seed;
  if (!s->boundary.valid() || s->range() > boundary_range) {
    delete s;
    return nullptr;
  }
  for (uint64_t b = 0; b < n; ++b) s->add_body(V3{pos[3 * b], pos[3 * b + 1], pos[3 * b + 2]}, read_quat(quat + 4 * b));
  return s;
}
void oracle_convex_destroy(void* h) { delete (ConvexSystem*)h; }
// tune_local with Translate (rotate 0) or Rotate<Versor> (1); counts = 2 per step
void oracle_convex_tune(void* h, int rotate, double target, uint64_t samples, uint32_t steps, double* size,
                        uint64_t* counts) {
  ((ConvexSystem*)h)->tune(rotate != 0, target, samples, steps, *size, counts);
}
int oracle_convex_apply(void* h, int rotate, double max_move, uint64_t* acc, uint64_t* rej) {
  return ((ConvexSystem*)h)->parallel_apply(rotate != 0, max_move, *acc, *rej) ? 0 : -1;
}
void oracle_convex_increment_step(void* h) { ((ConvexSystem*)h)->increment_step(); }
void oracle_convex_get_counter(void* h, uint64_t* step, uint32_t* substep) {
  *step = ((ConvexSystem*)h)->step;
  *substep = ((ConvexSystem*)h)->substep;
}
// bodies and sites in body order: positions n x 3, orientations n x 4
void oracle_convex_state(void* h, double* bpos, double* bq, double* spos, double* sq) {
  auto* S = (ConvexSystem*)h;
  for (size_t b = 0; b < S->bpos.size(); ++b) {
    const CSite& st = S->sites.items[S->sites.indices[b]];
    for (int i = 0; i < 3; ++i) {
      bpos[3 * b + i] = S->bpos[b][i];
      spos[3 * b + i] = st.pos[i];
    }
    bq[4 * b] = S->bq[b].s;
    sq[4 * b] = st.q.s;
    for (int i = 0; i < 3; ++i) {
      bq[4 * b + 1 + i] = S->bq[b].v[i];
      sq[4 * b + 1 + i] = st.q.v[i];
    }
  }
}
int oracle_convex_any_overlap(void* h) { return ((ConvexSystem*)h)->any_overlap() ? 1 : 0; }
void oracle_convex_trace_enable(void* h, int on) {
  ((ConvexSystem*)h)->tracing = on != 0;
  ((ConvexSystem*)h)->trace.clear();
}
uint64_t oracle_convex_trace_size(void* h) { return ((ConvexSystem*)h)->trace.size(); }
void oracle_convex_trace_read(void* h, void* out) {
  auto* S = (ConvexSystem*)h;
  std::memcpy(out, S->trace.data(), S->trace.size() * sizeof(TraceRecord));
  S->trace.clear();
}
// Timed loop of {translate apply; rotate apply; increment_step}, as the
// reference's convex-shape benchmarks advance (benchmarks/src/mc/octahedron.rs:115-127).
uint64_t oracle_convex_time(void* h, double d, double a, double seconds, double* elapsed, uint64_t* steps) {
  using clk = std::chrono::steady_clock;
  auto* S = (ConvexSystem*)h;
  auto t0 = clk::now();
  uint64_t total = 0, n = 0;
  for (;;) {
    uint64_t x = 0, y = 0;
    S->parallel_apply(false, d, x, y);
    total += x + y;
    S->parallel_apply(true, a, x, y);
    total += x + y;
    S->increment_step();
    ++n;
    double dt = std::chrono::duration<double>(clk::now() - t0).count();
    if (dt >= seconds) {
      *elapsed = dt;
      break;
    }
  }
  *steps = n;
  return total;
}

// ---- scale distribution function (oracle/sdf.hpp) ----------------------------
int oracle_sdf(int dim, const double* L, const double* pos, uint64_t n, double sigma, double xmax, double dx,
               uint32_t nbins, int all_pairs, uint64_t* hist, uint64_t* overlaps) {
  if ((dim != 2 && dim != 3) || !(sigma > 0.0) || !(xmax > 0.0 && xmax < 1.0) || !(dx > 0.0)) return -1;
  if (dim == 2)
    sdf_histogram<2>(pos, n, L, sigma, xmax, dx, nbins, all_pairs != 0, hist, overlaps);
  else
    sdf_histogram<3>(pos, n, L, sigma, xmax, dx, nbins, all_pairs != 0, hist, overlaps);
  return 0;
}

// Hypercuboid::map_point (hoomd-geometry/src/shape/cuboid.rs:553-567) of n
// points from edges Lold to Lnew: p' = clamp(p (Lnew / Lold), -Lnew/2,
// next_down(Lnew/2)) per axis.  A point outside the old box (the
// PointOutsideShape error; is_point_inside, cuboid.rs:449-454) is counted and
// left unmapped.  Returns the number of such points.
uint64_t oracle_map_points(int dim, const double* Lold, const double* Lnew, const double* pts, uint64_t n,
                           double* out) {
  uint64_t outside = 0;
  for (uint64_t i = 0; i < n; ++i) {
    bool inside = true;
    for (int a = 0; a < dim; ++a) {
      const double x = pts[i * dim + a];
      inside = inside && -Lold[a] / 2.0 <= x && x < Lold[a] / 2.0;
    }
    for (int a = 0; a < dim; ++a) {
      const double x = pts[i * dim + a];
      if (!inside) {
        out[i * dim + a] = x;
        continue;
      }
      const double scale = Lnew[a] / Lold[a];
      const double lo = -Lnew[a] / 2.0, hi = std::nextafter(Lnew[a] / 2.0, -INF);
      const double v = scale * x;
      out[i * dim + a] = v < lo ? lo : (v > hi ? hi : v);  // f64::clamp
    }
    outside += inside ? 0 : 1;
  }
  return outside;
}

}  // extern "C"
