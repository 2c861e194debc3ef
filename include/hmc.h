/* hmc.h -- C ABI of libhmc.so, the MI355X (gfx950) implementation of
 * hoomd-mc's checkerboard-parallel local trial-move sweep.
 *
 * The library replaces, for built-in shapes and interactions only, the body of
 *   impl Trial for ParallelSweep<Translate<Cartesian<D>>, HypercuboidCheckerboard<D>, Point, Point>
 *   (hoomd-mc/src/parallel_sweep.rs:317-421)
 * acting on a Microstate with a Periodic<Hypercuboid<D>> boundary
 * (hoomd-microstate/src/boundary/periodic/cuboid.rs) and a
 * PairwiseCutoff<HardSphere> Hamiltonian
 * (hoomd-interaction/src/pairwise/hard_shape.rs:127-174).
 * Everything else keeps using the reference's CPU path.
 *
 * Conventions: every function returns an int status (HMC_OK == 0, negative on
 * error) and never throws; positions are row-major (n_bodies x dim) doubles;
 * a context is single-threaded (one caller at a time, like &mut self).
 */
#ifndef HMC_H
#define HMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hmc_ctx hmc_ctx;

/* hoomd_mc::Count (hoomd-mc/src/lib.rs:205-233). */
typedef struct {
  uint64_t accepted;
  uint64_t rejected;
} hmc_count;

enum {
  HMC_OK = 0,
  HMC_ERR_INVALID = -1,     /* bad argument or configuration */
  HMC_ERR_UNSUPPORTED = -2, /* combination not offloaded: keep the CPU path */
  HMC_ERR_DEVICE = -3,      /* HIP runtime error */
  HMC_ERR_COVER = -4,       /* a body is outside the checkerboard (the reference panics) */
  HMC_ERR_STATE = -5        /* no state set, or trace buffer overflow */
};

/* Pair model (the Hamiltonian H of PairwiseCutoff<H>). */
enum {
  HMC_HARD_SPHERE = 0, /* HardSphere (pairwise/hard_shape.rs:127-174): params[0] = diameter */
  HMC_LJ = 1,          /* Isotropic<LennardJones<12,6>> (isotropic.rs, univariate/lennard_jones.rs):
                          params = {epsilon, sigma, r_cut, -} */
  HMC_LJ_SHIFTED = 2,  /* Isotropic<Shifted<LennardJones<12,6>>> (univariate/shifted.rs):
                          params = {epsilon, sigma, r_cut, r_shift} */
  HMC_CONVEX_POLYHEDRON = 3, /* HardShape<Convex<ConvexPolyhedron>> (pairwise/hard_shape.rs:79-118,
                                hoomd-geometry/src/convex.rs:146-170) on OrientedPoint<Cartesian<3>,
                                Versor> bodies: hmc_config.vertices; dim 3, HMC_FP64 only */
  HMC_OVERLAP_PENALTY = 4    /* Isotropic { Expanded { delta, OverlapPenalty }, r_cut: delta }
                                (univariate/overlap_penalty.rs:109-115, univariate/expanded.rs:62-66),
                                the placement Hamiltonian of benchmarks/src/mc/hard_sphere.rs:192-198:
                                params = {k, maximum_allowed_overlap, epsilon_shoulder, delta};
                                HMC_FP64 only */
};

/* Arithmetic of the pair test. Both modes make identical decisions. */
enum {
  HMC_FP64 = 0,  /* every pair in FP64, exactly as the reference */
  HMC_MIXED = 1  /* hard spheres: FP32 overlap test, pairs inside its rigorous
                    error band re-tested in FP64.  LJ: FP32 pair energies
                    (FP64 displacements and cutoff test), the trial's
                    Metropolis decision re-taken in FP64 when the FP32 error
                    bound on dH could change it.  Convex polyhedra: FP32
                    separating-axis tests (face normals, edge-pair cross
                    products, vertex containment) with a margin of 2^-12 R;
                    pairs inside it run XenoCollide in FP64 */
};

typedef struct {
  uint32_t dim;                  /* 2 or 3 (Cartesian<D>) */
  uint64_t n_bodies;             /* single-site point bodies */
  double box_edges[3];           /* Hypercuboid edge lengths (first dim used) */
  double boundary_range;         /* Periodic::new(maximum_interaction_range, ..)
                                    (boundary/periodic.rs:96-110): ghost criterion */
  double body_interaction_range; /* ParallelSweep::new(body_interaction_range, ..)
                                    (parallel_sweep.rs:138-146): checkerboard width */
  double spatial_cell_width;     /* cell width of the microstate's VecCell
                                    (hoomd-spatial/src/vec_cell.rs:286-299); decides
                                    which images are visited at exact contact */
  int32_t model;                 /* HMC_HARD_SPHERE, HMC_LJ, HMC_LJ_SHIFTED, HMC_CONVEX_POLYHEDRON
                                    or HMC_OVERLAP_PENALTY */
  double params[4];
  int32_t precision;             /* HMC_FP64 or HMC_MIXED */
  int32_t device;                /* HIP device ordinal */
  uint32_t flags;                /* HMC_FLAG_* (0: defaults) */
  const double* vertices;        /* HMC_CONVEX_POLYHEDRON: n_vertices x 3 vertex coordinates
                                    (ConvexPolytope<3, 32>, shape/convex_polytope.rs); read at
                                    hmc_create only */
  uint32_t n_vertices;           /* 1..32 */
} hmc_config;

/* hmc_config.flags */
enum {
  HMC_FLAG_GLOBAL_PASS = 1, /* run every trial from global memory instead of
                               staging each workgroup's neighbourhood box in LDS
                               (same decisions; for comparison and testing) */
  HMC_FLAG_NO_GRAPH = 2,    /* hmc_run_translate launches every apply directly
                               instead of replaying a captured hipGraph
                               (same results) */
  HMC_FLAG_PER_COLOUR = 4,  /* MIXED hard shapes: one launch per colour pass
                               instead of the default single persistent launch
                               per apply (same results) */
  HMC_FLAG_FP32_ERROR = 8   /* MIXED: also measure in FP64 and record the
                               largest FP32 error and its ratio to the error
                               bound (hmc_stats; same decisions, slower).
                               LJ: every trial's dH; hard spheres: every pair
                               within 1.01 sigma (runs the per-colour tile
                               passes) */
};

/* One trial of a trace: replaces the per-space BodyTrial status of
 * parallel_sweep.rs:26-52 (status 0 Rejected, 1 Accepted). */
typedef struct {
  uint32_t substep;
  uint32_t color;
  uint64_t space;
  int64_t body;
  int32_t status;
  int32_t pad;
} hmc_trace_record;

typedef struct {
  uint64_t applies;         /* apply calls since creation */
  uint64_t cycles;          /* colour cycles (= substeps consumed) */
  uint64_t pair_tests;      /* not maintained (0) */
  uint64_t band_retests;    /* MIXED: hard-sphere pairs, or LJ trials, re-tested in FP64 */
  uint64_t guard_cycles;    /* cycles run by the single-workgroup overflow guard */
  uint64_t tile_overflows;  /* LDS tiles whose box exceeded the staging capacity
                               (their trials ran from global memory) */
  uint32_t tiled;           /* 1: one launch per colour pass, staging each
                               workgroup's neighbourhood box in LDS; 2: one
                               persistent launch per apply, each workgroup
                               keeping its region and halo in LDS; 0: no LDS */
  uint32_t tile_shape[3];   /* colour spaces per workgroup along each axis */
  uint64_t ambiguous;       /* LJ: Metropolis decisions that the bound on the
                               summation-order difference could not settle
                               (decided from the GPU's own sum; 0 means every
                               decision provably equals the reference's) */
  double fp32_error_max;    /* HMC_FLAG_FP32_ERROR: largest |dH32 - dH64| of an LJ
                               trial, or |d2_32 - d2_64| of a hard-sphere pair */
  double fp32_error_bound;  /* HMC_FLAG_FP32_ERROR: the largest such error divided
                               by its bound (at most 1 when the bound holds) */
  uint32_t regions;         /* persistent apply: regions (workgroups that take part) */
  uint32_t region_threads;  /* persistent apply: lanes per region */
  uint32_t regions_per_cu;  /* persistent apply: resident workgroups per CU (occupancy) */
  uint32_t wait_timeout_at; /* first neighbour wait that timed out: (region << 12) | (pass + 1), or 0 */
  uint32_t rounds;          /* persistent batch apply: rounds of regions per step (0 otherwise) */
} hmc_stats;

/* Allocate device state for one system.  Replaces building a ParallelSweep and
 * a device mirror of the Microstate.  HMC_ERR_UNSUPPORTED for anything the
 * GPU path does not cover (the caller keeps the CPU sweep). */
int hmc_create(const hmc_config* cfg, hmc_ctx** out);

/* Upload body positions (the microstate's wrapped body properties, body index
 * order).  site_pos may be NULL: sites are then wrap(body + 0) as
 * Microstate::add_body / update_body_properties set them
 * (hoomd-microstate/src/microstate.rs:777-835, 985-1034). */
int hmc_set_state(hmc_ctx* ctx, const double* body_pos, const double* site_pos);

/* Download body and site positions in body index order (site_pos may be NULL). */
int hmc_get_state(hmc_ctx* ctx, double* body_pos, double* site_pos);

/* Microstate::clone_with_boundary(new_boundary, |_| true)
 * (hoomd-microstate/src/microstate.rs:1529-1565) for a Periodic<Hypercuboid<D>>
 * with new edge lengths, such as Periodic::scale_volume makes
 * (boundary/periodic.rs:288-300, hoomd-geometry/src/shape/cuboid.rs:514-518):
 * a new context with the same configuration on new_edges (first dim used),
 * each body mapped by Hypercuboid::map_point (cuboid.rs:553-567) and rebuilt by
 * update_body_properties (microstate.rs:985-1034).  ctx is left unchanged, as
 * QuickCompress keeps the initial microstate (hoomd-mc/src/quick_compress.rs,
 * Compressing state).  HMC_ERR_INVALID when the ranges do not fit the new
 * edges or a body lies outside the old box; free the new context with
 * hmc_destroy. */
int hmc_clone_with_boundary(hmc_ctx* ctx, const double* new_edges, hmc_ctx** out);

/* QuickCompress (hoomd-mc/src/quick_compress.rs:19-444): compress (or expand)
 * the box toward a target volume in steps the Hamiltonian tolerates, waiting
 * for the sweeps in between to remove the overlaps (total energy <= 0).  The
 * state machine runs on the host; the trial microstate is
 * hmc_clone_with_boundary and the energies hmc_total_energy. */
enum { HMC_QC_STARTUP = 0, HMC_QC_WAITING = 1, HMC_QC_COMPRESSING = 2, HMC_QC_COMPLETE = 3 };
typedef struct {
  double target_volume;           /* QuickCompress::with_target_volume */
  double maximum_energy_per_site; /* default 1000 */
  double maximum_delta;           /* default 0.05, in (0, 1) */
  int32_t state;                  /* HMC_QC_* */
  int32_t is_final;               /* Waiting: the last compression reached the target */
  double boundary[3];             /* Waiting: last known edges; Complete: final edges */
  uint64_t ambiguous;             /* energy comparisons inside the summation-order bound
                                     (decided by the GPU value; expected 0) */
} hmc_quick_compress;

/* QuickCompress::with_target_volume: defaults, Startup state. */
void hmc_quick_compress_init(hmc_quick_compress* qc, double target_volume);

/* QuickCompress::apply(microstate, hamiltonian, |_| true) with (step,
 * substep, seed) = microstate.counter(); it consumes one substep (the
 * caller's counter then calls increment_substep).  When a compression is
 * accepted, *ctx is replaced by the trial context and the old one destroyed.
 * HMC_ERR_INVALID when a trial box is too small for the interaction ranges
 * (the reference panics in Periodic::scale_volume). */
int hmc_quick_compress_apply(hmc_quick_compress* qc, hmc_ctx** ctx, uint64_t step, uint32_t substep,
                             uint32_t seed);

/* QuickInsert (hoomd-mc/src/quick_insert.rs:94-334) with a UniformIn
 * distribution over the periodic box (hoomd-mc/src/uniform_in.rs:99-110):
 * while the total energy is <= 0, insert bodies at uniformly random positions,
 * rejecting the insertions of infinite energy, until `target` bodies are in or
 * `allowed_overlaps` insertions of positive energy were made in one call.
 * Point bodies only (spheres, disks, LJ, OverlapPenalty). */
enum { HMC_QI_RUNNING = 0, HMC_QI_COMPLETE = 1 };
typedef struct {
  uint64_t target;           /* bodies to insert (QuickInsert::new) */
  uint64_t allowed_overlaps; /* max(target / 8, 1) */
  uint64_t inserted;         /* insertions so far */
  int32_t state;             /* HMC_QI_* */
  int32_t pad;
  uint64_t ambiguous;        /* sign tests of an energy, and candidate pairs within 2^-30 of the
                                cutoff (VecCell's edge rule), inside their rounding bound:
                                decided by the computed value; expected 0 */
} hmc_quick_insert;

/* QuickInsert::new(distribution, target): Running, nothing inserted. */
void hmc_quick_insert_init(hmc_quick_insert* qi, uint64_t target);

/* QuickInsert::apply(microstate, hamiltonian) with (step, substep, seed) =
 * microstate.counter().  Every candidate of the call is drawn on the host from
 * Counter(step, substep, seed).make_rng() (one Uniform(-L/2, L/2) per axis),
 * its insertion energy against the current bodies is computed on the device
 * (k_insert_energy, all candidates in one launch), and the candidates are then
 * taken in order on the host, each against the bodies inserted before it in
 * the same call.  When bodies were inserted, *ctx is replaced by a context
 * holding the old bodies followed by the new ones (Microstate::add_body order)
 * and the old context destroyed.  *count gets the inserted (accepted) and
 * rejected candidates.  When qi->state is not HMC_QI_COMPLETE after the call,
 * the caller then calls increment_substep, as the reference's apply does (a
 * call that finds the work complete returns before it). */
int hmc_quick_insert_apply(hmc_quick_insert* qi, hmc_ctx** ctx, uint64_t step, uint32_t substep, uint32_t seed,
                           hmc_count* count);

/* ParallelSweep::tune / Sweep::tune (hoomd-mc/src/tune_local.rs:21-89,
 * parallel_sweep.rs:486-506) with Translate (move HMC_MOVE_TRANSLATE) or,
 * for convex polyhedra, Rotate<Versor> (HMC_MOVE_ROTATE): `steps` rounds of
 * `samples` trial moves of random bodies against the unchanged microstate,
 * from one stream Counter(step, substep, seed).make_rng(); after each round
 * tune_by_scaling (lib.rs:514-524) multiplies *maximum_size by
 * (ratio + 1.5) / (target + 1.5) (a rotation size at most pi/2,
 * rotate/versor.rs:125-137).  The state is not modified and no substep is
 * consumed.  counts (optional, `steps` entries) get each round's accepted
 * and rejected samples; *ambiguous (optional) the samples whose draw or
 * decision the bound on the summation order could not fix (decided by the
 * GPU's dH; expected 0).  Every sample of a round is evaluated on the device
 * at once: the round's stream is drawn on the host, one lane takes a sample
 * from every offset of it (k_tune_eval, k_tune_eval_convex), and the samples
 * are the chain of offsets from 0. */
int hmc_tune_local(hmc_ctx* ctx, uint32_t move, uint64_t step, uint32_t substep, uint32_t seed, double kT,
                   double target, uint64_t samples, uint32_t steps, double* maximum_size, hmc_count* counts,
                   uint64_t* ambiguous);

/* GrandCanonical(UniformIn over the periodic box)::apply
 * (hoomd-mc/src/grand_canonical.rs:145-205): one insertion or removal trial,
 * chosen by a bool draw of Counter(step, substep, seed).make_rng().  An
 * insertion at a uniform position is accepted when
 * V f exp(-dE/kT) / (n + 1) > u, a removal of body random_range(0..n) when
 * n exp(-dE/kT) / (V f) > u (1 > u when f <= 0); the insertion and removal
 * energies come from the device.  An accepted move replaces *ctx by a
 * context of the new body list (add_body appends; remove_body moves the last
 * body to the removed index) and destroys the old one.  The caller then
 * increments the substep, as the reference's apply does.  *ambiguous
 * (optional): 1 when the bound on the summation order could flip the
 * decision (decided by the GPU's dE).  Point bodies only. */
typedef struct {
  uint64_t insert_accepted, insert_rejected, remove_accepted, remove_rejected;
} hmc_gc_count;
int hmc_grand_canonical_apply(hmc_ctx** ctx, uint64_t step, uint32_t substep, uint32_t seed, double kT,
                              double fugacity, hmc_gc_count* count, uint64_t* ambiguous);

/* One ParallelSweep::apply with Translate moves (parallel_sweep.rs:390-421).
 * (step, substep, seed) = microstate.counter(); kT = macrostate.temperature();
 * max_distance = Translate::maximum_distance.  On return *substeps_used holds
 * the number of colour cycles: the caller increments the microstate substep by
 * that much, as the reference's apply does. */
int hmc_apply_translate(hmc_ctx* ctx, uint64_t step, uint32_t substep, uint32_t seed, double kT,
                        double max_distance, hmc_count* out, uint32_t* substeps_used);

/* One ParallelSweep::apply with Rotate<Versor> moves (rotate/versor.rs:111-122;
 * parallel_sweep.rs:390-421) on oriented bodies; max_rotation =
 * Rotate::maximum_rotation.  Same conventions as hmc_apply_translate. */
int hmc_apply_rotate(hmc_ctx* ctx, uint64_t step, uint32_t substep, uint32_t seed, double kT,
                     double max_rotation, hmc_count* out, uint32_t* substeps_used);

/* Orientations of oriented bodies, (s, x, y, z) per body in body index order
 * (the microstate's stored Versors).  site_quat may be NULL: sites are then
 * body * identity as OrientedPoint::transform sets them
 * (property/oriented_point.rs:107-113).  Bodies start with the identity. */
int hmc_set_orientations(hmc_ctx* ctx, const double* body_quat, const double* site_quat);
int hmc_get_orientations(hmc_ctx* ctx, double* body_quat, double* site_quat);

/* n_steps iterations of { ParallelSweep::apply; Microstate::increment_step }
 * (the body of HardSphereSim::advance, benchmarks/src/mc/hard_sphere.rs:90-111)
 * starting at (step0, substep 0), with no host synchronisation between
 * applies.  With one launch per colour pass, the launches of one step after
 * the first are replayed from a captured hipGraph (HMC_FLAG_NO_GRAPH:
 * launched directly); the persistent apply is always launched directly.
 * *out receives the summed count. */
int hmc_run_translate(hmc_ctx* ctx, uint64_t step0, uint32_t seed, double kT, double max_distance,
                      uint32_t n_steps, hmc_count* out);

/* One trial move of a step: Translate (size = maximum_distance) or
 * Rotate<Versor> (size = maximum_rotation, oriented bodies only). */
enum { HMC_MOVE_TRANSLATE = 0, HMC_MOVE_ROTATE = 1 };
typedef struct {
  int32_t move; /* HMC_MOVE_* */
  int32_t pad;
  double size;
} hmc_move;

/* n_steps iterations of { for each move in order: ParallelSweep::apply of
 * that move (the step's substeps follow on from one apply to the next);
 * Microstate::increment_step } starting at (step0, substep 0), as a
 * Simulation's advance applies several trials per step (e.g. translate then
 * rotate, benchmarks/src/mc/octahedron.rs:109-134), with no host
 * synchronisation.  1 to 4 moves.  out[m] (n_moves entries, may be NULL)
 * receives the summed count of move m, as the reference keeps one Count per
 * trial.  Graph replay and direct launches as hmc_run_translate, which is
 * hmc_run with one Translate move. */
int hmc_run(hmc_ctx* ctx, uint64_t step0, uint32_t seed, double kT, const hmc_move* moves, uint32_t n_moves,
            uint32_t n_steps, hmc_count* out);

/* The side effects of Trial::apply on the caller's microstate
 * (Microstate::update_body_properties of every accepted body,
 * hoomd-microstate/src/microstate.rs:985-1056, called by
 * ParallelSweep::update_bodies, parallel_sweep.rs:283-314), synced lazily:
 * every accepted move stamps its body with the apply's epoch (applies since
 * hmc_create; hmc_set_state clears the stamps).  hmc_moved_since sets bit
 * i % 64 of mask[i / 64] (ceil(n_bodies / 64) words) for every body moved by
 * an apply after epoch `since`, *n_moved to their number and *now (may be
 * NULL) to the current epoch, the `since` of the next sync.
 * hmc_get_bodies copies the current body and site positions (n x dim) and,
 * for oriented bodies, orientations (n x 4) of the n listed bodies; any
 * output may be NULL.  The caller then runs update_body_properties on those
 * bodies only. */
int hmc_moved_since(hmc_ctx* ctx, uint32_t since, uint32_t* now, uint64_t* mask, uint64_t* n_moved);
int hmc_get_bodies(hmc_ctx* ctx, const uint32_t* index, uint64_t n, double* body_pos, double* site_pos,
                   double* body_quat, double* site_quat);

/* PairwiseCutoff::total_energy (hoomd-interaction/src/pairwise_cutoff.rs:720-745)
 * of the current state: +inf on any hard overlap.  The reference sums in its
 * VecCell order, the device in a fixed order of its own; *bound (may be NULL)
 * receives 2 gamma_{n-1} sum|terms|, which bounds the difference between any
 * two summation orders of the same n terms. */
int hmc_total_energy(hmc_ctx* ctx, double* energy, double* bound);

/* hmc_total_energy without allocating: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same kernels and
 * the same *energy and *bound, bit for bit; the block partials live in a
 * static device array of the library (1 MiB per GPU, held until the process
 * ends) instead of a per-call allocation.  HMC_ERR_INVALID above 8,388,608
 * bodies.  Calls that overlap on one GPU (several contexts on several
 * threads or streams) race on that array. */
int hmc_total_energy_noalloc(hmc_ctx* ctx, double* energy, double* bound);

/* PairwiseCutoff::delta_energy_insert (hoomd-interaction/src/pairwise_cutoff.rs:1017-1025,
 * with :420-488) of n candidate bodies (n x dim positions), each evaluated
 * alone against the current microstate: the energy of the site wrap(body + 0)
 * with every site, +inf on an overlap or an infinite term.  bound (may be
 * NULL) receives, per candidate, a bound on the difference from the
 * reference's summation order (0 when infinite); two calls may differ in the
 * last bits within it.  The microstate is unchanged.  HMC_ERR_UNSUPPORTED for
 * convex polyhedra (oriented bodies). */
int hmc_delta_energy_insert(hmc_ctx* ctx, const double* body_pos, uint64_t n, double* energy, double* bound);

/* Scale distribution function (SDF) of the current state, hard spheres and
 * disks: for every body the smallest compression x = 1 - sigma / r over the
 * other bodies (r the minimum-image site distance) at which the isotropically
 * scaled system would first overlap.  hist[k] (k < nbins) counts the bodies
 * with 0 <= x < xmax and floor(x / dx) = k; *overlaps (may be NULL) counts
 * those with x < 0.  Needs 0 < xmax < 1/2 and box edges above
 * 2 sigma / (1 - xmax); HMC_ERR_UNSUPPORTED for other models.  No counterpart
 * in the reference, which has no pressure compute: the definition is
 * HOOMD-blue's hpmc compute.SDF, from which beta P / rho = 1 + s(0+) / (2 d)
 * with s = hist / (N dx) extrapolated to x = 0. */
int hmc_sdf(hmc_ctx* ctx, double xmax, double dx, uint32_t nbins, uint64_t* hist, uint64_t* overlaps);

/* Record every non-empty trial of the following applies (testing/replay). */
int hmc_trace_enable(hmc_ctx* ctx, int on, uint64_t capacity);
int hmc_trace_read(hmc_ctx* ctx, hmc_trace_record* buf, uint64_t cap, uint64_t* n);

int hmc_get_stats(hmc_ctx* ctx, hmc_stats* out);

/* Time the build and pass launches with HIP events on the library's stream:
 * period <= 0 turns timing off; period k >= 1 times every k-th apply (a small
 * k adds event records between launches).  hmc_profile_read returns the
 * summed durations (ms) and launch counts since the last enable. */
int hmc_profile_enable(hmc_ctx* ctx, int period);
int hmc_profile_read(hmc_ctx* ctx, double* pass_ms_total, uint64_t* pass_launches, double* build_ms_total,
                     uint64_t* builds);

/* Test hook: the mailbox words of every region of the persistent apply, six
 * per region (the two pass-parity counters, the residency probe's counter in
 * region 0, then the isotropic apply's progress: staged, passes published,
 * waits passed). No reference counterpart. */
int hmc_debug_regions(hmc_ctx* ctx, uint32_t* out, uint64_t cap, uint64_t* n);

/* Evaluate the library's device math on n inputs: kind 0 cbrt, 1 log,
 * 2 exp, 5 the sine and 6 the cosine of one sincos call (the device's
 * restatements of the glibc functions Rust's std calls on Linux), 3 sqrt,
 * 4 division in[2i]/in[2i+1].  Test hook for bit-exactness checks of the
 * device arithmetic against the host C library. */
int hmc_debug_math(int device, int kind, const double* in, double* out, uint64_t n);

/* Evaluate the device convex-polyhedron routines on n inputs: kind 0
 * collide3d of the polytope with itself (in: v_ij[3], q_ij[4]; out 0/1),
 * kind 1 the global pair test (in: r_i[3], q_i[4], r_j[3], q_j[4]; out 0/1),
 * kind 2 the Rotate<Versor> displacement (in: step, substep, seed, index,
 * sigma; out: q[4]), kind 3 the mixed mode's FP32 separating-axis class of
 * a pair (in as for kind 1; out 0 apart, 1 overlap, 2 left to FP64).
 * Test hook. */
int hmc_debug_convex(int device, int kind, const double* vertices, uint32_t n_vertices, const double* in,
                     double* out, uint64_t n);

/* The mixed mode's separating-axis data of a polytope (host only):
 * counts = {face axes, edge directions, scanned vertices, symmetric,
 * FP32 tests used, edge axes used}; faces (16 x 5: unit normal, hmin, hmax)
 * and edges (8 x 3: unit direction) in FP32 values, either may be NULL.
 * Test hook; no reference counterpart. */
int hmc_debug_convex_axes(const double* vertices, uint32_t n_vertices, uint32_t* counts, double* faces,
                          double* edges);

/* Draw n standard normals (ziggurat) from stream (step, substep, seed, index). */
int hmc_debug_normals(int device, uint64_t step, uint32_t substep, uint32_t seed, uint64_t index, double* out,
                      uint64_t n);

const char* hmc_last_error(const hmc_ctx* ctx);
void hmc_destroy(hmc_ctx* ctx);

/* ---- GSD trajectory frames -------------------------------------------------
 * The GSD 2.1 file layer (hoomd-gsd/src/file_layer.rs: a 256-byte header,
 * an index of 32-byte entries sorted by (frame, id), a list of 64-byte chunk
 * names, and the chunk data) and the HOOMD schema frames of
 * hoomd-gsd/src/hoomd.rs.  Host code, except that hmc_gsd_append_frame packs
 * a context's sites to FP32 on the device before copying them back.  Every
 * frame is written through to the file (index and names included) when it
 * ends. */
typedef struct hmc_gsd hmc_gsd;
enum {
  HMC_GSD_U8 = 1, HMC_GSD_U16 = 2, HMC_GSD_U32 = 3, HMC_GSD_U64 = 4, HMC_GSD_I8 = 5, HMC_GSD_I16 = 6,
  HMC_GSD_I32 = 7, HMC_GSD_I64 = 8, HMC_GSD_F32 = 9, HMC_GSD_F64 = 10, HMC_GSD_STRING = 11
};

/* GsdFile::create (file_layer.rs:1008-1030): a new empty file (truncating),
 * with the application and schema names (< 64 bytes each) and the schema
 * version (major, minor); schema "hoomd" (1, 4) is HoomdGsdFile::create
 * (hoomd.rs:186-198).  HMC_ERR_INVALID for names too long or I/O errors. */
int hmc_gsd_create(const char* path, const char* application, const char* schema, uint32_t schema_major,
                   uint32_t schema_minor, hmc_gsd** out);

/* GsdFile::open (file_layer.rs:966-980): an existing file, for reading
 * (writable = 0) or for appending frames (writable = 1). */
int hmc_gsd_open(const char* path, int writable, hmc_gsd** out);

/* GsdFile::write_arrays / write_scalars (file_layer.rs): one chunk of `rows`
 * x `columns` values of GSD type `type` in the current frame; a name may
 * appear once per frame. */
int hmc_gsd_write_chunk(hmc_gsd* f, const char* name, uint8_t type, uint64_t rows, uint32_t columns,
                        const void* data);

/* GsdFile::end_frame (file_layer.rs:2043-2053) followed by sync_all
 * (file_layer.rs:2269-2330): the frame's data, new names and index entries
 * are written to the file. */
int hmc_gsd_end_frame(hmc_gsd* f);

/* HoomdGsdFile::append_frame(step) with configuration/box, configuration/
 * dimensions, particles/position (the sites in tag order as FP32, z = 0 in
 * 2D) and, for oriented bodies, particles/orientation, then Frame::end
 * (hoomd.rs:343-1013), from a context's current state. */
int hmc_gsd_append_frame(hmc_gsd* f, hmc_ctx* ctx, uint64_t step);

/* Frames in the file (GsdFile::n_frames). */
uint64_t hmc_gsd_n_frames(const hmc_gsd* f);

/* GsdFile::read of chunk `name` in `frame`: its rows, columns and type, and
 * its bytes copied to `out` when capacity_bytes holds them (out may be null
 * to query the shape).  HMC_ERR_INVALID when the chunk does not exist. */
int hmc_gsd_read_chunk(const hmc_gsd* f, uint64_t frame, const char* name, void* out, uint64_t capacity_bytes,
                       uint64_t* rows, uint32_t* columns, uint8_t* type);

/* The header's application, schema and schema version. */
int hmc_gsd_header(const hmc_gsd* f, char* application, char* schema, uint32_t* schema_major,
                   uint32_t* schema_minor);

/* Closes the file (a frame in progress is ended first, as Frame's Drop does). */
int hmc_gsd_close(hmc_gsd* f);

/* ---- replica batches ------------------------------------------------------
 * Many independent hard-sphere/disk systems (own box, checkerboard, counter,
 * seed and maximum distance each; same dim and precision) advanced together:
 * every step of every replica is one set of launches, captured once in a
 * hipGraph and replayed.  Each replica's trials and results are exactly those
 * of an hmc_ctx of its own.  No counterpart in the reference, which runs
 * replicas as separate Simulations. */
typedef struct hmc_batch hmc_batch;
int hmc_batch_create(uint32_t n_replicas, const hmc_config* cfgs, hmc_batch** out);
int hmc_batch_set_state(hmc_batch* batch, uint32_t replica, const double* body_pos, const double* site_pos);
int hmc_batch_get_state(hmc_batch* batch, uint32_t replica, double* body_pos, double* site_pos);
/* n_steps iterations of { apply; increment_step } for every replica r from
 * (step0[r], substep 0) with seeds[r] and max_distance[r]; out[r] receives the
 * summed counts.  use_graph = 0 launches every step directly (same results). */
int hmc_batch_run_translate(hmc_batch* batch, const uint64_t* step0, const uint32_t* seeds, double kT,
                            const double* max_distance, uint32_t n_steps, int use_graph, hmc_count* out);
/* Autotuner of the LDS-tiled batch pass (the tile size, i.e. spaces per
 * workgroup, of every replica).  For tile-size policies from the largest
 * tiles that fit down to small ones (tile shapes chosen as if each replica
 * had 1, 4, 16, 64 or 256 compute units to fill), it runs one step and then
 * n_steps timed steps of every replica from (step0, seeds, kT,
 * max_distance), restores the batch's positions, counters and counts after
 * each, and keeps the fastest layout: *best_ms (may be NULL) its time per
 * step, *best_policy (may be NULL) its compute-unit count.  The batch's
 * results do not depend on the layout.  HMC_ERR_UNSUPPORTED when the batch
 * does not run on LDS tiles.  No counterpart in the reference. */
int hmc_batch_autotune(hmc_batch* batch, const uint64_t* step0, const uint32_t* seeds, double kT,
                       const double* max_distance, uint32_t n_steps, double* best_ms, uint32_t* best_policy);
/* Statistics of one replica (applies and ambiguous are not kept for
 * batches: 0); tiled = 1 when the batch's passes run on LDS tiles (set
 * HMC_FLAG_GLOBAL_PASS in cfgs[0].flags for the lane-per-space pass). */
int hmc_batch_get_stats(hmc_batch* batch, uint32_t replica, hmc_stats* out);
/* hmc_sdf of every replica: hist holds n_replicas * nbins counts (replica
 * major), overlaps (may be NULL) n_replicas counts. */
int hmc_batch_sdf(hmc_batch* batch, double xmax, double dx, uint32_t nbins, uint64_t* hist, uint64_t* overlaps);
const char* hmc_batch_last_error(const hmc_batch* batch);
void hmc_batch_destroy(hmc_batch* batch);

#ifdef __cplusplus
}
#endif

#endif /* HMC_H */
