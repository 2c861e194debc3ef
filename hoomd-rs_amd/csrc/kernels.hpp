// Declarations of the sweep's kernels for the host translation units (each
// kernel is defined, and its template instances instantiated, in its k_*.hip).
#pragma once
#include "sweep_device.hpp"

namespace hmc {

__global__ void k_setup(Geometry g, ApplyState* st, uint64_t step, uint32_t substep, uint32_t seed, double kT, double max_move, int from_args, int advance, uint32_t move, int new_epoch);
__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v);
template <int D>
__global__ void k_bin(Geometry g, ApplyState* st, Bodies B, Work w);
__global__ void k_scan_blocks(const uint32_t* cnt, uint32_t* start, uint32_t* block_sums, uint32_t* block_ne, uint32_t n);
__global__ void k_scan_top(Geometry g, ApplyState* st, uint32_t* block_sums, const uint32_t* block_ne, uint32_t nblocks, uint32_t* start, int count);
__global__ void k_scan_lookback(Geometry g, ApplyState* st, uint32_t* cnt, uint32_t* start, unsigned long long* status, uint32_t n, uint32_t nblocks);
__global__ void k_scan_add(uint32_t* start, const uint32_t* block_sums, uint32_t* cnt, uint32_t n);
__global__ void k_scatter(Geometry g, Work w);
template <int PREC, bool COPY = true>
__global__ void k_place(Geometry g, Bodies B, Cells C, Work w);
template <int D, int PREC, bool FAST>
__global__ void k_pass(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, uint32_t xcd_chunk);
template <int D, int PREC, bool FP32_ERROR = false>
__global__ void k_pass_tile(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, TileCfg tc);
template <int D>
__global__ void k_apply_tile(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, TileCfg tc);
template <int D>
__global__ void k_bapply_tile(const BatchRep* reps, const TileCfg* tcs, const uint32_t* toff, const uint32_t* trep, ApplyState* sts, const uint32_t* rt0, const uint32_t* rn, uint32_t nrounds, uint32_t nwg, uint32_t xcd_chunk, uint32_t probe, Bodies B, Cells C, Work w);
template <int D>
__global__ void k_apply_lj(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, TileCfg tc);
template <int D, int PREC>
__global__ void k_guard(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t first_cycle);
template <int D, int PREC>
__global__ void k_pass_lj(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, uint32_t xcd_chunk);
template <int D>
__global__ void k_guard_lj(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t first_cycle);
template <bool MIXED>
__global__ void k_pass_convex(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, uint32_t xcd_chunk, SatShape sat);
__global__ void k_guard_convex(Geometry g, ApplyState* st, Bodies B, Cells C, Work w, uint32_t first_cycle);
__global__ void k_tune_eval_convex(Geometry g, const ApplyState* st, Bodies B, Work w, const uint64_t* __restrict__ xs, uint32_t M, uint32_t move, double size, uint32_t* __restrict__ next, uint8_t* __restrict__ flags);
__global__ void k_init_site_quats(Geometry g, Bodies B, int have_sites);
__global__ void k_debug_convex(int kind, const double* vx, uint32_t nv, double rcut2, const double* in, double* out, uint64_t n, SatShape sat);
__global__ void k_bsetup(const BatchRep* reps, ApplyState* sts, uint32_t R, const uint64_t* step, const uint32_t* seed, double kT, const double* max_move, int from_args, int advance);
template <int D>
__global__ void k_bbin(const BatchRep* reps, ApplyState* sts, const uint32_t* boff, uint32_t R, uint32_t n_total, Bodies B, Work w);
__global__ void k_bcycles(const BatchRep* reps, ApplyState* sts, uint32_t R);
template <int D, int PREC, bool FAST>
__global__ void k_bpass(const BatchRep* reps, ApplyState* sts, const uint32_t* coff, uint32_t R, uint32_t per_colour_total, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle, uint32_t xcd_chunk);
template <int D, int PREC>
__global__ void k_bpass_tile(const BatchRep* reps, const TileCfg* tcs, const uint32_t* toff, const uint32_t* trep, ApplyState* sts, uint32_t ntiles, uint32_t xcd_chunk, Bodies B, Cells C, Work w, uint32_t colour, uint32_t cycle);
template <int D, int PREC>
__global__ void k_bguard(const BatchRep* reps, ApplyState* sts, Bodies B, Cells C, Work w, uint32_t first_cycle);
template <int D>
__global__ void k_energy(Geometry g, const ApplyState* st, Bodies B, Work w, double* part);
__global__ void k_energy_final(const double* part, uint32_t nblocks, double* out);
template <int D>
__global__ void k_insert_energy(Geometry g, const ApplyState* st, Bodies B, Work w, const double* __restrict__ cand, uint32_t n, double* __restrict__ out, uint32_t* outside);
template <int D>
__global__ void k_tune_eval(Geometry g, const ApplyState* st, Bodies B, Work w, const uint64_t* __restrict__ xs, uint32_t M, double max_move, double kT, uint32_t* __restrict__ next, uint8_t* __restrict__ flags);
template <int D>
__global__ void k_remove_energy(Geometry g, const ApplyState* st, Bodies B, Work w, uint32_t body, double* out);
template <int D>
__global__ void k_sdf(const BatchRep* reps, const uint32_t* boff, uint32_t R, uint32_t n_total, Bodies B, Work w, double xmax, double dx, uint32_t nbins, uint32_t* hist, uint32_t* ovl);
template <int D>
__global__ void k_init_sites(Geometry g, Bodies B, int have_sites);
template <int D>
__global__ void k_map_bodies(Geometry g, double3 lold, const double4* __restrict__ src, Bodies B, uint32_t* outside);
__global__ void k_tally(ApplyState* st, const unsigned long long* acc_slots, uint32_t m);
__global__ void k_moved_mask(const double4* __restrict__ bpos, uint32_t n, uint32_t since, unsigned long long* mask,
                             unsigned long long* count);
__global__ void k_gather_bodies(Bodies B, const uint32_t* __restrict__ idx, uint32_t n, double4* __restrict__ out);
__global__ void k_pack_frame(Bodies B, uint32_t n, int dim, float* __restrict__ pos, float* __restrict__ quat);
__global__ void k_debug_math(int kind, const double* in, double* out, uint64_t n);
__global__ void k_debug_normals(uint64_t step, uint32_t substep, uint32_t seed, uint64_t index, double* out, uint64_t n);

}  // namespace hmc
