// Host-side state of a context and of a batch, and the helpers the C ABI's
// translation units (api_*.hip) share.
#pragma once
#include <array>
#include <mutex>

#include "kernels.hpp"

using namespace hmc;

struct hmc_ctx {
  hmc_config cfg;
  hmc_config cfg_user;              // as given to hmc_create (vertices -> verts_host)
  std::vector<double> verts_host;   // convex: the vertex table, for hmc_clone_with_boundary
  Geometry g;
  int device;
  hipStream_t stream = nullptr;
  Bodies bodies{};
  Cells cells{};
  Work w{};
  ApplyState* st = nullptr;
  bool have_state = false;
  uint32_t cycles_guess = 1;
  bool fast = false;  // every axis: >= 3 spaces and reach 1
  bool tiled = false;  // LDS-staged pass (requires fast)
  TileCfg tc{};
  bool persist = false;  // one cooperative launch per apply (k_apply_tile)
  TileCfg pc{};
  uint32_t pgrid = 0, pthreads = 0, pper_cu = 0;
  double* verts = nullptr;  // convex polyhedron vertices (device)
  SatShape sat{};           // convex, mixed precision: the FP32 separating axes (else use = 0)
  Geometry* gdev = nullptr;  // device copy of g (g.self)
  uint32_t n_scan_blocks = 0;
  uint64_t applies = 0;
  std::string err;
  bool profiling = false;
  bool capturing = false;       // recording a graph: no event timing
  uint32_t profile_period = 1;  // time the launches of every k-th apply
  // one apply (with increment_step) captured as a graph, and what it depends on
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  double graph_kT = 0.0;
  std::vector<hmc_move> graph_moves;  // the step the graph holds: one apply per move
  uint32_t graph_launched = 0;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pass_events, build_events;
  double pass_ms = 0.0, build_ms = 0.0;
  uint64_t pass_n = 0, build_n = 0;
};

struct hmc_batch {
  int device = 0;
  uint32_t R = 0;
  int dim = 3;
  int precision = HMC_FP64;
  bool fast = true;
  double kT = 1.0;
  std::vector<BatchRep> hreps;
  std::vector<uint32_t> boff, coff;  // R + 1 running offsets of bodies and colour spaces
  uint32_t n_total = 0, s_total = 0, pc_total = 0;
  Geometry gtot{};  // totals for the replica-agnostic kernels
  BatchRep* reps = nullptr;
  ApplyState* sts = nullptr;  // R + 1 (the last one is scratch)
  uint32_t *dboff = nullptr, *dcoff = nullptr;
  uint64_t* dstep = nullptr;
  uint32_t* dseed = nullptr;
  double* dmove = nullptr;
  Bodies bodies{};
  Cells cells{};
  Work w{};
  hipStream_t stream = nullptr;
  uint32_t n_scan_blocks = 0;
  uint32_t cycles_guess = 1;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint32_t graph_cycles = 0;
  // LDS-tiled passes (k_bpass_tile): per-replica tile shapes, first tiles
  bool tiled = false;
  TileCfg* dtcs = nullptr;
  uint32_t* dtoff = nullptr;
  uint32_t* dtrep = nullptr;  // replica of each tile
  uint32_t dtrep_cap = 0;
  uint32_t ntiles = 0, tile_threads = 0, tile_lds = 0;
  // persistent applies (mixed, 3D; k_bapply_tile): one region shape per
  // replica, the replicas' first regions, the replica of each region, and
  // the launch groups (whole replicas whose regions are resident together)
  bool persist = false;
  TileCfg* dptcs = nullptr;
  uint32_t* dptoff = nullptr;
  uint32_t* dptrep = nullptr;
  uint32_t pthreads = 0, plds = 0;
  std::vector<uint32_t> grp_tile0, grp_ntiles;  // the rounds (host copy)
  uint32_t* drt0 = nullptr;                     // the rounds' first regions and sizes (device)
  uint32_t* drn = nullptr;
  uint32_t nwg = 0;                             // workgroups of the launch: the largest round
  uint32_t pper_cu = 0, pregions = 0;           // resident regions per CU; all regions
  std::vector<TileCfg> hptcs;                   // the replicas' region shapes (host copy)
  // per replica and region shape T (index T): the largest body count of a
  // box of (2T + 2)^3 spaces over the state last set, every placement of the
  // box on the space grid (0: not known); sizes the regions' LDS capacity
  std::vector<std::array<uint32_t, 7>> box_max;
  bool regions_stale = false;  // a state was set: choose the regions again before the next step
  int pcus = 0;                // compute units the regions were chosen for
  hipError_t launch_err = hipSuccess;
  std::string err;
};

namespace hmc {

inline int fail(hmc_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, expr)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) return fail((ctx), HMC_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

template <class T>
hipError_t dalloc(T** p, size_t n) {
  return hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
}

inline uint32_t grid_for(uint64_t n, uint32_t block) { return (uint32_t)((n + block - 1) / block); }

hipEvent_t take_event(hmc_ctx* c);
int dispatch_apply(hmc_ctx* c, int from_args, int advance, uint64_t step, uint32_t substep, uint32_t seed,
                   double kT, double max_move, uint32_t move);
void drop_graph(hmc_ctx* c);
void collect_events(hmc_ctx* c, std::vector<std::pair<hipEvent_t, hipEvent_t>>& evs, double& ms_total,
                    uint64_t& n);
int read_state(hmc_ctx* c, ApplyState* out, bool check_errors = true);
bool contact_band(double E, double sigma, double sigma2, double edge2, int D, float* lo, float* hi);
bool choose_tiles(const Geometry& g, int D, double occ, size_t vt, int cus, TileCfg* out);
bool choose_regions(const Geometry& g, int D, double occ, int cus, TileCfg* out);
bool choose_regions_lj(const Geometry& g, int D, double occ, int cus, TileCfg* out);
int make_geometry(const hmc_config* cfg, Geometry& g, bool* fast, double* sigma_out, double* lmax_out);
const char* sdf_check(const Geometry& g, double xmax, double dx, uint32_t nbins);
Geometry sdf_geometry(const Geometry& g, double xmax);
int total_energy_into(hmc_ctx* c, double* part, double* energy, double* bound);
// Cell lists of the current state on c's stream (checkerboard spaces at
// origin 0, bodies of each space in body order: w.start, w.tmp_body).
void launch_cell_lists(hmc_ctx* c);

// The allocation-free total energy (hmc_total_energy_noalloc): its static
// scratch holds the partials of kNoAllocBlocks energy blocks; noalloc_setup
// records the scratch's device address per device (g_scratch).
constexpr uint32_t kNoAllocBlocks = 32768;
constexpr int kNoAllocDevices = 64;
extern void* g_scratch[kNoAllocDevices];
int noalloc_setup(int device);

}  // namespace hmc
