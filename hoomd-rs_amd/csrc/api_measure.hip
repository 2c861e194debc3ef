// Measurements through the C ABI: the SDF, insertion energies and the
// total energy.
#include "host.hpp"

namespace hmc {

// Arguments of the SDF for one system: hard spheres/disks, 0 < xmax < 1/2,
// dx > 0, a box wider than 2 r_max (one image within r_max) and spaces at
// least r_max / 2 wide (two neighbour spaces on each side suffice).
const char* sdf_check(const Geometry& g, double xmax, double dx, uint32_t nbins) {
  if (g.model != HMC_HARD_SPHERE) return "the SDF is built for hard spheres and disks";
  if (!(xmax > 0.0 && xmax < 0.5) || !(dx > 0.0) || !std::isfinite(dx) || nbins == 0 || nbins > (1u << 24))
    return "SDF needs 0 < xmax < 1/2, dx > 0 and 0 < nbins <= 2^24";
  const double rmax = g.sigma / (1.0 - xmax);
  for (int a = 0; a < g.dim; ++a) {
    if (!(g.L[a] > 2.0 * rmax * (1.0 + 0x1p-30))) return "box edge below 2 sigma / (1 - xmax)";
    if (!(2.0 * g.width[a] >= rmax * (1.0 + 0x1p-30))) return "checkerboard spaces narrower than r_max / 2";
  }
  return nullptr;
}

// The SDF's own cells: the checkerboard with spaces at least r_max (1 + 2^-29)
// wide (never more spaces than the sweep's, whose buffers it reuses), so
// that one neighbour space on each side covers r_max.
Geometry sdf_geometry(const Geometry& g, double xmax) {
  Geometry s = g;
  const double rmax = g.sigma / (1.0 - xmax);
  uint64_t spaces = 1;
  for (int a = 0; a < g.dim; ++a) {
    int64_t k = (int64_t)std::floor(g.L[a] / (rmax * (1.0 + 0x1p-29)));
    k = std::max<int64_t>(1, std::min<int64_t>(k, g.shape[a]));
    s.shape[a] = (int32_t)k;
    s.half_shape[a] = (uint32_t)std::max<int64_t>(1, k / 2);
    s.width[a] = g.L[a] / (double)k;
    spaces *= (uint64_t)k;
  }
  uint32_t str = 1;
  for (int a = g.dim - 1; a >= 0; --a) {
    s.stride[a] = str;
    str *= (uint32_t)s.shape[a];
  }
  s.n_spaces = (uint32_t)spaces;
  s.per_colour = s.n_spaces >> g.dim;
  return s;
}

// PairwiseCutoff::total_energy of the current state with the block partials
// in part (4 nblk + 4 doubles, nblk = ceil(N / kEnergyBlock)): the cells of
// the current counter's checkerboard, k_energy per block, k_energy_final in
// a fixed order.  Allocates nothing; reads back only the result and the
// state's error word.
int total_energy_into(hmc_ctx* c, double* part, double* energy, double* bound) {
  const Geometry& g = c->g;
  const uint32_t N = g.n_bodies;
  if (N == 0) {
    *energy = 0.0;
    if (bound) *bound = 0.0;
    return HMC_OK;
  }
  hipStream_t s = c->stream;
  const uint32_t nblk = grid_for(N, kEnergyBlock);
  // cells for the neighbour search: the checkerboard of the current counter
  k_setup<<<1, 64, 0, s>>>(g, c->st, 0, 0, 0, 1.0, 1.0, 0, 0, 0u, 0);
  if (g.dim == 2)
    k_bin<2><<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->st, c->bodies, c->w);
  else
    k_bin<3><<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->st, c->bodies, c->w);
  k_scan_blocks<<<c->n_scan_blocks, 256, 0, s>>>(c->w.cnt, c->w.start, c->w.block_sums, c->w.block_ne, g.n_spaces);
  k_scan_top<<<1, 1024, 0, s>>>(g, c->st, c->w.block_sums, c->w.block_ne, c->n_scan_blocks, c->w.start, 0);
  k_scan_add<<<grid_for(g.n_spaces, 256), 256, 0, s>>>(c->w.start, c->w.block_sums, c->w.cnt, g.n_spaces);
  k_scatter<<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->w);
  if (g.dim == 2)
    k_energy<2><<<nblk, kEnergyBlock, 0, s>>>(g, c->st, c->bodies, c->w, part);
  else
    k_energy<3><<<nblk, kEnergyBlock, 0, s>>>(g, c->st, c->bodies, c->w, part);
  k_energy_final<<<1, kEnergyFinal, 0, s>>>(part, nblk, part + 4 * (size_t)nblk);
  double out[4] = {0, 0, 0, 0};
  uint32_t err = 0;
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, part + 4 * (size_t)nblk, sizeof out, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(&err, &c->st->error, sizeof err, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(c, HMC_ERR_DEVICE, std::string("total energy: ") + hipGetErrorString(e));
  // (the next apply rebuilds the cells; the neighbour copies were not touched)
  if (err & kErrBarrier) return fail(c, HMC_ERR_DEVICE, "a device-side wait timed out");
  if (err) return fail(c, HMC_ERR_COVER, "a body lies outside the checkerboard");
  // any two summation orders of n terms differ by at most 2 gamma_{n-1} sum|terms|
  const double n = std::max(out[2], 1.0);
  const double gam = n * 0x1p-53 / (1.0 - n * 0x1p-53);
  *energy = out[3] != 0.0 ? std::numeric_limits<double>::infinity() : out[0];
  if (bound) *bound = out[3] != 0.0 ? 0.0 : 2.0 * gam * out[1];
  return HMC_OK;
}

// ------------------------------------------- static scratch of the no-alloc --
// hmc_total_energy_noalloc keeps the block partials of the total energy in
// this array instead of a per-call allocation: 4 kNoAllocBlocks + 4 doubles
// (1 MiB + 32 B), the partials of up to 8,388,608 bodies.  Each GPU the
// library runs on holds its own copy until the process ends.  In host code
// the array's name is only a host-side stand-in: kernels receive the device
// address taken once per device by noalloc_setup.  Calls on one GPU share it
// (each call writes every partial it reads): calls that overlap on one GPU
// race.
static __device__ __align__(256) double scratch[4 * kNoAllocBlocks + 4];
static std::mutex g_scratch_mu;
void* g_scratch[kNoAllocDevices];

// The device address of scratch on `device`, taken once per device with that
// device current (hmc_create calls it after selecting the context's device).
int noalloc_setup(int device) {
  if (device < 0 || device >= kNoAllocDevices) return HMC_ERR_INVALID;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return HMC_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return HMC_ERR_DEVICE;
  g_scratch[device] = p;
  return HMC_OK;
}

}  // namespace hmc

extern "C" {

int hmc_sdf(hmc_ctx* c, double xmax, double dx, uint32_t nbins, uint64_t* hist, uint64_t* overlaps) {
  if (!c || !hist) return HMC_ERR_INVALID;
  if (!c->have_state) return fail(c, HMC_ERR_STATE, "hmc_set_state was not called");
  if (const char* why = sdf_check(c->g, xmax, dx, nbins))
    return fail(c, c->g.model != HMC_HARD_SPHERE ? HMC_ERR_UNSUPPORTED : HMC_ERR_INVALID, why);
  const Geometry g = sdf_geometry(c->g, xmax);
  const uint32_t nsb = grid_for(g.n_spaces, kScanBlock);
  const uint32_t N = g.n_bodies;
  HIPCHK(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  // device: the system as a one-entry batch table, body offsets, counts
  const size_t bytes = sizeof(BatchRep) + 2 * sizeof(uint32_t) + ((size_t)nbins + 1) * sizeof(uint32_t);
  char* buf = nullptr;
  HIPCHK(c, hipMallocAsync((void**)&buf, bytes, s));
  BatchRep* drep = (BatchRep*)buf;
  uint32_t* dboff = (uint32_t*)(buf + sizeof(BatchRep));
  uint32_t* dh = dboff + 2;
  const BatchRep hrep{g, 0u, 0u, 0u, 0u};
  const uint32_t hboff[2] = {0u, N};
  hipError_t e = hipMemcpyAsync(drep, &hrep, sizeof hrep, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(dboff, hboff, sizeof hboff, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    k_fill_u32<<<grid_for(nbins + 1, 256), 256, 0, s>>>(dh, nbins + 1, 0u);
    if (N) {
      // cells of the current state on the SDF's checkerboard (origin drawn
      // from the current counter; the next apply rebuilds its own)
      k_setup<<<1, 64, 0, s>>>(g, c->st, 0, 0, 0, 1.0, 1.0, 0, 0, 0u, 0);
      if (g.dim == 2)
        k_bin<2><<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->st, c->bodies, c->w);
      else
        k_bin<3><<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->st, c->bodies, c->w);
      k_scan_blocks<<<nsb, 256, 0, s>>>(c->w.cnt, c->w.start, c->w.block_sums, c->w.block_ne, g.n_spaces);
      k_scan_top<<<1, 1024, 0, s>>>(g, c->st, c->w.block_sums, c->w.block_ne, nsb, c->w.start, 0);
      k_scan_add<<<grid_for(g.n_spaces, 256), 256, 0, s>>>(c->w.start, c->w.block_sums, c->w.cnt, g.n_spaces);
      k_scatter<<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->w);
      if (g.dim == 2)
        k_sdf<2><<<grid_for(N, kBlock), kBlock, 0, s>>>(drep, dboff, 1u, N, c->bodies, c->w, xmax, dx, nbins, dh,
                                                        dh + nbins);
      else
        k_sdf<3><<<grid_for(N, kBlock), kBlock, 0, s>>>(drep, dboff, 1u, N, c->bodies, c->w, xmax, dx, nbins, dh,
                                                        dh + nbins);
    }
    e = hipGetLastError();
  }
  std::vector<uint32_t> h(nbins + 1, 0u);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), dh, h.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFreeAsync(buf, s);
  if (e != hipSuccess) return fail(c, HMC_ERR_DEVICE, std::string("sdf: ") + hipGetErrorString(e));
  ApplyState st;  // a body outside the checkerboard is reported here
  int rc = read_state(c, &st);
  if (rc) return rc;
  for (uint32_t k = 0; k < nbins; ++k) hist[k] = h[k];
  if (overlaps) *overlaps = h[nbins];
  return HMC_OK;
}

}  // extern "C"

void hmc::launch_cell_lists(hmc_ctx* c) {
  const Geometry& g = c->g;
  const uint32_t N = g.n_bodies;
  hipStream_t s = c->stream;
  k_setup<<<1, 64, 0, s>>>(g, c->st, 0, 0, 0, 1.0, 1.0, 0, 0, 0u, 0);
  if (g.dim == 2)
    k_bin<2><<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->st, c->bodies, c->w);
  else
    k_bin<3><<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->st, c->bodies, c->w);
  k_scan_blocks<<<c->n_scan_blocks, 256, 0, s>>>(c->w.cnt, c->w.start, c->w.block_sums, c->w.block_ne, g.n_spaces);
  k_scan_top<<<1, 1024, 0, s>>>(g, c->st, c->w.block_sums, c->w.block_ne, c->n_scan_blocks, c->w.start, 0);
  k_scan_add<<<grid_for(g.n_spaces, 256), 256, 0, s>>>(c->w.start, c->w.block_sums, c->w.cnt, g.n_spaces);
  k_scatter<<<grid_for(N, kBlock), kBlock, 0, s>>>(g, c->w);
}

extern "C" {

int hmc_delta_energy_insert(hmc_ctx* c, const double* body_pos, uint64_t n, double* energy, double* bound) {
  if (!c || (n && (!body_pos || !energy))) return HMC_ERR_INVALID;
  if (!c->have_state) return fail(c, HMC_ERR_STATE, "hmc_set_state was not called");
  const Geometry& g = c->g;
  if (g.model == HMC_CONVEX_POLYHEDRON)
    return fail(c, HMC_ERR_UNSUPPORTED, "insertion energies of oriented bodies are not offloaded");
  if (n >= 0xffffffffull) return fail(c, HMC_ERR_INVALID, "too many candidates");
  if (n == 0) return HMC_OK;
  const int D = g.dim;
  const uint32_t N = g.n_bodies;
  if (N == 0) {  // no sites: every insertion energy is 0
    for (uint64_t k = 0; k < n; ++k) {
      energy[k] = 0.0;
      if (bound) bound[k] = 0.0;
    }
    return HMC_OK;
  }
  HIPCHK(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  double* dc = nullptr;
  double* dout = nullptr;
  uint32_t* flag = nullptr;
  hipError_t e = hipMallocAsync((void**)&dc, (size_t)n * D * sizeof(double), s);
  if (e == hipSuccess) e = hipMallocAsync((void**)&dout, 2 * (size_t)n * sizeof(double), s);
  if (e == hipSuccess) e = hipMallocAsync((void**)&flag, sizeof(uint32_t), s);
  if (e == hipSuccess) e = hipMemcpyAsync(dc, body_pos, (size_t)n * D * sizeof(double), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemsetAsync(flag, 0, sizeof(uint32_t), s);
  if (e == hipSuccess) {
    launch_cell_lists(c);  // cells of the current state, as for the total energy
    if (D == 2)
      k_insert_energy<2><<<grid_for(n, kEnergyBlock), kEnergyBlock, 0, s>>>(g, c->st, c->bodies, c->w, dc,
                                                                           (uint32_t)n, dout, flag);
    else
      k_insert_energy<3><<<grid_for(n, kEnergyBlock), kEnergyBlock, 0, s>>>(g, c->st, c->bodies, c->w, dc,
                                                                           (uint32_t)n, dout, flag);
    e = hipGetLastError();
  }
  std::vector<double> hout(2 * (size_t)n);
  uint32_t outside = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(hout.data(), dout, 2 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(&outside, flag, sizeof(uint32_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (dc) (void)hipFreeAsync(dc, s);
  if (dout) (void)hipFreeAsync(dout, s);
  if (flag) (void)hipFreeAsync(flag, s);
  if (e != hipSuccess) return fail(c, HMC_ERR_DEVICE, std::string("insertion energy: ") + hipGetErrorString(e));
  ApplyState st;
  int rc = read_state(c, &st);
  if (rc) return rc;
  if (outside) return fail(c, HMC_ERR_COVER, "a candidate lies outside the checkerboard");
  for (uint64_t k = 0; k < n; ++k) {
    energy[k] = hout[2 * k];
    if (bound) bound[k] = hout[2 * k + 1];
  }
  return HMC_OK;
}

int hmc_total_energy(hmc_ctx* c, double* energy, double* bound) {
  if (!c || !energy) return HMC_ERR_INVALID;
  if (!c->have_state) return fail(c, HMC_ERR_STATE, "hmc_set_state was not called");
  HIPCHK(c, hipSetDevice(c->device));
  if (c->g.n_bodies == 0) return total_energy_into(c, nullptr, energy, bound);
  const uint32_t nblk = grid_for(c->g.n_bodies, kEnergyBlock);
  double* part = nullptr;
  HIPCHK(c, hipMallocAsync((void**)&part, (4 * (size_t)nblk + 4) * sizeof(double), c->stream));
  const int rc = total_energy_into(c, part, energy, bound);
  (void)hipFreeAsync(part, c->stream);
  return rc;
}

int hmc_total_energy_noalloc(hmc_ctx* c, double* energy, double* bound) {
  if (!c || !energy) return HMC_ERR_INVALID;
  if (!c->have_state) return fail(c, HMC_ERR_STATE, "hmc_set_state was not called");
  if (grid_for(c->g.n_bodies, kEnergyBlock) > kNoAllocBlocks)
    return fail(c, HMC_ERR_INVALID,
                "hmc_total_energy_noalloc: more than 8,388,608 bodies (the static scratch holds the partials of "
                "32,768 blocks of 256); use hmc_total_energy");
  if (c->device < 0 || c->device >= kNoAllocDevices || !g_scratch[c->device])
    return fail(c, HMC_ERR_DEVICE, "hmc_total_energy_noalloc: the scratch of this device was not set up");
  HIPCHK(c, hipSetDevice(c->device));
  return total_energy_into(c, (double*)g_scratch[c->device], energy, bound);
}

}  // extern "C"
