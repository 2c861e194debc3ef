"""hoomd-rs_amd: MI355X-native hoomd-mc local trial-move sweep.

The product is libhmc.so (HIP kernels for gfx950 behind the C ABI declared in
include/hmc.h).  This module is a thin ctypes view of that ABI for Python
callers (tests, benchmark).  It has no fallback: if libhmc.so is missing or
no GPU is present the calls raise.
"""
import atexit
import ctypes as C
import itertools
import pathlib
import weakref

import numpy as np

PKG = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG / "libhmc.so"

HMC_OK = 0
HMC_ERR_INVALID = -1
HMC_ERR_UNSUPPORTED = -2
HMC_ERR_DEVICE = -3
HMC_ERR_COVER = -4
HMC_ERR_STATE = -5

HMC_HARD_SPHERE = 0
HMC_LJ = 1
HMC_LJ_SHIFTED = 2
HMC_CONVEX_POLYHEDRON = 3
HMC_OVERLAP_PENALTY = 4
HMC_FP64 = 0
HMC_MIXED = 1

# every entry point declared in include/hmc.h
EXPORTS = ["hmc_create", "hmc_set_state", "hmc_get_state", "hmc_clone_with_boundary", "hmc_apply_translate",
           "hmc_quick_compress_init", "hmc_quick_compress_apply", "hmc_quick_insert_init",
           "hmc_quick_insert_apply", "hmc_delta_energy_insert", "hmc_tune_local", "hmc_grand_canonical_apply", "hmc_gsd_create", "hmc_gsd_open",
           "hmc_gsd_write_chunk", "hmc_gsd_end_frame", "hmc_gsd_append_frame", "hmc_gsd_n_frames",
           "hmc_gsd_read_chunk", "hmc_gsd_header", "hmc_gsd_close",
           "hmc_apply_rotate",
           "hmc_set_orientations", "hmc_get_orientations", "hmc_run_translate", "hmc_total_energy",
           "hmc_total_energy_noalloc", "hmc_run", "hmc_moved_since", "hmc_get_bodies", "hmc_debug_regions",
           "hmc_trace_enable", "hmc_trace_read", "hmc_get_stats", "hmc_profile_enable", "hmc_profile_read",
           "hmc_debug_math", "hmc_debug_normals", "hmc_debug_convex", "hmc_debug_convex_axes", "hmc_last_error",
           "hmc_destroy",
           "hmc_batch_create", "hmc_batch_set_state", "hmc_batch_get_state", "hmc_batch_run_translate",
           "hmc_batch_last_error", "hmc_batch_destroy", "hmc_sdf", "hmc_batch_sdf", "hmc_batch_get_stats", "hmc_batch_autotune"]


# Contexts and batches still open when the interpreter exits are destroyed by
# an atexit hook, in creation order, while the HIP runtime (and any profiler
# attached to it) is still up -- not by __del__ during interpreter teardown.
_LIVE = weakref.WeakValueDictionary()
_LIVE_ID = itertools.count()


@atexit.register
def _close_all():
    for _k, obj in sorted(list(_LIVE.items())):
        obj.close()


class HmcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hmc error {code}: {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [("dim", C.c_uint32), ("n_bodies", C.c_uint64), ("box_edges", C.c_double * 3),
                ("boundary_range", C.c_double), ("body_interaction_range", C.c_double),
                ("spatial_cell_width", C.c_double),
                ("model", C.c_int32), ("params", C.c_double * 4), ("precision", C.c_int32),
                ("device", C.c_int32), ("flags", C.c_uint32),
                ("vertices", C.POINTER(C.c_double)), ("n_vertices", C.c_uint32)]

FLAG_GLOBAL_PASS = 1
FLAG_NO_GRAPH = 2
FLAG_PER_COLOUR = 4
FLAG_FP32_ERROR = 8


class Count(C.Structure):
    _fields_ = [("accepted", C.c_uint64), ("rejected", C.c_uint64)]


class GcCount(C.Structure):
    _fields_ = [("insert_accepted", C.c_uint64), ("insert_rejected", C.c_uint64),
                ("remove_accepted", C.c_uint64), ("remove_rejected", C.c_uint64)]


class Move(C.Structure):
    _fields_ = [("move", C.c_int32), ("pad", C.c_int32), ("size", C.c_double)]


HMC_MOVE_TRANSLATE = 0
HMC_MOVE_ROTATE = 1


class Stats(C.Structure):
    _fields_ = [("applies", C.c_uint64), ("cycles", C.c_uint64), ("pair_tests", C.c_uint64),
                ("band_retests", C.c_uint64), ("guard_cycles", C.c_uint64),
                ("tile_overflows", C.c_uint64), ("tiled", C.c_uint32), ("tile_shape", C.c_uint32 * 3),
                ("ambiguous", C.c_uint64), ("fp32_error_max", C.c_double), ("fp32_error_bound", C.c_double),
                ("regions", C.c_uint32), ("region_threads", C.c_uint32), ("regions_per_cu", C.c_uint32),
                ("wait_timeout_at", C.c_uint32), ("rounds", C.c_uint32)]


TRACE_DTYPE = np.dtype([("substep", "<u4"), ("color", "<u4"), ("space", "<u8"),
                        ("body", "<i8"), ("status", "<i4"), ("pad", "<i4")])

_lib = None
D_P = C.POINTER(C.c_double)


def lib():
    """Load libhmc.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(str(LIB_PATH))
        vp = C.c_void_p
        L.hmc_create.argtypes = [C.POINTER(Config), C.POINTER(vp)]
        L.hmc_set_state.argtypes = [vp, D_P, D_P]
        L.hmc_get_state.argtypes = [vp, D_P, D_P]
        L.hmc_clone_with_boundary.argtypes = [vp, D_P, C.POINTER(vp)]
        L.hmc_delta_energy_insert.argtypes = [vp, D_P, C.c_uint64, D_P, D_P]
        L.hmc_quick_compress_init.argtypes = [C.POINTER(QuickCompressState), C.c_double]
        L.hmc_quick_compress_init.restype = None
        L.hmc_quick_compress_apply.argtypes = [C.POINTER(QuickCompressState), C.POINTER(vp), C.c_uint64,
                                               C.c_uint32, C.c_uint32]
        L.hmc_quick_insert_init.argtypes = [C.POINTER(QuickInsertState), C.c_uint64]
        L.hmc_quick_insert_init.restype = None
        L.hmc_quick_insert_apply.argtypes = [C.POINTER(QuickInsertState), C.POINTER(vp), C.c_uint64,
                                             C.c_uint32, C.c_uint32, C.POINTER(Count)]
        L.hmc_tune_local.argtypes = [vp, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                     C.c_uint64, C.c_uint32, D_P, C.POINTER(Count), C.POINTER(C.c_uint64)]
        L.hmc_grand_canonical_apply.argtypes = [C.POINTER(vp), C.c_uint64, C.c_uint32, C.c_uint32, C.c_double,
                                                C.c_double, C.POINTER(GcCount), C.POINTER(C.c_uint64)]
        L.hmc_gsd_create.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_uint32, C.c_uint32, C.POINTER(vp)]
        L.hmc_gsd_open.argtypes = [C.c_char_p, C.c_int, C.POINTER(vp)]
        L.hmc_gsd_write_chunk.argtypes = [vp, C.c_char_p, C.c_uint8, C.c_uint64, C.c_uint32, vp]
        L.hmc_gsd_end_frame.argtypes = [vp]
        L.hmc_gsd_append_frame.argtypes = [vp, vp, C.c_uint64]
        L.hmc_gsd_n_frames.argtypes = [vp]
        L.hmc_gsd_n_frames.restype = C.c_uint64
        L.hmc_gsd_read_chunk.argtypes = [vp, C.c_uint64, C.c_char_p, vp, C.c_uint64, C.POINTER(C.c_uint64),
                                         C.POINTER(C.c_uint32), C.POINTER(C.c_uint8)]
        L.hmc_gsd_header.argtypes = [vp, C.c_char_p, C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.hmc_gsd_close.argtypes = [vp]
        L.hmc_apply_translate.argtypes = [vp, C.c_uint64, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                          C.POINTER(Count), C.POINTER(C.c_uint32)]
        L.hmc_run_translate.argtypes = [vp, C.c_uint64, C.c_uint32, C.c_double, C.c_double, C.c_uint32,
                                        C.POINTER(Count)]
        L.hmc_apply_rotate.argtypes = [vp, C.c_uint64, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                       C.POINTER(Count), C.POINTER(C.c_uint32)]
        L.hmc_set_orientations.argtypes = [vp, D_P, D_P]
        L.hmc_total_energy.argtypes = [vp, D_P, D_P]
        L.hmc_total_energy_noalloc.argtypes = [vp, D_P, D_P]
        L.hmc_run.argtypes = [vp, C.c_uint64, C.c_uint32, C.c_double, C.POINTER(Move), C.c_uint32, C.c_uint32,
                              C.POINTER(Count)]
        L.hmc_moved_since.argtypes = [vp, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                      C.POINTER(C.c_uint64)]
        L.hmc_get_bodies.argtypes = [vp, C.POINTER(C.c_uint32), C.c_uint64, D_P, D_P, D_P, D_P]
        L.hmc_debug_regions.argtypes = [vp, C.POINTER(C.c_uint32), C.c_uint64, C.POINTER(C.c_uint64)]
        u64p = C.POINTER(C.c_uint64)
        L.hmc_sdf.argtypes = [vp, C.c_double, C.c_double, C.c_uint32, u64p, u64p]
        L.hmc_batch_sdf.argtypes = [vp, C.c_double, C.c_double, C.c_uint32, u64p, u64p]
        L.hmc_batch_get_stats.argtypes = [vp, C.c_uint32, C.POINTER(Stats)]
        L.hmc_batch_autotune.argtypes = [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32), C.c_double, D_P,
                                         C.c_uint32, D_P, C.POINTER(C.c_uint32)]
        L.hmc_batch_create.argtypes = [C.c_uint32, C.POINTER(Config), C.POINTER(vp)]
        L.hmc_batch_set_state.argtypes = [vp, C.c_uint32, D_P, D_P]
        L.hmc_batch_get_state.argtypes = [vp, C.c_uint32, D_P, D_P]
        L.hmc_batch_run_translate.argtypes = [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32), C.c_double, D_P,
                                              C.c_uint32, C.c_int, C.POINTER(Count)]
        L.hmc_batch_last_error.argtypes = [vp]
        L.hmc_batch_last_error.restype = C.c_char_p
        L.hmc_batch_destroy.argtypes = [vp]
        L.hmc_get_orientations.argtypes = [vp, D_P, D_P]
        L.hmc_debug_convex.argtypes = [C.c_int, C.c_int, D_P, C.c_uint32, D_P, D_P, C.c_uint64]
        L.hmc_debug_convex_axes.argtypes = [D_P, C.c_uint32, C.POINTER(C.c_uint32), D_P, D_P]
        L.hmc_trace_enable.argtypes = [vp, C.c_int, C.c_uint64]
        L.hmc_trace_read.argtypes = [vp, vp, C.c_uint64, C.POINTER(C.c_uint64)]
        L.hmc_get_stats.argtypes = [vp, C.POINTER(Stats)]
        L.hmc_profile_enable.argtypes = [vp, C.c_int]
        L.hmc_profile_read.argtypes = [vp, D_P, C.POINTER(C.c_uint64), D_P, C.POINTER(C.c_uint64)]
        L.hmc_debug_math.argtypes = [C.c_int, C.c_int, D_P, D_P, C.c_uint64]
        L.hmc_debug_normals.argtypes = [C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64, D_P,
                                        C.c_uint64]
        L.hmc_last_error.argtypes = [vp]
        L.hmc_last_error.restype = C.c_char_p
        L.hmc_destroy.argtypes = [vp]
        L.hmc_destroy.restype = None
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(D_P)


class GpuParallelSweep:
    """A system resident on the GPU, advanced by ParallelSweep::apply with
    Translate moves (hoomd-mc/src/parallel_sweep.rs:390-421)."""

    def __init__(self, box_edges, positions, *, sites=None, diameter=1.0, boundary_range=None,
                 body_interaction_range=None, spatial_cell_width=None, precision="fp64", device=0,
                 lds_tiles=True, model="hard_sphere", params=None, vertices=None, orientations=None,
                 site_orientations=None, graphs=True, persistent=True, fp32_error=False):
        """positions / sites: the microstate's body and site positions (already
        wrapped into the box, as Microstate::add_body stores them); sites=None
        derives them as wrap(body + 0).  lds_tiles=False runs every trial from
        global memory (same decisions).  graphs=False makes run() launch every
        apply directly instead of replaying a captured hipGraph (same results).
        persistent=False keeps mixed-precision hard shapes on one launch per
        colour pass instead of one persistent launch per apply (same results;
        the persistent launch needs cooperative launch support and a region
        shape whose workgroups all fit on the device at once).
        fp32_error=True (mixed precision) also measures in FP64 -- LJ: every
        trial's dH; hard spheres: every pair within 1.01 sigma, on the
        per-colour tile passes -- and records the largest FP32 error and its
        ratio to the error bound (stats()["fp32_error_max"],
        ["fp32_error_bound"]).

        model: "hard_sphere" (params = (diameter,), default (diameter,)),
        "lj" (params = (epsilon, sigma, r_cut)) or "lj_shifted"
        (params = (epsilon, sigma, r_cut, r_shift)) or "convex" (hard convex
        polyhedra: vertices (k x 3), orientations (n x 4, s x y z; identity by
        default)) or "overlap_penalty" (the placement Hamiltonian
        Isotropic { Expanded { delta, OverlapPenalty }, r_cut: delta }:
        params = (k, maximum_allowed_overlap, epsilon_shoulder, delta),
        default OverlapPenalty::default() with delta 1; FP64 only); the
        ranges default to the model's interaction range (diameter, r_cut,
        twice the circumsphere radius or delta)."""
        pos = np.ascontiguousarray(positions, np.float64)
        self.dim = pos.shape[1]
        self.n = pos.shape[0]
        self.box_edges = [float(box_edges[a]) for a in range(self.dim)]
        cfg = Config()
        cfg.dim = self.dim
        cfg.n_bodies = self.n
        for a in range(self.dim):
            cfg.box_edges[a] = float(box_edges[a])
        model_id = {"hard_sphere": HMC_HARD_SPHERE, "lj": HMC_LJ, "lj_shifted": HMC_LJ_SHIFTED,
                    "convex": HMC_CONVEX_POLYHEDRON, "overlap_penalty": HMC_OVERLAP_PENALTY}[model]
        if params is None:
            params = {"hard_sphere": (diameter,),
                      "overlap_penalty": (10000.0, 0.2, 100.0, 1.0)}.get(model, (1.0, 1.0, 2.5, 2.5))
        if model == "convex":
            self._verts = np.ascontiguousarray(vertices, np.float64).reshape(-1, 3)
            cfg.vertices = _p(self._verts)
            cfg.n_vertices = len(self._verts)
            rng = 2.0 * float(np.sqrt(np.max(np.sum(self._verts**2, axis=1))))
            params = ()
        else:
            rng = {"hard_sphere": 0, "overlap_penalty": 3}.get(model, 2)
            rng = params[rng]
        cfg.boundary_range = rng if boundary_range is None else boundary_range
        cfg.body_interaction_range = rng if body_interaction_range is None else body_interaction_range
        # VecCell::with_search_radius(maximum interaction range), as the
        # reference's benchmarks build the microstate
        cfg.spatial_cell_width = rng if spatial_cell_width is None else spatial_cell_width
        cfg.model = model_id
        for i, v in enumerate(params):
            cfg.params[i] = float(v)
        cfg.precision = {"fp64": HMC_FP64, "mixed": HMC_MIXED}[precision]
        cfg.device = device
        cfg.flags = ((0 if lds_tiles else FLAG_GLOBAL_PASS) | (0 if graphs else FLAG_NO_GRAPH) |
                     (0 if persistent else FLAG_PER_COLOUR) | (FLAG_FP32_ERROR if fp32_error else 0))
        h = C.c_void_p()
        rc = lib().hmc_create(C.byref(cfg), C.byref(h))
        if rc != HMC_OK:
            raise HmcError(rc, "hmc_create failed")
        self.h = h
        _LIVE[next(_LIVE_ID)] = self
        self.set_positions(pos, sites)
        if orientations is not None:
            self.set_orientations(orientations, site_orientations)

    def _check(self, rc):
        if rc != HMC_OK:
            raise HmcError(rc, lib().hmc_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().hmc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def delta_energy_insert(self, bodies):
        """(energy, bound) arrays: PairwiseCutoff::delta_energy_insert
        (pairwise_cutoff.rs:1017-1025) of each candidate body position alone
        against the current state, with a bound on the summation-order
        difference."""
        b = np.ascontiguousarray(bodies, np.float64).reshape(-1, self.dim)
        e = np.zeros(len(b))
        bd = np.zeros(len(b))
        self._check(lib().hmc_delta_energy_insert(self.h, _p(b), len(b), _p(e), _p(bd)))
        return e, bd

    def clone_with_boundary(self, box_edges):
        """Microstate::clone_with_boundary(new_boundary, |_| true)
        (hoomd-microstate/src/microstate.rs:1529-1565) onto new box edges: a
        new GpuParallelSweep with every body mapped by Hypercuboid::map_point
        (hoomd-geometry/src/shape/cuboid.rs:553-567) on the device; this one is
        unchanged.  HmcError(HMC_ERR_INVALID) when the interaction ranges do
        not fit the new edges."""
        e = np.zeros(3)
        e[: self.dim] = [float(x) for x in box_edges[: self.dim]]
        h = C.c_void_p()
        self._check(lib().hmc_clone_with_boundary(self.h, _p(e), C.byref(h)))
        new = object.__new__(type(self))
        new.__dict__.update({k: v for k, v in self.__dict__.items() if k != "h"})
        new.h = h
        new.box_edges = [float(x) for x in e[: self.dim]]
        return new

    def apply(self, step, substep, seed, kT, max_distance):
        cnt = Count()
        used = C.c_uint32()
        self._check(lib().hmc_apply_translate(self.h, step, substep, seed, kT, max_distance, C.byref(cnt),
                                              C.byref(used)))
        return int(cnt.accepted), int(cnt.rejected), int(used.value)

    def apply_rotate(self, step, substep, seed, kT, max_rotation):
        cnt = Count()
        used = C.c_uint32()
        self._check(lib().hmc_apply_rotate(self.h, step, substep, seed, kT, max_rotation, C.byref(cnt),
                                           C.byref(used)))
        return int(cnt.accepted), int(cnt.rejected), int(used.value)

    def total_energy(self):
        """(energy, bound): PairwiseCutoff::total_energy of the current state and
        the bound on its difference from any other summation order."""
        e = C.c_double()
        b = C.c_double()
        self._check(lib().hmc_total_energy(self.h, C.byref(e), C.byref(b)))
        return e.value, b.value

    def total_energy_noalloc(self):
        """total_energy through hmc_total_energy_noalloc (no device allocation:
        the partials live in the library's static scratch)."""
        e = C.c_double()
        b = C.c_double()
        self._check(lib().hmc_total_energy_noalloc(self.h, C.byref(e), C.byref(b)))
        return e.value, b.value

    def sdf(self, xmax=0.02, dx=1e-4, nbins=None):
        """(hist, overlaps): the scale distribution function histogram of the
        current state (hmc_sdf), nbins = ceil(xmax / dx) bins by default."""
        nbins = sdf_nbins(xmax, dx) if nbins is None else int(nbins)
        hist = np.zeros(nbins, np.uint64)
        ov = C.c_uint64()
        self._check(lib().hmc_sdf(self.h, xmax, dx, nbins, hist.ctypes.data_as(C.POINTER(C.c_uint64)),
                                  C.byref(ov)))
        return hist, int(ov.value)

    def orientations(self):
        b = np.zeros((self.n, 4))
        s = np.zeros((self.n, 4))
        self._check(lib().hmc_get_orientations(self.h, _p(b), _p(s)))
        return b, s

    def set_orientations(self, body, site=None):
        body = np.ascontiguousarray(body, np.float64)
        site_p = None if site is None else _p(np.ascontiguousarray(site, np.float64))
        self._check(lib().hmc_set_orientations(self.h, _p(body), site_p))

    def run(self, step0, seed, kT, max_distance, n_steps):
        cnt = Count()
        self._check(lib().hmc_run_translate(self.h, step0, seed, kT, max_distance, n_steps, C.byref(cnt)))
        return int(cnt.accepted), int(cnt.rejected)

    def tune(self, move, step, substep, seed, kT, target, samples, steps, size):
        """ParallelSweep::tune (tune_local.rs:21-89) with move "translate" or
        (convex) "rotate" against the current state, from
        Counter(step, substep, seed).make_rng(): returns (the tuned maximum
        size, [(accepted, rejected)] per step, ambiguous samples).  The state
        is not modified."""
        m = C.c_double(size)
        cnt = (Count * max(int(steps), 1))()
        amb = C.c_uint64()
        mv = {"translate": HMC_MOVE_TRANSLATE, "rotate": HMC_MOVE_ROTATE}[move]
        self._check(lib().hmc_tune_local(self.h, mv, step, substep, seed, kT, target, int(samples), int(steps),
                                         C.byref(m), cnt, C.byref(amb)))
        return m.value, [(int(c.accepted), int(c.rejected)) for c in cnt[:steps]], int(amb.value)

    def tune_translate(self, step, substep, seed, kT, target, samples, steps, max_move):
        """tune("translate", ...)."""
        return self.tune("translate", step, substep, seed, kT, target, samples, steps, max_move)

    def grand_canonical(self, step, substep, seed, kT, fugacity):
        """GrandCanonical(UniformIn over the box)::apply (grand_canonical.rs:145-205)
        with (step, substep, seed) = the microstate's counter: returns
        (insert accepted, insert rejected, remove accepted, remove rejected,
        ambiguous).  An accepted move changes the body list (a new body last,
        or the last body moved to the removed index) and self.n; the caller
        then increments the substep."""
        cnt = GcCount()
        amb = C.c_uint64()
        rc = lib().hmc_grand_canonical_apply(C.byref(self.h), step, substep, seed, kT, fugacity, C.byref(cnt),
                                             C.byref(amb))
        if rc != HMC_OK:
            raise HmcError(rc, lib().hmc_last_error(self.h).decode())
        self.n += int(cnt.insert_accepted) - int(cnt.remove_accepted)
        return (int(cnt.insert_accepted), int(cnt.insert_rejected), int(cnt.remove_accepted),
                int(cnt.remove_rejected), int(amb.value))

    def run_moves(self, step0, seed, kT, moves, n_steps):
        """n_steps of {each (kind, size) of moves in order; increment_step}
        (hmc_run); kind is "translate" or "rotate".  Returns one (accepted,
        rejected) per move."""
        arr = (Move * len(moves))()
        for i, (kind, size) in enumerate(moves):
            arr[i].move = {"translate": HMC_MOVE_TRANSLATE, "rotate": HMC_MOVE_ROTATE}[kind]
            arr[i].size = size
        out = (Count * len(moves))()
        self._check(lib().hmc_run(self.h, step0, seed, kT, arr, len(moves), n_steps, out))
        return [(int(c.accepted), int(c.rejected)) for c in out]

    def debug_regions(self):
        """Mailbox words of the persistent apply's regions, (regions, 6):
        parity counters, probe counter, staged, passes published, waits passed."""
        st = self.stats()
        cap = 6 * max(st["regions"], 1)
        out = np.zeros(cap, np.uint32)
        n = C.c_uint64()
        self._check(lib().hmc_debug_regions(self.h, out.ctypes.data_as(C.POINTER(C.c_uint32)), cap, C.byref(n)))
        return out[: 6 * n.value].reshape(-1, 6)

    def moved_since(self, since):
        """(indices of the bodies moved by applies after epoch `since`, the
        current epoch) -- hmc_moved_since."""
        words = (self.n + 63) // 64
        mask = np.zeros(max(words, 1), np.uint64)
        now = C.c_uint32()
        nm = C.c_uint64()
        self._check(lib().hmc_moved_since(self.h, since, C.byref(now), mask.ctypes.data_as(C.POINTER(C.c_uint64)),
                                          C.byref(nm)))
        bits = np.unpackbits(mask.view(np.uint8), bitorder="little")[: self.n]
        idx = np.flatnonzero(bits).astype(np.uint32)
        assert len(idx) == nm.value
        return idx, int(now.value)

    def get_bodies(self, index, orientations=False):
        """Body and site positions (and, with orientations, body and site
        quaternions) of the listed bodies (hmc_get_bodies)."""
        idx = np.ascontiguousarray(index, np.uint32)
        n = len(idx)
        b = np.zeros((n, self.dim))
        s = np.zeros((n, self.dim))
        bq = np.zeros((n, 4)) if orientations else None
        sq = np.zeros((n, 4)) if orientations else None
        self._check(lib().hmc_get_bodies(self.h, idx.ctypes.data_as(C.POINTER(C.c_uint32)), n, _p(b), _p(s),
                                         None if bq is None else _p(bq), None if sq is None else _p(sq)))
        return (b, s, bq, sq) if orientations else (b, s)

    def positions(self):
        b = np.zeros((self.n, self.dim))
        s = np.zeros((self.n, self.dim))
        self._check(lib().hmc_get_state(self.h, _p(b), _p(s)))
        return b, s

    def set_positions(self, body, site=None):
        body = np.ascontiguousarray(body, np.float64)
        site_p = None if site is None else _p(np.ascontiguousarray(site, np.float64))
        self._check(lib().hmc_set_state(self.h, _p(body), site_p))

    def trace(self, on=True, capacity=1 << 22):
        self._check(lib().hmc_trace_enable(self.h, 1 if on else 0, capacity))

    def read_trace(self, capacity=1 << 22):
        out = np.zeros(capacity, TRACE_DTYPE)
        n = C.c_uint64()
        self._check(lib().hmc_trace_read(self.h, out.ctypes.data_as(C.c_void_p), capacity, C.byref(n)))
        return out[: n.value]

    def stats(self):
        s = Stats()
        self._check(lib().hmc_get_stats(self.h, C.byref(s)))
        out = {k: getattr(s, k) for k, _ in Stats._fields_}
        out["tile_shape"] = tuple(int(v) for v in s.tile_shape)
        return {k: v if k in ("tile_shape", "fp32_error_max", "fp32_error_bound") else int(v)
                for k, v in out.items()}

    def profile(self, on=True, period=1):
        """Time every period-th apply's launches with HIP events (on=False: off)."""
        self._check(lib().hmc_profile_enable(self.h, int(period) if on else 0))

    def read_profile(self):
        pm = np.zeros(1)
        bm = np.zeros(1)
        pn = C.c_uint64()
        bn = C.c_uint64()
        self._check(lib().hmc_profile_read(self.h, _p(pm), C.byref(pn), _p(bm), C.byref(bn)))
        return float(pm[0]), int(pn.value), float(bm[0]), int(bn.value)


def debug_math(kind, x, device=0):
    kinds = {"cbrt": 0, "log": 1, "exp": 2, "sqrt": 3, "div": 4, "sin": 5, "cos": 6}
    x = np.ascontiguousarray(x, np.float64)
    n = x.size // 2 if kind == "div" else x.size
    out = np.zeros(n)
    rc = lib().hmc_debug_math(device, kinds[kind], _p(x), _p(out), n)
    if rc != HMC_OK:
        raise HmcError(rc, "hmc_debug_math failed")
    return out


def debug_normals(step, substep, seed, index, n, device=0):
    out = np.zeros(n)
    rc = lib().hmc_debug_normals(device, step, substep, seed, index, _p(out), n)
    if rc != HMC_OK:
        raise HmcError(rc, "hmc_debug_normals failed")
    return out


def debug_convex(kind, vertices, rows, device=0):
    """Device convex routines on rows of inputs: "collide" (v_ij[3], q_ij[4]),
    "global" (r_i[3], q_i[4], r_j[3], q_j[4]), "sat" (as "global"; the mixed
    mode's FP32 class: 0 apart, 1 overlap, 2 left to FP64) or "rotate" (step,
    substep, seed, index, sigma)."""
    k = {"collide": 0, "global": 1, "rotate": 2, "sat": 3}[kind]
    v = np.ascontiguousarray(vertices, np.float64).reshape(-1, 3)
    x = np.ascontiguousarray(rows, np.float64)
    n = x.shape[0]
    out = np.zeros((n, 4) if k == 2 else n)
    rc = lib().hmc_debug_convex(device, k, _p(v), len(v), _p(x), _p(out), n)
    if rc != HMC_OK:
        raise HmcError(rc, "hmc_debug_convex failed")
    return out


def convex_axes(vertices):
    """The mixed mode's separating-axis data of a polytope (host code): dict
    of the counts (faces, edges, scanned vertices, sym, use, use_edges), the
    face axes (unit normal, hmin, hmax) and the edge directions, as FP32
    values."""
    v = np.ascontiguousarray(vertices, np.float64).reshape(-1, 3)
    cnt = (C.c_uint32 * 6)()
    faces = np.zeros((16, 5))
    edges = np.zeros((8, 3))
    rc = lib().hmc_debug_convex_axes(_p(v), len(v), cnt, _p(faces), _p(edges))
    if rc != HMC_OK:
        raise HmcError(rc, "hmc_debug_convex_axes failed")
    nf, ne, nh, sym, use, use_edges = (int(x) for x in cnt)
    return dict(n_faces=nf, n_edges=ne, n_scanned=nh, sym=bool(sym), use=bool(use), use_edges=bool(use_edges),
                faces=faces[:min(nf, 16)], edges=edges[:min(ne, 8)])


class GpuBatch:
    """Independent hard-sphere/disk replicas advanced together (one set of
    launches per step, replayed from a captured hipGraph)."""

    def __init__(self, systems, *, diameter=1.0, precision="fp64", device=0, lds_tiles=True, persistent=True):
        """systems: list of (box_edges, body_positions, site_positions or None).
        lds_tiles=False runs every pass a lane per space from global memory
        instead of on LDS tiles (same decisions).  persistent=False keeps a
        mixed-precision 3D batch on one launch per colour pass instead of
        persistent applies (one cooperative launch per group of replicas per
        step, not graph-captured; same decisions)."""
        self.R = len(systems)
        cfgs = (Config * self.R)()
        self.dims = []
        for r, (L, pos, _site) in enumerate(systems):
            pos = np.asarray(pos, np.float64)
            c = cfgs[r]
            c.dim = pos.shape[1]
            c.n_bodies = pos.shape[0]
            for a in range(c.dim):
                c.box_edges[a] = float(L[a])
            c.boundary_range = c.body_interaction_range = c.spatial_cell_width = diameter
            c.model = HMC_HARD_SPHERE
            c.params[0] = diameter
            c.precision = {"fp64": HMC_FP64, "mixed": HMC_MIXED}[precision]
            c.device = device
            c.flags = (0 if lds_tiles else FLAG_GLOBAL_PASS) | (0 if persistent else FLAG_PER_COLOUR)
            self.dims.append((pos.shape[0], pos.shape[1]))
        h = C.c_void_p()
        rc = lib().hmc_batch_create(self.R, cfgs, C.byref(h))
        if rc != HMC_OK:
            raise HmcError(rc, "hmc_batch_create failed")
        self.h = h
        _LIVE[next(_LIVE_ID)] = self
        for r, (_L, pos, site) in enumerate(systems):
            self.set_positions(r, pos, site)

    def _check(self, rc):
        if rc != HMC_OK:
            raise HmcError(rc, lib().hmc_batch_last_error(self.h).decode())

    def set_positions(self, r, body, site=None):
        body = np.ascontiguousarray(body, np.float64)
        site_p = None if site is None else _p(np.ascontiguousarray(site, np.float64))
        self._check(lib().hmc_batch_set_state(self.h, r, _p(body), site_p))

    def positions(self, r):
        n, d = self.dims[r]
        b = np.zeros((n, d))
        s = np.zeros((n, d))
        self._check(lib().hmc_batch_get_state(self.h, r, _p(b), _p(s)))
        return b, s

    def run(self, step0, seeds, kT, max_distance, n_steps, use_graph=True):
        st = (C.c_uint64 * self.R)(*[int(x) for x in np.broadcast_to(step0, (self.R,))])
        sd = (C.c_uint32 * self.R)(*[int(x) for x in np.broadcast_to(seeds, (self.R,))])
        md = np.ascontiguousarray(np.broadcast_to(np.asarray(max_distance, np.float64), (self.R,)))
        out = (Count * self.R)()
        self._check(lib().hmc_batch_run_translate(self.h, st, sd, kT, _p(md), n_steps, 1 if use_graph else 0, out))
        return [(int(c.accepted), int(c.rejected)) for c in out]

    def autotune(self, step0, seeds, kT, max_distance, n_steps=3):
        """Pick the fastest tile layout for these run parameters
        (hmc_batch_autotune; the batch's state is left as it was).
        Returns (ms per step, policy)."""
        st = (C.c_uint64 * self.R)(*[int(x) for x in np.broadcast_to(step0, (self.R,))])
        sd = (C.c_uint32 * self.R)(*[int(x) for x in np.broadcast_to(seeds, (self.R,))])
        md = np.ascontiguousarray(np.broadcast_to(np.asarray(max_distance, np.float64), (self.R,)))
        ms = np.zeros(1)
        pol = C.c_uint32()
        self._check(lib().hmc_batch_autotune(self.h, st, sd, kT, _p(md), n_steps, _p(ms), C.byref(pol)))
        return float(ms[0]), int(pol.value)

    def stats(self, r):
        s = Stats()
        self._check(lib().hmc_batch_get_stats(self.h, r, C.byref(s)))
        out = {k: getattr(s, k) if k.startswith("fp32") else int(getattr(s, k))
               for k, _ in Stats._fields_ if k != "tile_shape"}
        out["tile_shape"] = tuple(int(v) for v in s.tile_shape)
        return out

    def sdf(self, xmax=0.02, dx=1e-4, nbins=None):
        """(hist[R, nbins], overlaps[R]): hmc_sdf of every replica."""
        nbins = sdf_nbins(xmax, dx) if nbins is None else int(nbins)
        hist = np.zeros((self.R, nbins), np.uint64)
        ov = np.zeros(self.R, np.uint64)
        u64p = C.POINTER(C.c_uint64)
        self._check(lib().hmc_batch_sdf(self.h, xmax, dx, nbins, hist.ctypes.data_as(u64p), ov.ctypes.data_as(u64p)))
        return hist, ov

    def close(self):
        if getattr(self, "h", None):
            lib().hmc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class QuickCompressState(C.Structure):
    _fields_ = [("target_volume", C.c_double), ("maximum_energy_per_site", C.c_double),
                ("maximum_delta", C.c_double), ("state", C.c_int32), ("is_final", C.c_int32),
                ("boundary", C.c_double * 3), ("ambiguous", C.c_uint64)]


QC_STATES = ("startup", "waiting", "compressing", "complete")


class QuickCompress:
    """QuickCompress (hoomd-mc/src/quick_compress.rs:19-444) over a
    GpuParallelSweep: hmc_quick_compress_apply runs the state machine on the
    host, the trial microstate (hmc_clone_with_boundary) and the energies
    (hmc_total_energy) on the device."""

    def __init__(self, target_volume, maximum_energy_per_site=1000.0, maximum_delta=0.05):
        self.st = QuickCompressState()
        lib().hmc_quick_compress_init(C.byref(self.st), float(target_volume))
        self.st.maximum_energy_per_site = float(maximum_energy_per_site)
        self.st.maximum_delta = float(maximum_delta)

    def apply(self, sim, step, substep, seed):
        """QuickCompress::apply with (step, substep, seed) = the microstate's
        counter; it consumes one substep.  An accepted compression replaces
        sim's device state by the trial microstate (sim.box_edges follow)."""
        before = sim.h.value
        rc = lib().hmc_quick_compress_apply(C.byref(self.st), C.byref(sim.h), step, substep, seed)
        if rc != HMC_OK:
            raise HmcError(rc, lib().hmc_last_error(sim.h).decode())
        if sim.h.value != before:
            sim.box_edges = [float(self.st.boundary[a]) for a in range(sim.dim)]

    def is_complete(self):
        return self.st.state == 3

    @property
    def state(self):
        return QC_STATES[self.st.state]

    @property
    def ambiguous(self):
        return int(self.st.ambiguous)


class QuickInsertState(C.Structure):
    _fields_ = [("target", C.c_uint64), ("allowed_overlaps", C.c_uint64), ("inserted", C.c_uint64),
                ("state", C.c_int32), ("pad", C.c_int32), ("ambiguous", C.c_uint64)]


class QuickInsert:
    """QuickInsert (hoomd-mc/src/quick_insert.rs:94-334) with UniformIn over
    the periodic box, over a GpuParallelSweep: the candidates' insertion
    energies against the current bodies on the device (one launch per call),
    the candidates taken in order on the host (hmc_quick_insert_apply)."""

    def __init__(self, target):
        self.st = QuickInsertState()
        lib().hmc_quick_insert_init(C.byref(self.st), int(target))

    def apply(self, sim, step, substep, seed):
        """QuickInsert::apply with (step, substep, seed) = the microstate's
        counter.  Returns (accepted, rejected, substeps used: 0 or 1).  The
        inserted bodies follow sim's bodies in order (sim.n grows)."""
        cnt = Count()
        rc = lib().hmc_quick_insert_apply(C.byref(self.st), C.byref(sim.h), step, substep, seed, C.byref(cnt))
        if rc != HMC_OK:
            raise HmcError(rc, lib().hmc_last_error(sim.h).decode())
        sim.n += int(cnt.accepted)
        return int(cnt.accepted), int(cnt.rejected), 0 if self.is_complete() else 1

    def is_complete(self):
        return self.st.state == 1

    @property
    def inserted(self):
        return int(self.st.inserted)

    @property
    def target(self):
        return int(self.st.target)

    @property
    def ambiguous(self):
        return int(self.st.ambiguous)


GSD_TYPES = {1: np.uint8, 2: np.uint16, 3: np.uint32, 4: np.uint64, 5: np.int8, 6: np.int16, 7: np.int32,
             8: np.int64, 9: np.float32, 10: np.float64, 11: np.uint8}


class GsdFile:
    """A GSD 2.1 file (hoomd-gsd/src/file_layer.rs) through hmc_gsd_*:
    GsdFile.create(path) makes a HOOMD trajectory (schema "hoomd" 1.4, as
    HoomdGsdFile::create), GsdFile.open(path, writable) opens one.
    append_frame(sim, step) writes a GpuParallelSweep's state as a HOOMD
    frame; write_chunk / end_frame write any chunks; read(frame, name)
    returns a chunk as an array of shape (rows, columns)."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def create(cls, path, application="hoomd-rs_amd", schema="hoomd", schema_version=(1, 4)):
        h = C.c_void_p()
        rc = lib().hmc_gsd_create(str(path).encode(), application.encode(), schema.encode(),
                                  schema_version[0], schema_version[1], C.byref(h))
        if rc != HMC_OK:
            raise HmcError(rc, f"cannot create {path}")
        return cls(h)

    @classmethod
    def open(cls, path, writable=False):
        h = C.c_void_p()
        rc = lib().hmc_gsd_open(str(path).encode(), 1 if writable else 0, C.byref(h))
        if rc != HMC_OK:
            raise HmcError(rc, f"cannot open {path}")
        return cls(h)

    def write_chunk(self, name, array):
        a = np.ascontiguousarray(array)
        if a.ndim == 0:
            a = a.reshape(1, 1)
        elif a.ndim == 1:
            a = a.reshape(-1, 1)
        code = {v: k for k, v in GSD_TYPES.items() if k != 11}[a.dtype.type]
        rc = lib().hmc_gsd_write_chunk(self.h, name.encode(), code, a.shape[0], a.shape[1], a.ctypes.data)
        if rc != HMC_OK:
            raise HmcError(rc, f"cannot write chunk {name}")

    def end_frame(self):
        rc = lib().hmc_gsd_end_frame(self.h)
        if rc != HMC_OK:
            raise HmcError(rc, "cannot end the frame")

    def append_frame(self, sim, step):
        rc = lib().hmc_gsd_append_frame(self.h, sim.h, step)
        if rc != HMC_OK:
            raise HmcError(rc, lib().hmc_last_error(sim.h).decode())

    @property
    def n_frames(self):
        return int(lib().hmc_gsd_n_frames(self.h))

    def header(self):
        app = C.create_string_buffer(64)
        sch = C.create_string_buffer(64)
        ma, mi = C.c_uint32(), C.c_uint32()
        lib().hmc_gsd_header(self.h, app, sch, C.byref(ma), C.byref(mi))
        return app.value.decode(), sch.value.decode(), (ma.value, mi.value)

    def read(self, frame, name):
        rows, cols, typ = C.c_uint64(), C.c_uint32(), C.c_uint8()
        rc = lib().hmc_gsd_read_chunk(self.h, frame, name.encode(), None, 0, C.byref(rows), C.byref(cols),
                                      C.byref(typ))
        if rc != HMC_OK:
            raise KeyError(f"chunk {name} not in frame {frame}")
        out = np.zeros((rows.value, cols.value), GSD_TYPES[typ.value])
        rc = lib().hmc_gsd_read_chunk(self.h, frame, name.encode(), out.ctypes.data, out.nbytes, C.byref(rows),
                                      C.byref(cols), C.byref(typ))
        if rc != HMC_OK:
            raise HmcError(rc, f"cannot read chunk {name}")
        return out

    def close(self):
        if self.h:
            rc = lib().hmc_gsd_close(self.h)
            self.h = None
            if rc != HMC_OK:
                raise HmcError(rc, "closing the file failed")


def scale_volume(box_edges, v):
    """Edges of a hypercuboid whose volume is scaled by v
    (Hypercuboid::scale_volume, hoomd-geometry/src/shape/cuboid.rs:514-518):
    every edge times v^(1/D), as the reference forms it."""
    f = float(v) ** (1.0 / len(box_edges))
    return [float(e) * f for e in box_edges]


def sdf_nbins(xmax, dx):
    """Bins of width dx covering [0, xmax)."""
    return int(np.ceil(xmax / dx))


def sdf_beta_p_over_rho(hist, n_bodies, dx, dim, n_frames=1, degree=5):
    """beta P / rho = 1 + s(0+) / (2 dim) from SDF counts summed over n_frames
    frames of n_bodies bodies, as HOOMD-blue's hpmc compute.SDF.betaP: s =
    counts / (n_frames n_bodies dx) at the bin centres, a least-squares
    polynomial of the given degree through them, evaluated at x = 0."""
    hist = np.asarray(hist, np.float64)
    s = hist / (n_frames * n_bodies * dx)
    x = (np.arange(len(hist)) + 0.5) * dx
    s0 = np.polyval(np.polyfit(x, s, degree), 0.0)
    return 1.0 + s0 / (2.0 * dim)
