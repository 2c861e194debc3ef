"""ctypes view of the oracle (oracle/liboracle.so), the CPU restatement of the
reference sweep.  Test infrastructure only."""
import ctypes as C
import pathlib
import subprocess

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent.parent
LIB_PATH = ROOT / "oracle" / "liboracle.so"

_lib = None

D_P = C.POINTER(C.c_double)
U64_P = C.POINTER(C.c_uint64)
I64_P = C.POINTER(C.c_int64)
U32_P = C.POINTER(C.c_uint32)


def _p(a, t):
    return a.ctypes.data_as(t)


class TraceRecord(C.Structure):
    _fields_ = [("substep", C.c_uint32), ("color", C.c_uint32), ("space", C.c_uint64),
                ("body", C.c_int64), ("status", C.c_int32), ("pad", C.c_int32)]


TRACE_DTYPE = np.dtype([("substep", "<u4"), ("color", "<u4"), ("space", "<u8"),
                        ("body", "<i8"), ("status", "<i4"), ("pad", "<i4")])


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            subprocess.run(["make", "-C", str(ROOT / "oracle")], check=True,
                           stdout=subprocess.DEVNULL)
        L = C.CDLL(str(LIB_PATH))
        L.oracle_sfc64_stream.argtypes = [C.c_uint64] * 4 + [C.c_int, U64_P, C.c_uint64]
        L.oracle_counter_stream.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64, U64_P, C.c_uint64]
        L.oracle_sample.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64, C.c_int, C.c_uint32,
                                    D_P, C.c_uint64]
        L.oracle_zig_tables.argtypes = [D_P, D_P]
        L.oracle_translate.argtypes = [C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32,
                                       D_P, C.c_double, D_P, U32_P]
        L.oracle_wrap.argtypes = [C.c_int, D_P, D_P, D_P, C.c_uint64]
        L.oracle_ghosts.argtypes = [C.c_int, D_P, C.c_double, D_P, D_P]
        L.oracle_ghosts.restype = C.c_int
        L.oracle_compute_dimensions.argtypes = [C.c_int, D_P, C.c_double, C.POINTER(C.c_int), D_P, I64_P]
        L.oracle_compute_dimensions.restype = C.c_int
        L.oracle_multi_to_index.argtypes = [C.c_int, I64_P, I64_P]
        L.oracle_multi_to_index.restype = C.c_uint64
        L.oracle_by_color.argtypes = [C.c_int, I64_P, U64_P]
        L.oracle_point_to_space_fixed.argtypes = [C.c_int, C.c_double, D_P, C.c_int, D_P, I64_P, C.c_uint64]
        L.oracle_checkerboard_origin.argtypes = [C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, C.c_double,
                                                 D_P, D_P, D_P, I64_P]
        L.oracle_stencil.argtypes = [C.c_int, C.c_uint32, I64_P]
        L.oracle_stencil.restype = C.c_uint64
        L.oracle_cell_index.argtypes = [C.c_int, C.c_double, D_P, I64_P]
        L.oracle_system_create.argtypes = [C.c_int, D_P, C.c_double, C.c_double, C.c_int, D_P,
                                           C.c_uint64, C.c_uint32, D_P, C.c_uint64]
        L.oracle_system_create.restype = C.c_void_p
        L.oracle_system_destroy.argtypes = [C.c_void_p]
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_max_threads.restype = C.c_int
        L.oracle_parallel_apply.argtypes = [C.c_void_p, C.c_double, C.c_double, U64_P, U64_P]
        L.oracle_parallel_apply.restype = C.c_int
        L.oracle_sweep_apply.argtypes = [C.c_void_p, C.c_double, C.c_double, U64_P, U64_P]
        L.oracle_increment_step.argtypes = [C.c_void_p]
        L.oracle_get_counter.argtypes = [C.c_void_p, U64_P, U32_P]
        L.oracle_set_counter.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32]
        L.oracle_n_ghosts.argtypes = [C.c_void_p]
        L.oracle_n_ghosts.restype = C.c_uint64
        L.oracle_get_positions.argtypes = [C.c_void_p, D_P, D_P]
        L.oracle_total_energy.argtypes = [C.c_void_p]
        L.oracle_total_energy.restype = C.c_double
        L.oracle_delta_energy_one.argtypes = [C.c_void_p, C.c_uint64, D_P]
        L.oracle_delta_energy_one.restype = C.c_double
        L.oracle_quick_insert_apply.argtypes = [C.c_void_p, U64_P, U64_P, U64_P]
        L.oracle_quick_insert_apply.restype = None
        L.oracle_tune_translate.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_uint64, C.c_uint32, D_P, U64_P]
        L.oracle_tune_translate.restype = None
        L.oracle_grand_canonical_apply.argtypes = [C.c_void_p, C.c_double, C.c_double, U64_P]
        L.oracle_grand_canonical_apply.restype = None
        L.oracle_n_bodies.argtypes = [C.c_void_p]
        L.oracle_n_bodies.restype = C.c_uint64
        L.oracle_delta_energy_insert.argtypes = [C.c_void_p, D_P]
        L.oracle_delta_energy_insert.restype = C.c_double
        L.oracle_trace_enable.argtypes = [C.c_void_p, C.c_int]
        L.oracle_trace_size.argtypes = [C.c_void_p]
        L.oracle_trace_size.restype = C.c_uint64
        L.oracle_trace_read.argtypes = [C.c_void_p, C.c_void_p]
        L.oracle_set_external_harmonic.argtypes = [C.c_void_p, C.c_double, D_P]
        L.oracle_run_accumulate.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_uint64, D_P, D_P]
        L.oracle_time_applies.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, D_P, U64_P]
        L.oracle_time_applies.restype = C.c_uint64
        _lib = L
    return _lib


def sfc64_stream(a, b, c, ctr, n, warm=True):
    out = np.zeros(n, np.uint64)
    lib().oracle_sfc64_stream(a, b, c, ctr, 1 if warm else 0, _p(out, U64_P), n)
    return out


def counter_stream(step, substep, seed, index, n):
    out = np.zeros(n, np.uint64)
    lib().oracle_counter_stream(step, substep, seed, index, _p(out, U64_P), n)
    return out


def sample(step, substep, seed, index, kind, n, param=0):
    out = np.zeros(n, np.float64)
    kinds = {"f64": 0, "open01": 1, "normal": 2, "index": 3}
    lib().oracle_sample(step, substep, seed, index, kinds[kind], param, _p(out, D_P), n)
    return out


def zig_tables():
    x = np.zeros(257)
    f = np.zeros(257)
    lib().oracle_zig_tables(_p(x, D_P), _p(f, D_P))
    return x, f


def wrap(L, pts):
    L = np.ascontiguousarray(L, np.float64)
    pts = np.ascontiguousarray(pts, np.float64)
    out = np.zeros_like(pts)
    dim = len(L)
    lib().oracle_wrap(dim, _p(L, D_P), _p(pts, D_P), _p(out, D_P), pts.size // dim)
    return out


def ghosts(L, rng, pt):
    L = np.ascontiguousarray(L, np.float64)
    pt = np.ascontiguousarray(pt, np.float64)
    dim = len(L)
    out = np.zeros((8, dim))
    n = lib().oracle_ghosts(dim, _p(L, D_P), rng, _p(pt, D_P), _p(out, D_P))
    return out[:n]


def compute_dimensions(edges, rng, periodic):
    edges = np.ascontiguousarray(edges, np.float64)
    dim = len(edges)
    w = np.zeros(dim)
    s = np.zeros(dim, np.int64)
    per = (C.c_int * 1)(1 if periodic else 0)
    rc = lib().oracle_compute_dimensions(dim, _p(edges, D_P), rng, per, _p(w, D_P), _p(s, I64_P))
    if rc != 0:
        return None
    return w, s


def multi_to_index(m, shape):
    m = np.ascontiguousarray(m, np.int64)
    shape = np.ascontiguousarray(shape, np.int64)
    return int(lib().oracle_multi_to_index(len(m), _p(m, I64_P), _p(shape, I64_P)))


def by_color(shape):
    shape = np.ascontiguousarray(shape, np.int64)
    dim = len(shape)
    total = int(np.prod(shape // 2))
    out = np.zeros((2**dim, total), np.uint64)
    lib().oracle_by_color(dim, _p(shape, I64_P), _p(out, U64_P))
    return out


def point_to_space_fixed(rng, edges, periodic, pts):
    edges = np.ascontiguousarray(edges, np.float64)
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, len(edges))
    out = np.zeros(len(pts), np.int64)
    lib().oracle_point_to_space_fixed(len(edges), rng, _p(edges, D_P), 1 if periodic else 0,
                                      _p(pts, D_P), _p(out, I64_P), len(pts))
    return out


def checkerboard(dim, step, substep, seed, rng, edges):
    edges = np.ascontiguousarray(edges, np.float64)
    o = np.zeros(dim)
    w = np.zeros(dim)
    s = np.zeros(dim, np.int64)
    lib().oracle_checkerboard_origin(dim, step, substep, seed, rng, _p(edges, D_P), _p(o, D_P),
                                     _p(w, D_P), _p(s, I64_P))
    return o, w, s


def stencil(dim, radius):
    out = np.zeros((2 * radius + 1) ** dim * dim, np.int64)
    n = lib().oracle_stencil(dim, radius, _p(out, I64_P))
    return out[: n * dim].reshape(n, dim)


def cell_index(w, p):
    p = np.ascontiguousarray(p, np.float64)
    out = np.zeros(len(p), np.int64)
    lib().oracle_cell_index(len(p), w, _p(p, D_P), _p(out, I64_P))
    return out


class OracleSystem:
    """The restated Microstate + PairwiseCutoff + ParallelSweep/Sweep."""

    MODELS = {"hard_sphere": 0, "lj": 1, "lj_shifted": 2, "overlap_penalty": 4}

    def __init__(self, L, positions, *, model="hard_sphere", params=(1.0,), boundary_range=None,
                 body_range=None, step=0, seed=0):
        L = np.ascontiguousarray(L, np.float64)
        self.dim = len(L)
        pos = np.ascontiguousarray(positions, np.float64).reshape(-1, self.dim)
        self.n = len(pos)
        self.seed = seed
        prm = np.zeros(4)
        prm[: len(params)] = params
        rng = {"hard_sphere": params[0], "overlap_penalty": prm[3]}.get(model, prm[2])
        boundary_range = rng if boundary_range is None else boundary_range
        body_range = rng if body_range is None else body_range
        self._keep = (L, pos, prm)
        self.h = lib().oracle_system_create(self.dim, _p(L, D_P), boundary_range, body_range,
                                            self.MODELS[model], _p(prm, D_P), step, seed,
                                            _p(pos, D_P), self.n)
        if not self.h:
            raise ValueError("invalid oracle system configuration")

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_system_destroy(self.h)
            self.h = None

    def parallel_apply(self, kt, max_move):
        a = np.zeros(1, np.uint64)
        r = np.zeros(1, np.uint64)
        rc = lib().oracle_parallel_apply(self.h, kt, max_move, _p(a, U64_P), _p(r, U64_P))
        if rc != 0:
            raise RuntimeError("checkerboard does not cover the system")
        return int(a[0]), int(r[0])

    def sweep_apply(self, kt, max_move):
        a = np.zeros(1, np.uint64)
        r = np.zeros(1, np.uint64)
        lib().oracle_sweep_apply(self.h, kt, max_move, _p(a, U64_P), _p(r, U64_P))
        return int(a[0]), int(r[0])

    def increment_step(self):
        lib().oracle_increment_step(self.h)

    def counter(self):
        s = np.zeros(1, np.uint64)
        ss = np.zeros(1, np.uint32)
        lib().oracle_get_counter(self.h, _p(s, U64_P), _p(ss, U32_P))
        return int(s[0]), int(ss[0])

    def set_counter(self, step, substep):
        lib().oracle_set_counter(self.h, step, substep)

    def positions(self):
        b = np.zeros((self.n, self.dim))
        s = np.zeros((self.n, self.dim))
        lib().oracle_get_positions(self.h, _p(b, D_P), _p(s, D_P))
        return b, s

    def n_ghosts(self):
        return int(lib().oracle_n_ghosts(self.h))

    def total_energy(self):
        return lib().oracle_total_energy(self.h)

    def delta_energy_one(self, body, trial):
        t = np.ascontiguousarray(trial, np.float64)
        return lib().oracle_delta_energy_one(self.h, body, _p(t, D_P))

    def quick_insert(self, qi):
        """QuickInsert::apply (quick_insert.rs:286-334) with UniformIn over the
        box; qi is the state [target, allowed_overlaps, inserted, complete]
        (np.uint64, updated in place).  Returns (accepted, rejected); the
        substep advances by one."""
        acc = np.zeros(1, np.uint64)
        rej = np.zeros(1, np.uint64)
        lib().oracle_quick_insert_apply(self.h, _p(qi, U64_P), _p(acc, U64_P), _p(rej, U64_P))
        self.n = int(lib().oracle_n_bodies(self.h))
        return int(acc[0]), int(rej[0])

    def tune_translate(self, kt, target, samples, steps, max_move):
        """Sweep::tune / ParallelSweep::tune with Translate (tune_local.rs:21-89):
        returns (the tuned maximum distance, [(accepted, rejected)] per step).
        The microstate and its counter are unchanged."""
        m = np.array([max_move], np.float64)
        cnt = np.zeros(2 * steps, np.uint64)
        lib().oracle_tune_translate(self.h, kt, target, samples, steps, _p(m, D_P), _p(cnt, U64_P))
        return float(m[0]), [(int(cnt[2 * t]), int(cnt[2 * t + 1])) for t in range(steps)]

    def grand_canonical(self, kt, fugacity):
        """GrandCanonical(UniformIn over the box)::apply (grand_canonical.rs:145-205):
        returns (insert accepted, insert rejected, remove accepted, remove
        rejected); the substep advances by one."""
        cnt = np.zeros(4, np.uint64)
        lib().oracle_grand_canonical_apply(self.h, kt, fugacity, _p(cnt, U64_P))
        self.n = int(lib().oracle_n_bodies(self.h))
        return tuple(int(x) for x in cnt)

    def delta_energy_insert(self, body):
        """PairwiseCutoff::delta_energy_insert of a new body at `body`."""
        t = np.ascontiguousarray(body, np.float64)
        return lib().oracle_delta_energy_insert(self.h, _p(t, D_P))

    def trace(self, on=True):
        lib().oracle_trace_enable(self.h, 1 if on else 0)

    def read_trace(self):
        n = lib().oracle_trace_size(self.h)
        out = np.zeros(n, TRACE_DTYPE)
        lib().oracle_trace_read(self.h, out.ctypes.data_as(C.c_void_p))
        return out

    def set_external_harmonic(self, k, centre):
        """Adds External(Harmonic): 1/2 k |x - centre|^2 per site."""
        c = np.ascontiguousarray(centre, np.float64)
        lib().oracle_set_external_harmonic(self.h, k, _p(c, D_P))

    def run_accumulate(self, parallel, kt, max_move, n):
        """n steps of apply + increment_step: (sum of body 0's position, sum of
        the external energy)."""
        ps = np.zeros(self.dim)
        es = np.zeros(1)
        lib().oracle_run_accumulate(self.h, 1 if parallel else 0, kt, max_move, n, _p(ps, D_P), _p(es, D_P))
        return ps, float(es[0])

    def time_applies(self, parallel, kt, max_move, seconds):
        el = np.zeros(1)
        na = np.zeros(1, np.uint64)
        moves = lib().oracle_time_applies(self.h, 1 if parallel else 0, kt, max_move, seconds,
                                          _p(el, D_P), _p(na, U64_P))
        return int(moves), float(el[0]), int(na[0])


def set_threads(n):
    lib().oracle_set_threads(n)


# ---------------------------------------------------------- convex polyhedra --
CUBE = np.array([[x, y, z] for z in (-0.5, 0.5) for y in (-0.5, 0.5) for x in (-0.5, 0.5)], np.float64)
CUBE = CUBE[[0, 1, 3, 2, 4, 5, 7, 6]]  # the reference fixture's vertex order


def _convex_lib():
    L = lib()
    if not getattr(L, "_convex_ready", False):
        L.oracle_convex_collide.argtypes = [D_P, C.c_uint32, D_P, D_P]
        L.oracle_convex_collide.restype = C.c_int
        L.oracle_convex_intersects_global.argtypes = [D_P, C.c_uint32, D_P, D_P, D_P, D_P]
        L.oracle_convex_intersects_global.restype = C.c_int
        L.oracle_quat_op.argtypes = [C.c_int, D_P, D_P, D_P]
        L.oracle_versor_displacement.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64, C.c_double,
                                                 D_P, C.c_uint64]
        L.oracle_convex_create.argtypes = [D_P, C.c_double, C.c_double, D_P, C.c_uint32, C.c_uint64, C.c_uint32,
                                           D_P, D_P, C.c_uint64]
        L.oracle_convex_create.restype = C.c_void_p
        L.oracle_convex_destroy.argtypes = [C.c_void_p]
        L.oracle_convex_apply.argtypes = [C.c_void_p, C.c_int, C.c_double, U64_P, U64_P]
        L.oracle_convex_tune.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_uint64, C.c_uint32, D_P, U64_P]
        L.oracle_convex_tune.restype = None
        L.oracle_convex_apply.restype = C.c_int
        L.oracle_convex_increment_step.argtypes = [C.c_void_p]
        L.oracle_convex_get_counter.argtypes = [C.c_void_p, U64_P, U32_P]
        L.oracle_convex_state.argtypes = [C.c_void_p, D_P, D_P, D_P, D_P]
        L.oracle_convex_any_overlap.argtypes = [C.c_void_p]
        L.oracle_convex_any_overlap.restype = C.c_int
        L.oracle_convex_trace_enable.argtypes = [C.c_void_p, C.c_int]
        L.oracle_convex_trace_size.argtypes = [C.c_void_p]
        L.oracle_convex_trace_size.restype = C.c_uint64
        L.oracle_convex_trace_read.argtypes = [C.c_void_p, C.c_void_p]
        L.oracle_convex_time.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, D_P, U64_P]
        L.oracle_convex_time.restype = C.c_uint64
        L._convex_ready = True
    return L


def _d(a, n=None):
    a = np.ascontiguousarray(a, np.float64)
    return a


def convex_collide(verts, vij, qij):
    v = _d(verts)
    return bool(_convex_lib().oracle_convex_collide(_p(v, D_P), len(v), _p(_d(vij), D_P), _p(_d(qij), D_P)))


def convex_intersects_global(verts, ri, qi, rj, qj):
    v = _d(verts)
    args = [_p(_d(x), D_P) for x in (ri, qi, rj, qj)]
    return bool(_convex_lib().oracle_convex_intersects_global(_p(v, D_P), len(v), *args))


def quat_op(op, a, b):
    out = np.zeros(4)
    _convex_lib().oracle_quat_op(op, _p(_d(a), D_P), _p(_d(b), D_P), _p(out, D_P))
    return out


def versor_displacement(step, substep, seed, index, sigma, n):
    out = np.zeros((n, 4))
    _convex_lib().oracle_versor_displacement(step, substep, seed, index, sigma, _p(out, D_P), n)
    return out


LIBM_KINDS = {"cbrt": 0, "log": 1, "exp": 2, "sin": 5, "cos": 6}


def libm_eval(kind, x):
    """The host C library's cbrt / log / exp, or the sine or cosine of its
    sincos, on every element of x (the functions the reference calls)."""
    L = lib()
    if not getattr(L, "_libm_ready", False):
        L.oracle_libm_eval.argtypes = [C.c_int, D_P, D_P, C.c_uint64]
        L._libm_ready = True
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty_like(x)
    L.oracle_libm_eval(LIBM_KINDS[kind], _p(x, D_P), _p(out, D_P), x.size)
    return out


def sdf(L, positions, sigma, xmax, dx, nbins, all_pairs=False):
    """oracle/sdf.hpp: (hist, overlaps) of one configuration of site positions;
    all_pairs selects the all-pairs neighbour search instead of the cell grid."""
    L = _d(L)
    dim = len(L)
    pos = _d(positions).reshape(-1, dim)
    hist = np.zeros(int(nbins), np.uint64)
    ov = np.zeros(1, np.uint64)
    f = lib().oracle_sdf
    f.argtypes = [C.c_int, D_P, D_P, C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_uint32, C.c_int,
                  U64_P, U64_P]
    f.restype = C.c_int
    rc = f(dim, _p(L, D_P), _p(pos, D_P), len(pos), sigma, xmax, dx, int(nbins), 1 if all_pairs else 0,
           _p(hist, U64_P), _p(ov, U64_P))
    if rc:
        raise ValueError("invalid SDF arguments")
    return hist, int(ov[0])


class ConvexOracle:
    """The restated Microstate<OrientedPoint> + HardShape<Convex> + ParallelSweep."""

    def __init__(self, L, positions, orientations, verts, *, boundary_range=None, body_range=None, step=0, seed=0):
        self.L = _d(L)
        pos = _d(positions).reshape(-1, 3)
        q = _d(orientations).reshape(-1, 4)
        v = _d(verts).reshape(-1, 3)
        r = float(np.sqrt(np.max(np.sum(v * v, axis=1)))) * 2.0
        self.range = r
        self.n = len(pos)
        self._keep = (pos, q, v)
        self.h = _convex_lib().oracle_convex_create(_p(self.L, D_P), r if boundary_range is None else boundary_range,
                                                    r if body_range is None else body_range, _p(v, D_P), len(v),
                                                    step, seed, _p(pos, D_P), _p(q, D_P), self.n)
        if not self.h:
            raise ValueError("invalid convex oracle configuration")

    def __del__(self):
        if getattr(self, "h", None):
            _convex_lib().oracle_convex_destroy(self.h)
            self.h = None

    def apply(self, rotate, max_move):
        a = np.zeros(1, np.uint64)
        r = np.zeros(1, np.uint64)
        rc = _convex_lib().oracle_convex_apply(self.h, 1 if rotate else 0, max_move, _p(a, U64_P), _p(r, U64_P))
        if rc:
            raise RuntimeError("a body lies outside the checkerboard")
        return int(a[0]), int(r[0])

    def increment_step(self):
        _convex_lib().oracle_convex_increment_step(self.h)

    def counter(self):
        s = np.zeros(1, np.uint64)
        ss = np.zeros(1, np.uint32)
        _convex_lib().oracle_convex_get_counter(self.h, _p(s, U64_P), _p(ss, U32_P))
        return int(s[0]), int(ss[0])

    def state(self):
        bp = np.zeros((self.n, 3))
        bq = np.zeros((self.n, 4))
        sp = np.zeros((self.n, 3))
        sq = np.zeros((self.n, 4))
        _convex_lib().oracle_convex_state(self.h, _p(bp, D_P), _p(bq, D_P), _p(sp, D_P), _p(sq, D_P))
        return bp, bq, sp, sq

    def any_overlap(self):
        return bool(_convex_lib().oracle_convex_any_overlap(self.h))

    def tune(self, rotate, target, samples, steps, size):
        """ParallelSweep::tune with Translate or Rotate<Versor> (tune_local.rs:21-89):
        returns (the tuned size, [(accepted, rejected)] per step)."""
        m = np.array([size], np.float64)
        cnt = np.zeros(2 * steps, np.uint64)
        _convex_lib().oracle_convex_tune(self.h, 1 if rotate else 0, target, samples, steps, _p(m, D_P),
                                         _p(cnt, U64_P))
        return float(m[0]), [(int(cnt[2 * t]), int(cnt[2 * t + 1])) for t in range(steps)]

    def trace(self, on=True):
        _convex_lib().oracle_convex_trace_enable(self.h, 1 if on else 0)

    def read_trace(self):
        n = _convex_lib().oracle_convex_trace_size(self.h)
        out = np.zeros(n, TRACE_DTYPE)
        _convex_lib().oracle_convex_trace_read(self.h, out.ctypes.data)
        return out

    def time_steps(self, d, a, seconds):
        el = np.zeros(1)
        st = np.zeros(1, np.uint64)
        moves = _convex_lib().oracle_convex_time(self.h, d, a, seconds, _p(el, D_P), _p(st, U64_P))
        return int(moves), float(el[0]), int(st[0])


def map_points(L_old, L_new, pts):
    """Hypercuboid::map_point of every point (oracle_map_points): (mapped,
    number outside the old box, left unmapped)."""
    L_old = np.ascontiguousarray(L_old, np.float64)
    L_new = np.ascontiguousarray(L_new, np.float64)
    dim = len(L_old)
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, dim)
    out = np.zeros_like(pts)
    f = lib().oracle_map_points
    f.argtypes = [C.c_int, D_P, D_P, D_P, C.c_uint64, D_P]
    f.restype = C.c_uint64
    n_out = f(dim, _p(L_old, D_P), _p(L_new, D_P), _p(pts, D_P), len(pts), _p(out, D_P))
    return out, int(n_out)


def uniform_f64(bits, low, high):
    """rand 0.10 Uniform::new(low, high).sample for f64 (a third-party crate,
    not in the reference tree; parity unpinned): scale = high - low, lowered
    one ulp at a time while low + scale (1 - 2^-52) >= high; a sample is
    (float with mantissa bits >> 12 in [1, 2) - 1) scale + low."""
    import math
    import struct
    scale = high - low
    while scale * (1.0 - 2.0**-52) + low >= high:
        scale = math.nextafter(scale, 0.0)
    v12 = struct.unpack("<d", struct.pack("<Q", (int(bits) >> 12) | 0x3FF0000000000000))[0]
    return (v12 - 1.0) * scale + low


class OracleQuickCompress:
    """QuickCompress::apply (hoomd-mc/src/quick_compress.rs:19-444) over
    OracleSystem microstates: the trial microstate is a new OracleSystem on
    the scaled edges (Hypercuboid::scale_volume, cuboid.rs:514-518) holding
    the bodies mapped by map_points (clone_with_boundary, microstate.rs:
    1529-1565); delta_energy_total follows pairwise_cutoff.rs:793-840."""

    def __init__(self, target_volume, maximum_energy_per_site=1000.0, maximum_delta=0.05):
        self.target = float(target_volume)
        self.max_e = float(maximum_energy_per_site)
        self.max_delta = float(maximum_delta)
        self.state = "startup"
        self.is_final = False
        self.boundary = None

    def apply(self, sysm, L, seed, model="hard_sphere", params=(1.0,)):
        """Returns the (possibly new) system and its edges; the counter's
        substep advances by one."""
        step, sub = sysm.counter()
        bits = counter_stream(step, sub, seed, 0, 1)[0]
        sysm.set_counter(step, sub + 1)
        L = [float(x) for x in L]
        if self.state == "startup":
            self.state, self.is_final, self.boundary = "waiting", False, L
        elif self.state == "complete":
            if self.boundary != L:
                self.state, self.is_final, self.boundary = "waiting", False, L
        elif self.state == "waiting":
            if sysm.total_energy() <= 0.0:
                if self.is_final and self.boundary == L:
                    self.state, self.boundary = "complete", L
                else:
                    self.state = "compressing"
        else:
            delta = uniform_f64(bits, 0.0, self.max_delta)
            current = 1.0
            for x in L:
                current *= x
            if self.target > current:
                trial = min(current * (1.0 + delta), self.target)
            else:
                trial = max(current * (1.0 - delta), self.target)
            f = (trial / current) ** (1.0 / len(L))
            L2 = [x * f for x in L]
            body, _site = sysm.positions()
            mapped, n_out = map_points(L, L2, body)
            assert n_out == 0
            new = OracleSystem(L2, mapped, model=model, params=params, seed=sysm.seed, step=step)
            new.set_counter(step, sub + 1)
            ef = new.total_energy()
            if ef == float("inf"):
                de = ef
            else:
                ei = 0.0 if model == "hard_sphere" else sysm.total_energy()
                de = -ei if ei == float("inf") else ef - ei
            if not de > self.max_e * len(body):
                self.state, self.is_final, self.boundary = "waiting", trial == self.target, L2
                return new, L2
        return sysm, L
