"""GPU parity for hard convex polyhedra (HardShape<Convex<ConvexPolyhedron>>)
with Translate and Rotate<Versor> moves, through the C ABI, against the
oracle: the device XenoCollide and pair test on random and near-contact
pairs, the rotation sampler, and bit-identical replays (counts, sorted
traces, FP64 positions and orientations), in FP64 and in the mixed mode
(FP32 separating axes, FP64 XenoCollide inside their margin), whose every
FP32 class must equal the FP64 decision.
"""
import ctypes as C
import math
import sys
import time

import numpy as np
import pytest

import _oracle as O
import _pkg

pytestmark = pytest.mark.gpu

OCTA = np.array([[-0.5, -0.5, 0.0], [0.5, -0.5, 0.0], [0.5, 0.5, 0.0], [-0.5, 0.5, 0.0],
                 [0.0, 0.0, 1.0 / math.sqrt(2.0)], [0.0, 0.0, -1.0 / math.sqrt(2.0)]])
TETRA = np.array([[1.0, 1.0, 1.0], [1.0, -1.0, -1.0], [-1.0, 1.0, -1.0], [-1.0, -1.0, 1.0]]) * 0.5


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hmc():
    p = _pkg.pkg()
    p.lib()
    return p


def _random_quats(rng, n):
    q = rng.normal(size=(n, 4))
    return q / np.linalg.norm(q, axis=1)[:, None]


@pytest.mark.parametrize("verts", [O.CUBE, OCTA], ids=["cube", "octahedron"])
def test_device_collide_matches_oracle(hmc, verts):
    rng = np.random.default_rng(5)
    n = 4000
    q = _random_quats(rng, n)
    # separations spread across contact: many pairs within rounding of touching
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    v = dirs * rng.uniform(0.3, 1.8, size=(n, 1))
    rows = np.hstack([v, q])
    got = hmc.debug_convex("collide", verts, rows)
    want = np.array([O.convex_collide(verts, r[:3], r[3:]) for r in rows], float)
    assert np.array_equal(got, want)
    assert 0.2 < want.mean() < 0.8


def test_device_global_pair_test_matches_oracle(hmc):
    rng = np.random.default_rng(9)
    n = 4000
    ri = rng.uniform(-20, 20, size=(n, 3))
    rj = ri + rng.normal(size=(n, 3)) * 0.9
    qi, qj = _random_quats(rng, n), _random_quats(rng, n)
    rows = np.hstack([ri, qi, rj, qj])
    got = hmc.debug_convex("global", O.CUBE, rows)
    want = np.array([O.convex_intersects_global(O.CUBE, *np.split(r, [3, 7, 10])) for r in rows], float)
    assert np.array_equal(got, want)


def _contact_distance(verts, qi, qj, u, r2):
    """Distance t along unit u at which the pair (0, qi), (t u, qj) stops
    overlapping in FP64, by bisection on the oracle's pair test."""
    lo, hi = 0.0, r2
    for _ in range(45):
        mid = 0.5 * (lo + hi)
        if O.convex_intersects_global(verts, np.zeros(3), qi, mid * u, qj):
            lo = mid
        else:
            hi = mid
    return hi


@pytest.mark.parametrize("verts", [O.CUBE, OCTA, TETRA], ids=["cube", "octahedron", "tetrahedron"])
def test_fp32_separating_axes_agree_with_fp64(hmc, verts):
    # pairs within the circumsphere cutoff, at random distances and at
    # relative offsets of 1e-3 .. 1e-9 from contact: every pair the FP32
    # tests decide (class 0 apart, 1 overlap) has that FP64 decision
    rng = np.random.default_rng(17)
    R = float(np.max(np.linalg.norm(verts, axis=1)))
    n = 3000
    qi, qj = _random_quats(rng, n), _random_quats(rng, n)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    t = rng.uniform(0.0, 2.0 * R, size=n)
    nc = 600
    rel = rng.choice([1e-3, -1e-3, 1e-4, -1e-4, 1e-6, -1e-6, 1e-9, -1e-9], size=nc)
    for k in range(nc):
        t[k] = _contact_distance(verts, qi[k], qj[k], u[k], 2.0 * R) * (1.0 + rel[k])
    ri = rng.uniform(-5, 5, size=(n, 3))
    rows = np.hstack([ri, qi, ri + t[:, None] * u, qj])
    cls = hmc.debug_convex("sat", verts, rows)
    want = np.array([O.convex_intersects_global(verts, *np.split(r, [3, 7, 10])) for r in rows])
    assert np.array_equal(hmc.debug_convex("global", verts, rows), want.astype(float))
    assert set(np.unique(cls)) <= {0.0, 1.0, 2.0}
    assert not np.any((cls == 0) & want), "FP32 'apart' on an overlapping pair"
    assert not np.any((cls == 1) & ~want), "FP32 'overlap' on a separated pair"
    far = cls[nc:]
    decided = np.mean(far != 2)
    near = cls[:nc][np.abs(rel) <= 1e-6]
    log(f"{len(verts)} vertices: FP32 decides {decided:.3f} of random pairs, {np.mean(near != 2):.3f} of pairs "
        f"within 1e-6 of contact; overlap share {want[nc:].mean():.2f}")
    assert decided > 0.9
    assert np.mean(near == 2) > 0.9  # near contact is left to FP64


def test_device_rotation_sampler_matches_oracle(hmc):
    rows = []
    want = []
    for index in range(300):
        for sigma in (0.1, 0.8, 2.5):
            rows.append([7, 3, 42, index, sigma])
            want.append(O.versor_displacement(7, 3, 42, index, sigma, 1)[0])
    got = hmc.debug_convex("rotate", O.CUBE, np.array(rows, float))
    assert np.array_equal(got.view(np.uint64), np.array(want).view(np.uint64))


def _sc(n, phi, vol=1.0):
    L = (n**3 * vol / phi) ** (1.0 / 3.0)
    a = L / n
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    pos = (np.stack([i.ravel(), j.ravel(), k.ravel()], 1) + 0.5) * a - L / 2
    return np.array([L, L, L]), pos


def _replay(hmc, L, pos, quats, verts, d, a, steps, seed=42, trace=True, precision="fp64"):
    ref = O.ConvexOracle(L, pos, quats, verts, seed=seed)
    bp, bq, sp, sq = ref.state()
    gpu = hmc.GpuParallelSweep(L, bp, sites=sp, model="convex", vertices=verts, orientations=bq,
                               site_orientations=sq, precision=precision)
    cap = 8 * len(pos) + 1024
    if trace:
        gpu.trace(True, capacity=cap)
        ref.trace(True)
    key = ["substep", "color", "space", "body", "status"]

    def sorted_trace(t):
        t = t[t["status"] != 2]
        return t[np.lexsort((t["space"], t["color"], t["substep"]))][key]

    for k in range(steps):
        for rotate in (False, True):
            step, sub = ref.counter()
            if rotate:
                g = gpu.apply_rotate(step, sub, seed, 1.0, a)
            else:
                g = gpu.apply(step, sub, seed, 1.0, d)
            oa, orj = ref.apply(rotate, a if rotate else d)
            _, sub2 = ref.counter()
            assert g == (oa, orj, sub2 - sub), (k, rotate, g, oa, orj, sub2 - sub)
            if trace:
                assert np.array_equal(sorted_trace(gpu.read_trace(capacity=cap)), sorted_trace(ref.read_trace())), \
                    (k, rotate)
        ref.increment_step()
    gb, gs = gpu.positions()
    gqb, gqs = gpu.orientations()
    ob, oq, os_, osq = ref.state()
    for x, y in ((gb, ob), (gs, os_), (gqb, oq), (gqs, osq)):
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64))
    ref.band = gpu.stats()["band_retests"]
    gpu.close()
    return ref


def test_replay_cubes(hmc):
    # 216 cubes, simple cubic at packing fraction 0.55, random orientations
    # from rotation sweeps of an aligned start
    L, pos = _sc(6, 0.55)
    q = np.tile([1.0, 0.0, 0.0, 0.0], (len(pos), 1))
    t = time.time()
    ref = _replay(hmc, L, pos, q, O.CUBE, 0.05, 0.1, 8)
    assert not ref.any_overlap()
    log(f"cubes N=216: 8 translate+rotate steps bit-identical ({time.time() - t:.1f}s)")


@pytest.mark.parametrize("verts", [O.CUBE, TETRA], ids=["cubes", "tetrahedra"])
def test_replay_mixed(hmc, verts):
    # the mixed mode replays the oracle bit for bit: 216 shapes, simple cubic
    # (cubes at packing fraction 0.55, tetrahedra at 0.3), aligned start
    # randomised by the rotation moves
    vol = 1.0 if verts is O.CUBE else 1.0 / 3.0
    L, pos = _sc(6, 0.55 if verts is O.CUBE else 0.3, vol=vol)
    q = np.tile([1.0, 0.0, 0.0, 0.0], (len(pos), 1))
    assert not O.ConvexOracle(L, pos, q, verts).any_overlap()
    t = time.time()
    ref = _replay(hmc, L, pos, q, verts, 0.05, 0.3, 8, precision="mixed")
    log(f"{len(verts)}-vertex shapes N=216 mixed: 8 steps bit-identical, {ref.band} pairs to FP64 "
        f"({time.time() - t:.1f}s)")


def test_replay_octahedra_large_rotations(hmc):
    # 125 octahedra (the reference's fixture shape), simple cubic at packing
    # fraction 0.15 (spacing 1.46 > the 1.41 tip-to-tip height: a valid
    # aligned start), large rotations
    L, pos = _sc(5, 0.15, vol=math.sqrt(2.0) / 3.0)
    q = np.tile([1.0, 0.0, 0.0, 0.0], (len(pos), 1))
    assert not O.ConvexOracle(L, pos, q, OCTA).any_overlap()
    _replay(hmc, L, pos, q, OCTA, 0.1, 0.8, 6, seed=3)
    _replay(hmc, L, pos, q, OCTA, 0.1, 0.8, 6, seed=3, precision="mixed")


def test_replay_config_c3_full(hmc):
    # N = 262,144 unit cubes, simple cubic 64^3 at packing fraction 0.55
    # (box 78.11), aligned start, translate 0.05 + rotate 0.1 per step, 10 steps
    L, pos = _sc(64, 0.55)
    q = np.tile([1.0, 0.0, 0.0, 0.0], (len(pos), 1))
    t = time.time()
    _replay(hmc, L, pos, q, O.CUBE, 0.05, 0.1, 10, trace=False)
    log(f"C3 N={len(pos)}: 10 translate+rotate steps bit-identical ({time.time() - t:.1f}s)")


def test_config_c3_mixed_equals_fp64(hmc):
    # C3 (N = 262,144 unit cubes, packing fraction 0.55): 10 steps of
    # translate 0.05 + rotate 0.1 in the mixed mode and in FP64 (which the
    # oracle pins above) give the same FP64 positions and orientations
    L, pos = _sc(64, 0.55)
    q = np.tile([1.0, 0.0, 0.0, 0.0], (len(pos), 1))
    out = []
    for prec in ("fp64", "mixed"):
        gpu = hmc.GpuParallelSweep(L, pos, model="convex", vertices=O.CUBE, orientations=q, precision=prec)
        counts = gpu.run_moves(0, 42, 1.0, [("translate", 0.05), ("rotate", 0.1)], 10)
        out.append((counts, *gpu.positions(), *gpu.orientations(), gpu.stats()["band_retests"]))
        gpu.close()
    assert out[0][0] == out[1][0]
    for x, y in zip(out[0][1:5], out[1][1:5]):
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64))
    log(f"C3 mixed == FP64 over 10 steps; {out[1][5]} pairs left to FP64 XenoCollide")
