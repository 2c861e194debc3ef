"""GPU parity for hard convex polyhedra (HardShape<Convex<ConvexPolyhedron>>)
with Translate and Rotate<Versor> moves, through the C ABI, against the
oracle: the device XenoCollide and pair test on random and near-contact
pairs, the rotation sampler, and bit-identical replays (counts, sorted
traces, FP64 positions and orientations).
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


def _replay(hmc, L, pos, quats, verts, d, a, steps, seed=42, trace=True):
    ref = O.ConvexOracle(L, pos, quats, verts, seed=seed)
    bp, bq, sp, sq = ref.state()
    gpu = hmc.GpuParallelSweep(L, bp, sites=sp, model="convex", vertices=verts, orientations=bq,
                               site_orientations=sq)
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


def test_replay_octahedra_large_rotations(hmc):
    # 125 octahedra (the reference's fixture shape), simple cubic at packing
    # fraction 0.15 (spacing 1.46 > the 1.41 tip-to-tip height: a valid
    # aligned start), large rotations
    L, pos = _sc(5, 0.15, vol=math.sqrt(2.0) / 3.0)
    q = np.tile([1.0, 0.0, 0.0, 0.0], (len(pos), 1))
    assert not O.ConvexOracle(L, pos, q, OCTA).any_overlap()
    _replay(hmc, L, pos, q, OCTA, 0.1, 0.8, 6, seed=3)


def test_replay_config_c3_full(hmc):
    # N = 262,144 unit cubes, simple cubic 64^3 at packing fraction 0.55
    # (box 78.11), aligned start, translate 0.05 + rotate 0.1 per step, 10 steps
    L, pos = _sc(64, 0.55)
    q = np.tile([1.0, 0.0, 0.0, 0.0], (len(pos), 1))
    t = time.time()
    _replay(hmc, L, pos, q, O.CUBE, 0.05, 0.1, 10, trace=False)
    log(f"C3 N={len(pos)}: 10 translate+rotate steps bit-identical ({time.time() - t:.1f}s)")
