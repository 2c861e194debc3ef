"""The Trial boundary's side effects and several-move steps, through the C ABI.

* Lazy sync (hmc_moved_since + hmc_get_bodies): after every apply the bodies
  the device reports as moved are exactly the bodies the oracle accepted in
  that apply (ParallelSweep::update_bodies -> update_body_properties,
  parallel_sweep.rs:283-314), their gathered positions equal the oracle's,
  and a host copy updated on those bodies only stays equal to the oracle's
  whole microstate.
* hmc_run with several moves per step (translate then rotate, or two
  translates) equals the oracle applying the same trials in order with the
  step's substeps following on, then increment_step: per-move counts and
  final state bit for bit, with the later steps replayed from a graph.
"""
import math
import sys

import numpy as np
import pytest

import _oracle as O
import _pkg

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hmc():
    p = _pkg.pkg()
    p.lib()
    return p


@pytest.fixture(scope="module")
def wl():
    return _pkg.workloads()


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _sc(n, phi):
    L = (n**3 / phi) ** (1.0 / 3.0)
    a = L / n
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    pos = (np.stack([i.ravel(), j.ravel(), k.ravel()], 1) + 0.5) * a - L / 2
    return np.array([L, L, L]), pos


@pytest.mark.parametrize("case", ["spheres_mixed", "spheres_fp64", "disks_fp64", "lj_mixed"])
def test_moved_bodies_sync_the_host_microstate(hmc, wl, case):
    log(f"moved-body sync: {case}")
    kT, d, model, params = 1.0, 0.08, "hard_sphere", (1.0,)
    if case == "spheres_mixed":
        L, pos = wl.fcc_spheres(n_cells=(10, 10, 10), packing_fraction=0.45)
        precision = "mixed"
    elif case == "spheres_fp64":
        L, pos = wl.fcc_spheres(n_cells=(10, 10, 10), packing_fraction=0.45)
        precision = "fp64"
    elif case == "disks_fp64":
        L, pos = wl.square_lattice_disks(n_side=40, packing_fraction=0.6)
        precision, d = "fp64", 0.15
    else:
        L, pos = wl.fcc_number_density((6, 6, 6), 0.8)
        model, params, precision, kT, d = "lj_shifted", (1.0, 1.0, 2.5, 2.5), "mixed", 1.2, 0.1
    ref = O.OracleSystem(L, pos, model=model, params=params, seed=42)
    hb, hs = ref.positions()  # the host microstate, synced lazily
    gpu = hmc.GpuParallelSweep(L, hb, sites=hs, model=model, params=params, precision=precision)
    ref.trace(True)
    hb, hs = hb.copy(), hs.copy()
    since = 0
    for k in range(8):
        step, sub = ref.counter()
        g = gpu.apply(step, sub, 42, kT, d)
        oa, orj = ref.parallel_apply(kT, d)
        ref.increment_step()
        assert g[:2] == (oa, orj), (k, g, oa, orj)
        t = ref.read_trace()
        want = np.unique(t["body"][t["status"] == 1]).astype(np.uint32)
        idx, now = gpu.moved_since(since)
        assert now == since + 1
        assert np.array_equal(idx, want), k
        b, s = gpu.get_bodies(idx)
        hb[idx], hs[idx] = b, s
        ob, os_ = ref.positions()
        assert np.array_equal(_bits(b), _bits(ob[idx])) and np.array_equal(_bits(s), _bits(os_[idx])), k
        assert np.array_equal(_bits(hb), _bits(ob)) and np.array_equal(_bits(hs), _bits(os_)), k
        since = now
    # nothing moved since the last sync; everything moved since creation is a superset
    idx, now = gpu.moved_since(since)
    assert len(idx) == 0 and now == since
    all_idx, _ = gpu.moved_since(0)
    assert len(all_idx) >= len(want)
    gpu.close()


def test_moved_bodies_and_orientations_of_cubes(hmc):
    L, pos = _sc(6, 0.55)
    q = np.tile([1.0, 0.0, 0.0, 0.0], (len(pos), 1))
    ref = O.ConvexOracle(L, pos, q, O.CUBE, seed=5)
    bp, bq, sp, sq = ref.state()
    gpu = hmc.GpuParallelSweep(L, bp, sites=sp, model="convex", vertices=O.CUBE, orientations=bq,
                               site_orientations=sq)
    ref.trace(True)
    since = 0
    for k in range(6):
        for rotate in (False, True):
            step, sub = ref.counter()
            g = gpu.apply_rotate(step, sub, 5, 1.0, 0.1) if rotate else gpu.apply(step, sub, 5, 1.0, 0.05)
            oa, orj = ref.apply(rotate, 0.1 if rotate else 0.05)
            assert g[:2] == (oa, orj)
            t = ref.read_trace()
            want = np.unique(t["body"][t["status"] == 1]).astype(np.uint32)
            idx, now = gpu.moved_since(since)
            assert np.array_equal(idx, want), (k, rotate)
            b, s, gbq, gsq = gpu.get_bodies(idx, orientations=True)
            ob, oq, os_, osq = ref.state()
            for x, y in ((b, ob[idx]), (s, os_[idx]), (gbq, oq[idx]), (gsq, osq[idx])):
                assert np.array_equal(_bits(x), _bits(y)), (k, rotate)
            since = now
        ref.increment_step()
    gpu.close()


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "direct"])
def test_run_translate_then_rotate_matches_oracle(hmc, graph):
    # 216 cubes at packing fraction 0.55: 8 steps of {translate 0.05; rotate 0.1}
    L, pos = _sc(6, 0.55)
    q = np.tile([1.0, 0.0, 0.0, 0.0], (len(pos), 1))
    ref = O.ConvexOracle(L, pos, q, O.CUBE, seed=42)
    bp, bq, sp, sq = ref.state()
    gpu = hmc.GpuParallelSweep(L, bp, sites=sp, model="convex", vertices=O.CUBE, orientations=bq,
                               site_orientations=sq, graphs=graph)
    counts = gpu.run_moves(0, 42, 1.0, [("translate", 0.05), ("rotate", 0.1)], 8)
    tot = [[0, 0], [0, 0]]
    for _ in range(8):
        for m, rotate in enumerate((False, True)):
            a, r = ref.apply(rotate, 0.1 if rotate else 0.05)
            tot[m][0] += a
            tot[m][1] += r
        ref.increment_step()
    assert counts == [tuple(x) for x in tot], (counts, tot)
    gb, gs = gpu.positions()
    gqb, gqs = gpu.orientations()
    ob, oq, os_, osq = ref.state()
    for x, y in ((gb, ob), (gs, os_), (gqb, oq), (gqs, osq)):
        assert np.array_equal(_bits(x), _bits(y))
    gpu.close()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_run_two_translates_per_step_matches_oracle(hmc, wl, precision):
    L, pos = wl.fcc_spheres(n_cells=(8, 8, 8), packing_fraction=0.45)
    ref = O.OracleSystem(L, pos, seed=3)
    b, s = ref.positions()
    gpu = hmc.GpuParallelSweep(L, b, sites=s, precision=precision)
    counts = gpu.run_moves(0, 3, 1.0, [("translate", 0.08), ("translate", 0.03)], 6)
    tot = [[0, 0], [0, 0]]
    for _ in range(6):
        for m, d in enumerate((0.08, 0.03)):
            a, r = ref.parallel_apply(1.0, d)
            tot[m][0] += a
            tot[m][1] += r
        ref.increment_step()
    assert counts == [tuple(x) for x in tot], (counts, tot)
    gb, gs = gpu.positions()
    ob, os_ = ref.positions()
    assert np.array_equal(_bits(gb), _bits(ob)) and np.array_equal(_bits(gs), _bits(os_))
    # one move per step is hmc_run_translate
    a = hmc.GpuParallelSweep(L, b, sites=s, precision=precision)
    c = hmc.GpuParallelSweep(L, b, sites=s, precision=precision)
    assert a.run_moves(0, 9, 1.0, [("translate", 0.08)], 5) == [c.run(0, 9, 1.0, 0.08, 5)]
    assert np.array_equal(_bits(a.positions()[1]), _bits(c.positions()[1]))
    for x in (gpu, a, c):
        x.close()


def test_run_rejects_bad_move_lists(hmc, wl):
    L, pos = wl.fcc_spheres(n_cells=(4, 4, 4), packing_fraction=0.45)
    gpu = hmc.GpuParallelSweep(L, wl.wrap(L, pos))
    for moves in ([], [("translate", 0.1)] * 5, [("translate", -1.0)], [("rotate", 0.1)]):
        with pytest.raises(hmc.HmcError):
            gpu.run_moves(0, 1, 1.0, moves, 2)
    with pytest.raises(hmc.HmcError):
        gpu.get_bodies(np.array([len(pos)], np.uint32))
    gpu.close()
