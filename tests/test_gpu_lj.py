"""GPU parity for Lennard-Jones Metropolis sweeps (PairwiseCutoff<Isotropic<
[Shifted<]LennardJones>>>) through the C ABI, against the oracle.

Bar: per-apply counts, sorted per-trial traces and final FP64 positions bit
for bit, and no ambiguous decision.  The GPU sums the pair energies in
another order than the reference's VecCell order; every decision is taken
only when the rigorous bound on that difference cannot change it, and the
library counts the ones where it could (expected: none).  In mixed
precision the pair energies are FP32 first; decisions the FP32 error bound
leaves open are re-taken in FP64 (counted as band re-tests), so the bar is
the same.
"""
import ctypes as C
import sys
import time

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


def _replay_lj(hmc, L, pos, model, params, kT, d, n_applies, seed=42, trace=True, precision="fp64"):
    ref = O.OracleSystem(L, pos, model=model, params=params, seed=seed)
    b, s = ref.positions()
    gpu = hmc.GpuParallelSweep(L, b, sites=s, model=model, params=params, precision=precision)
    cap = 64 * len(pos) + 1024
    if trace:
        gpu.trace(True, capacity=cap)
        ref.trace(True)
    for k in range(n_applies):
        step, sub = ref.counter()
        ga, gr, used = gpu.apply(step, sub, seed, kT, d)
        oa, orj = ref.parallel_apply(kT, d)
        _, sub2 = ref.counter()
        assert (ga, gr, used) == (oa, orj, sub2 - sub), (k, ga, gr, used, oa, orj, sub2 - sub)
        if trace:
            gt = gpu.read_trace(capacity=cap)
            ot = ref.read_trace()
            gt = gt[gt["status"] != 2]
            ot = ot[ot["status"] != 2]
            key = ["substep", "color", "space", "body", "status"]
            gt = gt[np.lexsort((gt["space"], gt["color"], gt["substep"]))][key]
            ot = ot[np.lexsort((ot["space"], ot["color"], ot["substep"]))][key]
            assert np.array_equal(gt, ot), k
        ref.increment_step()
    gb, gs = gpu.positions()
    ob, os_ = ref.positions()
    assert np.array_equal(gb.view(np.uint64), ob.view(np.uint64))
    assert np.array_equal(gs.view(np.uint64), os_.view(np.uint64))
    st = gpu.stats()
    gpu.close()
    assert st["ambiguous"] == 0, st
    return st, ref


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lj_shifted_fcc_3d(hmc, wl, precision):
    # 864 LJ particles, FCC at number density 0.8, kT 1.2, r_cut = r_shift = 2.5
    L, pos = wl.fcc_number_density((6, 6, 6), 0.8)
    t = time.time()
    st, _ = _replay_lj(hmc, L, pos, "lj_shifted", (1.0, 1.0, 2.5, 2.5), 1.2, 0.1, 10, precision=precision)
    log(f"LJ shifted N=864 {precision}: 10 applies bit-identical ({time.time() - t:.1f}s) {st}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lj_plain_square_2d(hmc, wl, precision):
    # 2D LJ without shift on a 24 x 24 square lattice at density 0.7
    L, pos = wl.square_lattice_disks(n_side=24, packing_fraction=0.7 * np.pi / 4)
    st, _ = _replay_lj(hmc, L, pos, "lj", (1.0, 1.0, 2.5), 1.0, 0.15, 10, seed=3, precision=precision)
    log(f"LJ 2D N=576 {precision}: 10 applies bit-identical {st}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lj_space_width_equal_to_cutoff(hmc, wl, precision):
    # box 10 with r_cut 2.5: four spaces of width exactly r_cut per axis, so the
    # neighbour scan reaches two spaces out (and wraps onto itself), and the
    # FCC second shell sits exactly at the cutoff
    L, pos = wl.fcc_at_volume((4, 4, 4), 1000.0)
    st, _ = _replay_lj(hmc, L, pos, "lj_shifted", (1.0, 1.0, 2.5, 2.5), 1.5, 0.2, 8, seed=9, precision=precision)
    log(f"LJ reach-2 N=256 {precision}: 8 applies bit-identical {st}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lj_config_c4_full_size(hmc, wl, precision):
    # N = 524,288 LJ particles: FCC 64 x 64 x 32 cells at number density 0.8
    # (box 109.44 x 109.44 x 54.72), kT 1.2, shifted at r_cut 2.5, d = 0.1
    L, pos = wl.fcc_number_density((64, 64, 32), 0.8)
    t = time.time()
    st, ref = _replay_lj(hmc, L, pos, "lj_shifted", (1.0, 1.0, 2.5, 2.5), 1.2, 0.1, 8, trace=False,
                         precision=precision)
    log(f"C4 N={len(pos)} {precision}: 8 applies bit-identical ({time.time() - t:.1f}s) {st}")
    if precision == "mixed":
        # the FP32 bound leaves few decisions open
        assert st["band_retests"] < 0.01 * 8 * len(pos), st


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lj_more_candidates_than_the_slot_table(hmc, wl, precision):
    # box 15 with r_cut 2.5: six spaces of width r_cut per axis, so the scan
    # reaches two spaces out (5^3 spaces, about 1,200 candidates per trial):
    # more than the LDS slot table holds, so the segment search runs
    L, pos = wl.fcc_at_volume((8, 8, 8), 15.0**3)
    st, _ = _replay_lj(hmc, L, pos, "lj_shifted", (1.0, 1.0, 2.5, 2.5), 1.2, 0.15, 4, seed=21, precision=precision)
    log(f"LJ reach-2 dense N={len(pos)} {precision}: 4 applies bit-identical {st}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lj_nearly_coincident_pair(hmc, wl, precision):
    # two sites 4e-4 apart: (sigma / r)^12 overflows FP32 (the mixed path
    # must re-take those decisions in FP64) while FP64 stays finite
    L, pos = wl.fcc_number_density((6, 6, 6), 0.8)
    pos = pos.copy()
    pos[1] = pos[0] + np.array([4e-4, 0.0, 0.0])
    st, _ = _replay_lj(hmc, L, pos, "lj_shifted", (1.0, 1.0, 2.5, 2.5), 1.2, 0.1, 6, seed=5, precision=precision)
    log(f"LJ near-coincident pair {precision}: 6 applies bit-identical {st}")
    if precision == "mixed":
        assert st["band_retests"] > 0


def test_lj_fp32_error_stays_within_its_bound(hmc, wl):
    # with the diagnostic on, every trial also forms dH in FP64: the largest
    # FP32 error must lie within the rigorous bound (ratio <= 1), and the
    # decisions stay bit-identical to the oracle's
    L, pos = wl.fcc_number_density((6, 6, 6), 0.8)
    ref = O.OracleSystem(L, pos, model="lj_shifted", params=(1.0, 1.0, 2.5, 2.5), seed=42)
    b, s = ref.positions()
    gpu = hmc.GpuParallelSweep(L, b, sites=s, model="lj_shifted", params=(1.0, 1.0, 2.5, 2.5),
                               precision="mixed", fp32_error=True)
    for _ in range(6):
        step, sub = ref.counter()
        ga, gr, used = gpu.apply(step, sub, 42, 1.2, 0.1)
        oa, orj = ref.parallel_apply(1.2, 0.1)
        assert (ga, gr) == (oa, orj)
        ref.increment_step()
    st = gpu.stats()
    log(f"LJ FP32 dH error: max {st['fp32_error_max']:.3g}, max error / bound {st['fp32_error_bound']:.3g}")
    assert 0.0 < st["fp32_error_max"] and 0.0 < st["fp32_error_bound"] <= 1.0
    gb, _ = gpu.positions()
    assert np.array_equal(gb, ref.positions()[0])
    gpu.close()


def test_total_energy_matches_oracle(hmc, wl):
    # PairwiseCutoff::total_energy on the device: within the summation-order
    # bound of the oracle's VecCell-order sum, before and after moves
    L, pos = wl.fcc_number_density((6, 6, 6), 0.8)
    prm = (1.0, 1.0, 2.5, 2.5)
    ref = O.OracleSystem(L, pos, model="lj_shifted", params=prm, seed=4)
    b, s = ref.positions()
    gpu = hmc.GpuParallelSweep(L, b, sites=s, model="lj_shifted", params=prm)
    for k in range(3):
        e, bound = gpu.total_energy()
        eo = ref.total_energy()
        assert 0.0 < bound < 1e-6 * abs(eo) and abs(e - eo) <= bound, (k, e, eo, bound)
        step, sub = ref.counter()
        gpu.apply(step, sub, 4, 1.2, 0.1)
        ref.parallel_apply(1.2, 0.1)
        ref.increment_step()
    gpu.close()


def test_total_energy_hard_shapes(hmc, wl):
    # 0 for a valid hard-sphere state, +inf once two spheres overlap
    L, pos = wl.fcc_spheres((5, 5, 5), packing_fraction=0.5)
    ref = O.OracleSystem(L, pos, seed=1)
    b, s = ref.positions()
    gpu = hmc.GpuParallelSweep(L, b, sites=s)
    assert gpu.total_energy()[0] == 0.0 == ref.total_energy()
    bad = b.copy()
    bad[7] = bad[3] + np.array([0.3, 0.0, 0.0])
    ref2 = O.OracleSystem(L, bad, seed=1)
    gpu.set_positions(ref2.positions()[0], ref2.positions()[1])
    assert gpu.total_energy()[0] == np.inf == ref2.total_energy()
    gpu.close()


# ------------------------------------------------- persistent apply (mixed) --
def _lj_persistent_replay(hmc, L, pos, model, params, kT, d, n_applies, seed=42, trace=True):
    """Mixed precision (one persistent launch per apply) against the oracle
    and against the per-colour FP64 passes; returns the mixed stats."""
    st, _ = _replay_lj(hmc, L, pos, model, params, kT, d, n_applies, seed=seed, trace=trace, precision="mixed")
    assert st["tiled"] == 2, st
    return st


def test_lj_persistent_shifted_3d(hmc, wl):
    # 6,912 shifted-LJ bodies, FCC 12^3 at number density 0.8 (8 spaces of
    # width 2.56 per axis), kT 1.2: one persistent launch per apply
    L, pos = wl.fcc_number_density((12, 12, 12), 0.8)
    t = time.time()
    st = _lj_persistent_replay(hmc, L, pos, "lj_shifted", (1.0, 1.0, 2.5, 2.5), 1.2, 0.1, 12)
    log(f"LJ persistent 3D N={len(pos)}: 12 applies bit-identical ({time.time() - t:.1f}s) {st}")
    assert st["tile_overflows"] == 0


def test_lj_persistent_plain_2d(hmc, wl):
    # 4,096 plain LJ disks (square lattice, r_cut 2.5): the 2D persistent apply
    L, pos = wl.square_lattice_disks(n_side=64, packing_fraction=0.55)
    st = _lj_persistent_replay(hmc, L, pos, "lj", (1.0, 1.0, 2.5), 1.0, 0.15, 10)
    log(f"LJ persistent 2D N={len(pos)}: 10 applies bit-identical {st}")


def test_lj_persistent_near_coincident_pair(hmc, wl):
    # two sites 4e-4 apart: FP32 cannot decide their trials (re-taken in FP64)
    L, pos = wl.fcc_number_density((12, 12, 12), 0.8)
    pos = pos.copy()
    pos[1] = pos[0] + np.array([4e-4, 0.0, 0.0])
    st = _lj_persistent_replay(hmc, L, pos, "lj_shifted", (1.0, 1.0, 2.5, 2.5), 1.2, 0.1, 6, seed=5)
    log(f"LJ persistent near-coincident pair: 6 applies bit-identical {st}")
    assert st["band_retests"] > 0


def test_lj_persistent_overflowing_region(hmc, wl):
    # a dense FCC block in one corner of a mostly empty box: the regions
    # around it hold more sites than their LDS capacity and run their trials
    # from the cell lists in global memory (L = 31: 12 spaces of width 2.58
    # per axis, so the neighbour scan reaches one space out)
    L1, block = wl.fcc_number_density((6, 6, 6), 0.8)
    L = np.array([31.0, 31.0, 31.0])
    pos = block - (-L1 / 2) + (-L / 2 + 0.6)
    st = _lj_persistent_replay(hmc, L, pos, "lj_shifted", (1.0, 1.0, 2.5, 2.5), 1.2, 0.1, 6, seed=8)
    log(f"LJ persistent overflow: 6 applies bit-identical {st}")
    assert st["tile_overflows"] > 0


def test_lj_persistent_equals_per_colour(hmc, wl):
    # the same run with persistent=False (one launch per colour pass)
    L, pos = wl.fcc_number_density((12, 12, 12), 0.8)
    prm = (1.0, 1.0, 2.5, 2.5)
    a = hmc.GpuParallelSweep(L, wl.wrap(L, pos), model="lj_shifted", params=prm, precision="mixed")
    b = hmc.GpuParallelSweep(L, wl.wrap(L, pos), model="lj_shifted", params=prm, precision="mixed",
                             persistent=False)
    assert a.stats()["tiled"] == 2 and b.stats()["tiled"] != 2
    assert a.run(0, 11, 1.2, 0.1, 8) == b.run(0, 11, 1.2, 0.1, 8)
    assert np.array_equal(a.positions()[1].view(np.uint64), b.positions()[1].view(np.uint64))
    a.close()
    b.close()
