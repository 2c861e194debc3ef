"""GPU parity: the HIP path (through the C ABI) against the oracle.

Bit-exact bar: per-apply counts, per-trial (substep, colour, space, body,
status) traces, and final FP64 body and site positions.  The oracle calls the
host C library, as the reference does; the device computes the same values
(its restatements of glibc's log, exp, cbrt and sincos).
"""
import ctypes as C
import json
import math
import pathlib
import sys
import time

import numpy as np
import pytest

import _oracle as O
import _pkg

pytestmark = pytest.mark.gpu
GOLDEN = pathlib.Path(__file__).with_name("golden")


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


# ------------------------------------------------------------ device math --
def test_device_cbrt_matches_host_library(hmc):
    # the 3D ball sampler's cbrt(u), u = k * 2^-53, against the C library that
    # Rust's f64::cbrt calls on Linux
    libm = C.CDLL("libm.so.6")
    libm.cbrt.argtypes = [C.c_double]
    libm.cbrt.restype = C.c_double
    rng = np.random.default_rng(3)
    u = rng.integers(0, 2**53, size=200000, dtype=np.uint64).astype(np.float64) * 2.0**-53
    u = np.concatenate([u, [0.0, 2.0**-53, 0.125, 0.5, 1.0 - 2.0**-53, 1e-300, 8.0, 27.0]])
    got = hmc.debug_math("cbrt", u)
    want = np.array([libm.cbrt(float(v)) for v in u])
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


def test_device_sqrt_and_division_are_ieee(hmc):
    rng = np.random.default_rng(5)
    a = rng.uniform(0.0, 10.0, size=1 << 18) * 10.0 ** rng.integers(-30, 30, size=1 << 18)
    assert np.array_equal(hmc.debug_math("sqrt", a), np.sqrt(a))
    b = rng.uniform(0.5, 200.0, size=(1 << 18, 2))
    b[:, 0] *= rng.choice([-1.0, 1.0], size=1 << 18)
    assert np.array_equal(hmc.debug_math("div", b.ravel()), b[:, 0] / b[:, 1])


def test_device_normals_match_oracle(hmc):
    # streams chosen to include the ziggurat tail (|z| > R) and wedge tests
    R = 3.6541528853610088
    tails = 0
    for index in range(24):
        want = O.sample(7, 1, 42, index, "normal", 4096)
        got = hmc.debug_normals(7, 1, 42, index, 4096)
        tails += int(np.sum(np.abs(want) > R))
        assert np.array_equal(got, want), index
    assert tails > 0


# ------------------------------------------------------------------ replay --
def _sorted_trace(t):
    t = t[t["status"] != 2]
    order = np.lexsort((t["space"], t["color"], t["substep"]))
    return t[order][["substep", "color", "space", "body", "status"]]


# pass layouts: one persistent launch per apply (mixed hard shapes; FP64
# falls back to the per-colour tiles), one LDS-tiled launch per colour, or
# every trial from global memory
MODES = {"persist": dict(), "percolour": dict(persistent=False), "global": dict(lds_tiles=False)}
# layouts not in the general parametrisations
EXTRA_MODES = {"fp32err": dict(fp32_error=True)}


def expected_tiled(mode, precision):
    return {"persist": 2 if precision == "mixed" else 1, "percolour": 1, "global": 0}[mode]


def _mirror(hmc, ref, L, precision, mode="persist"):
    """GPU copy of the oracle microstate (its wrapped bodies and sites)."""
    b, s = ref.positions()
    return hmc.GpuParallelSweep(L, b, sites=s, precision=precision, **{**MODES, **EXTRA_MODES}[mode])


def _replay(hmc, L, pos, d, n_applies, precision, seed=42, trace=True, step0=0, mode="persist"):
    ref = O.OracleSystem(L, pos, seed=seed, step=step0)
    gpu = _mirror(hmc, ref, L, precision, mode)
    if trace:
        gpu.trace(True, capacity=4 * len(pos) * 4 + 1024)
        ref.trace(True)
    for k in range(n_applies):
        step, sub = ref.counter()
        ga, gr, used = gpu.apply(step, sub, seed, 1.0, d)
        oa, orj = ref.parallel_apply(1.0, d)
        _, sub2 = ref.counter()
        assert (ga, gr, used) == (oa, orj, sub2 - sub), (k, ga, gr, used, oa, orj, sub2 - sub)
        if trace:
            gt = _sorted_trace(gpu.read_trace(capacity=4 * len(pos) * 4 + 1024))
            ot = _sorted_trace(ref.read_trace())
            assert len(gt) == len(ot) == oa + orj
            assert np.array_equal(gt, ot), k
        ref.increment_step()
    gb, gs = gpu.positions()
    ob, os_ = ref.positions()
    assert np.array_equal(gb.view(np.uint64), ob.view(np.uint64))
    assert np.array_equal(gs.view(np.uint64), os_.view(np.uint64))
    st = gpu.stats()
    gpu.close()
    return st


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_replay_hard_disks(hmc, wl, precision, mode):
    L, pos = wl.square_lattice_disks(n_side=64, packing_fraction=0.70)
    t = time.time()
    st = _replay(hmc, L, pos, 0.15, 20, precision, mode=mode)
    assert st["tiled"] == expected_tiled(mode, precision) and st["tile_overflows"] == 0
    log(f"hard disks N=4096 {precision}: 20 applies bit-identical ({time.time() - t:.1f}s) {st}")


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_replay_hard_spheres(hmc, wl, precision, mode):
    L, pos = wl.fcc_spheres((10, 10, 10), packing_fraction=0.50)
    t = time.time()
    st = _replay(hmc, L, pos, 0.08, 12, precision, mode=mode)
    assert st["tiled"] == expected_tiled(mode, precision) and st["tile_overflows"] == 0
    log(f"hard spheres N=4000 {precision}: 12 applies bit-identical ({time.time() - t:.1f}s) {st}")


def test_fp32_error_within_the_contact_band(hmc, wl):
    # with the FP32 error report on (per-colour tile passes), every pair within
    # 1.01 sigma is also measured in FP64: the largest |d2_32 - d2_64| lies
    # within the band's error bound, and the replay stays bit-identical
    L, pos = wl.fcc_spheres((10, 10, 10), packing_fraction=0.50)
    st = _replay(hmc, L, pos, 0.08, 8, "mixed", mode="fp32err")
    log(f"hard spheres FP32 d2 error: max {st['fp32_error_max']:.3g}, max error / bound {st['fp32_error_bound']:.3g}")
    assert st["tiled"] == 1
    assert 0.0 < st["fp32_error_max"] and 0.0 < st["fp32_error_bound"] <= 1.0


@pytest.mark.parametrize("precision,mode", [("fp64", "percolour"), ("mixed", "percolour"), ("mixed", "persist")])
def test_replay_clustered_tile_overflow(hmc, wl, precision, mode):
    # a dense FCC block in one corner of a mostly empty box: the LDS capacity
    # is sized from the mean occupancy, so the tiles (regions) around the
    # block overflow and run their trials from global memory, next to tiles
    # that run from LDS
    L1, block = wl.fcc_spheres((6, 6, 6), packing_fraction=0.50)
    L = np.array([24.5, 24.5, 24.5])  # 24 spaces of width 1.02 per axis
    pos = block - (-L1 / 2) + (-L / 2 + 0.75)
    st = _replay(hmc, L, pos, 0.08, 8, precision, mode=mode)
    assert st["tiled"] == expected_tiled(mode, precision) and st["tile_overflows"] > 0, st
    log(f"clustered N={len(pos)} {precision}: 8 applies bit-identical {st}")


def test_replay_large_moves_dilute(hmc, wl):
    # many trials leave their space (reject path) and cross the periodic faces
    L = np.array([9.3, 8.4, 7.7])
    pos = wl.random_dilute(120, L, 3, 1.0, seed=11)
    _replay(hmc, L, pos, 0.9, 10, "mixed", seed=7)


def test_replay_two_spaces_per_axis(hmc, wl):
    # shape 2 on the first axis: neighbour spaces -1 and +1 coincide
    L = np.array([2.6, 6.3, 5.9])
    pos = wl.random_dilute(18, L, 3, 1.0, seed=3)
    _replay(hmc, L, pos, 0.4, 10, "fp64", seed=5)


def test_replay_config_c1_full(hmc, wl):
    # hard disks N=16384 on a 128x128 lattice at packing fraction 0.70, d=0.15
    L, pos = wl.square_lattice_disks()
    t = time.time()
    _replay(hmc, L, pos, 0.15, 40, "mixed", trace=False)
    log(f"config C1 N=16384: 40 applies bit-identical ({time.time() - t:.1f}s)")


def test_single_body_counts_one_trial(hmc):
    gpu = hmc.GpuParallelSweep([10.0, 10.0], np.zeros((1, 2)))
    a, r, used = gpu.apply(0, 0, 1, 1.0, 0.1)
    assert (a + r, used) == (1, 1)


def test_body_outside_checkerboard_is_an_error(hmc):
    pos = np.array([[0.0, 0.0, 0.0], [40.0, 0.0, 0.0]])
    gpu = hmc.GpuParallelSweep([10.0, 10.0, 10.0], pos)
    with pytest.raises(hmc.HmcError) as e:
        gpu.apply(0, 0, 1, 1.0, 0.1)
    assert e.value.code == hmc.HMC_ERR_COVER


@pytest.mark.parametrize("graphs", [True, False], ids=["graph", "direct"])
def test_run_equals_repeated_apply(hmc, wl, graphs):
    L, pos = wl.fcc_spheres((8, 8, 8), packing_fraction=0.5)
    a = hmc.GpuParallelSweep(L, pos, precision="mixed")
    b = hmc.GpuParallelSweep(L, pos, precision="mixed", graphs=graphs)
    tot = [0, 0]
    for step in range(6):
        x, y, _ = a.apply(100 + step, 0, 9, 1.0, 0.08)
        tot[0] += x
        tot[1] += y
    assert list(b.run(100, 9, 1.0, 0.08, 6)) == tot
    assert np.array_equal(a.positions()[0], b.positions()[0])
    assert np.array_equal(a.positions()[1], b.positions()[1])


def test_run_graph_replays_across_parameter_changes(hmc, wl):
    # runs in pieces with the move size, kT and model changing between them
    # (each change re-captures the step graph), host reads in between, against
    # directly launched applies; LJ also launches a spare cycle per apply
    for model, prec, params, pieces in (("hard_sphere", "mixed", None,
                                         [(0.08, 1.0, 5), (0.08, 1.0, 3), (0.12, 1.0, 4)]),
                                        ("lj_shifted", "fp64", (1.0, 1.0, 2.5, 2.5),
                                         [(0.1, 1.2, 4), (0.1, 0.9, 4), (0.15, 0.9, 3)])):
        if model == "hard_sphere":
            L, pos = wl.fcc_spheres((7, 7, 7), packing_fraction=0.5)
        else:
            L, pos = wl.fcc_number_density((6, 6, 6), 0.8)
        g = hmc.GpuParallelSweep(L, pos, precision=prec, model=model, params=params)
        d = hmc.GpuParallelSweep(L, pos, precision=prec, model=model, params=params, graphs=False)
        step = 10
        for dist, kT, n in pieces:
            assert g.run(step, 5, kT, dist, n) == d.run(step, 5, kT, dist, n), (model, step)
            assert np.array_equal(g.positions()[1], d.positions()[1]), (model, step)
            step += n
        g.close()
        d.close()


# ---------------------------------------------------------- full size (C2) --
def test_config_c2_full_size(hmc, wl):
    # N = 1,000,188 hard spheres, FCC 4*63^3 at packing fraction 0.50, d = 0.08.
    L, pos = wl.fcc_spheres()
    t = time.time()
    ref = O.OracleSystem(L, pos, seed=42)
    fp64 = _mirror(hmc, ref, L, "fp64")
    mixed = _mirror(hmc, ref, L, "mixed")
    log(f"C2 setup {time.time() - t:.1f}s")
    # 40 applies (10 sweeps) against the oracle, positions compared bit for
    # bit every 10 applies
    for step in range(40):
        c64 = fp64.apply(step, 0, 42, 1.0, 0.08)
        cmx = mixed.apply(step, 0, 42, 1.0, 0.08)
        oa, orj = ref.parallel_apply(1.0, 0.08)
        _, used = ref.counter()
        ref.increment_step()
        assert c64 == cmx == (oa, orj, used), (step, c64, cmx, oa, orj, used)
        if step % 10 == 9:
            gb, gs = mixed.positions()
            ob, os_ = ref.positions()
            assert np.array_equal(gb.view(np.uint64), ob.view(np.uint64)), step
            assert np.array_equal(gs.view(np.uint64), os_.view(np.uint64)), step
            assert np.array_equal(fp64.positions()[0].view(np.uint64), gb.view(np.uint64)), step
            log(f"C2 apply {step}: accepted {oa} rejected {orj} cycles {used}, positions bit-identical "
                f"({time.time() - t:.1f}s)")
    # size-independent properties over more applies: FP64 and mixed stay
    # identical, and the state stays overlap-free
    r64 = fp64.run(40, 42, 1.0, 0.08, 8)
    rmx = mixed.run(40, 42, 1.0, 0.08, 8)
    assert r64 == rmx
    b64 = fp64.positions()[0]
    bmx, smx = mixed.positions()
    assert np.array_equal(b64, bmx)
    check = O.OracleSystem(L, smx)
    assert check.total_energy() == 0.0
    log(f"C2: 8 more applies FP64 == mixed, overlap-free ({time.time() - t:.1f}s); "
        f"band re-tests {mixed.stats()['band_retests']}")
