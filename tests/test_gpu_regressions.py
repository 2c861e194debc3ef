"""GPU regression tests for fixed defects, each against the oracle.

* The single-pass look-back scan of an apply's build tags its status words
  with the apply's epoch.  Energy, insertion-energy and SDF calls also run the
  build (with a three-kernel scan) and must not advance that epoch: if they
  did, 255 of them between two applies would make the previous apply's words
  read as current, and the next build would take a stale prefix.
* Insertion energies (PairwiseCutoff::delta_energy_insert,
  pairwise_cutoff.rs:1017-1025) after applies: the state the sweeps wrote
  (device-scope site stores of the persistent apply, the tile-overflow path)
  is the state the energies read.
"""
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


def _apply_both(gpu, ref, seed, kT, d):
    step, sub = ref.counter()
    got = gpu.apply(step, sub, seed, kT, d)
    oa, orj = ref.parallel_apply(kT, d)
    _, sub2 = ref.counter()
    ref.increment_step()
    assert got == (oa, orj, sub2 - sub), (got, oa, orj, sub2 - sub)


@pytest.mark.parametrize("between", ["total_energy", "insert", "sdf"])
@pytest.mark.parametrize("n_calls", [255, 256])
def test_non_apply_builds_between_applies_keep_replay(hmc, wl, between, n_calls):
    # an apply, then n_calls builds of another kind, then more applies: counts
    # and final positions stay bit-identical to the oracle's
    log(f"look-back epoch: {n_calls} x {between} between applies")
    L, pos = wl.fcc_spheres((8, 8, 8), packing_fraction=0.50)
    ref = O.OracleSystem(L, pos, seed=42)
    b, s = ref.positions()
    gpu = hmc.GpuParallelSweep(L, b, sites=s, precision="mixed")
    cand = np.zeros((4, 3))
    for rnd in range(3):
        _apply_both(gpu, ref, 42, 1.0, 0.08)
        for _ in range(n_calls):
            if between == "total_energy":
                e, _ = gpu.total_energy()
                assert e == 0.0
            elif between == "insert":
                gpu.delta_energy_insert(cand)
            else:
                gpu.sdf(xmax=0.02, dx=1e-3)
    _apply_both(gpu, ref, 42, 1.0, 0.08)
    gb, gs = gpu.positions()
    ob, os_ = ref.positions()
    assert np.array_equal(gb.view(np.uint64), ob.view(np.uint64))
    assert np.array_equal(gs.view(np.uint64), os_.view(np.uint64))
    assert gpu.stats()["applies"] == 4
    gpu.close()


def _clustered(wl):
    # a dense FCC block in one corner of a mostly empty box (tiles overflow)
    L1, block = wl.fcc_spheres((6, 6, 6), packing_fraction=0.50)
    L = np.array([24.5, 24.5, 24.5])
    return L, block - (-L1 / 2) + (-L / 2 + 0.75)


@pytest.mark.parametrize("case", ["mixed_persistent", "mixed_overflow", "lj_mixed"])
def test_insert_energy_after_applies(hmc, wl, case):
    log(f"insertion energies after applies: {case}")
    kT, d = 1.0, 0.08
    if case == "mixed_persistent":
        # dilute enough that some candidates fit between the lattice sites
        L, pos = wl.fcc_spheres((10, 10, 10), packing_fraction=0.15)
        model, params, precision = "hard_sphere", (1.0,), "mixed"
    elif case == "mixed_overflow":
        L, pos = _clustered(wl)
        model, params, precision = "hard_sphere", (1.0,), "mixed"
    else:
        L, pos = wl.fcc_number_density((8, 8, 8), 0.8)
        model, params, precision = "lj_shifted", (1.0, 1.0, 2.5, 2.5), "mixed"
        kT, d = 1.2, 0.1
    ref = O.OracleSystem(L, pos, model=model, params=params, seed=42)
    b, s = ref.positions()
    gpu = hmc.GpuParallelSweep(L, b, sites=s, model=model, params=params, precision=precision)
    for _ in range(6):
        _apply_both(gpu, ref, 42, kT, d)
    st = gpu.stats()
    if case == "mixed_persistent":
        assert st["tiled"] == 2
    if case == "mixed_overflow":
        assert st["tile_overflows"] > 0
    rng = np.random.default_rng(17)
    half = np.asarray(L) / 2
    cand = rng.uniform(-half, half, size=(3000, len(L)))
    # also candidates right next to bodies that moved
    gb, _ = gpu.positions()
    near = gb[rng.integers(0, len(gb), 500)] + rng.normal(0.0, 0.6, size=(500, len(L)))
    cand = np.concatenate([cand, ((near + half) % (2 * half)) - half])
    e, bound = gpu.delta_energy_insert(cand)
    want = np.array([ref.delta_energy_insert(c) for c in cand])
    inf = np.isinf(want)
    assert np.array_equal(np.isinf(e), inf)
    if model == "hard_sphere":
        assert np.array_equal(e, want)
        assert inf.any() and (~inf).any()
    else:
        assert (np.abs(e[~inf] - want[~inf]) <= bound[~inf]).all()
    gpu.close()
