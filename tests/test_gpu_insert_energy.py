"""GPU parity for hmc_delta_energy_insert (PairwiseCutoff::delta_energy_insert,
pairwise_cutoff.rs:1017-1025) against the oracle: random candidate bodies
over the whole box, each evaluated alone against the current state.  Hard
spheres: identical (0 or +inf); isotropic energies: within the returned
summation-order bound, and +inf exactly where the oracle's is."""
import ctypes as C
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


@pytest.mark.parametrize("dim,n,L,model,params,precision", [
    (3, 3000, [20.0, 21.0, 22.0], "hard_sphere", (1.0,), "mixed"),
    (2, 2000, [60.0, 62.0], "hard_sphere", (1.0,), "fp64"),
    (3, 2000, [18.0, 18.0, 18.0], "lj_shifted", (1.0, 1.0, 2.5, 2.5), "mixed"),
    (3, 2000, [18.0, 18.0, 18.0], "overlap_penalty", (1e4, 0.2, 100.0, 1.0), "fp64"),
    (2, 1500, [60.0, 60.0], "overlap_penalty", (50.0, 0.3, 2.0, 1.5), "fp64"),
])
def test_insert_energy_matches_oracle(hmc, wl, dim, n, L, model, params, precision):
    log(f"delta_energy_insert: D={dim} N={n} {model} {precision}")
    pos = wl.random_dilute(n, L, dim, 1.0, 21)
    ref = O.OracleSystem(L, pos, model=model, params=params, seed=4)
    b, s = ref.positions()
    gpu = hmc.GpuParallelSweep(L, b, sites=s, model=model, params=params, precision=precision)
    rng = np.random.default_rng(9)
    half = np.asarray(L) / 2
    cand = rng.uniform(-half, half, size=(4000, dim))
    e, bound = gpu.delta_energy_insert(cand)
    want = np.array([ref.delta_energy_insert(c) for c in cand])
    inf = np.isinf(want)
    assert np.array_equal(np.isinf(e), inf)
    assert (np.abs(e[~inf] - want[~inf]) <= bound[~inf]).all()
    if model == "hard_sphere":
        assert np.array_equal(e, want)
    assert (~inf).any() and (inf.any() or model == "lj_shifted")
    # the state is unchanged; a second call agrees within the bounds (the
    # order of the sites within a cell, and so the rounding, may differ)
    before = gpu.positions()
    e2, b2 = gpu.delta_energy_insert(cand[:10])
    fin = ~inf[:10]
    assert np.array_equal(np.isinf(e2), inf[:10])
    assert (np.abs(e2[fin] - e[:10][fin]) <= bound[:10][fin] + b2[fin]).all()
    after = gpu.positions()
    assert all(np.array_equal(x, y) for x, y in zip(before, after))
    gpu.close()


@pytest.mark.parametrize("model,params,precision", [
    ("hard_sphere", (1.0,), "mixed"),
    ("lj_shifted", (1.0, 1.0, 2.5, 2.5), "mixed"),
])
def test_insert_energy_after_applies(hmc, wl, model, params, precision):
    # insertion energies of a state the device has moved: the mixed
    # persistent applies write the FP64 sites with device-scope stores
    L = [16.0, 16.0, 16.0]
    pos = wl.random_dilute(1500, L, 3, 1.0, 5)
    ref = O.OracleSystem(L, pos, model=model, params=params, seed=6)
    b, s = ref.positions()
    gpu = hmc.GpuParallelSweep(L, b, sites=s, model=model, params=params, precision=precision)
    kT = 1.0 if model == "hard_sphere" else 1.5
    for _ in range(4):
        step, sub = ref.counter()
        g = gpu.apply(step, sub, 6, kT, 0.15)
        assert g[:2] == ref.parallel_apply(kT, 0.15)
        ref.increment_step()
    assert all(np.array_equal(x.view(np.uint64), y.view(np.uint64)) for x, y in zip(gpu.positions(),
                                                                                 ref.positions()))
    rng = np.random.default_rng(3)
    cand = rng.uniform(-8.0, 8.0, size=(3000, 3))
    e, bound = gpu.delta_energy_insert(cand)
    want = np.array([ref.delta_energy_insert(c) for c in cand])
    inf = np.isinf(want)
    assert np.array_equal(np.isinf(e), inf)
    assert (np.abs(e[~inf] - want[~inf]) <= bound[~inf]).all()
    if model == "hard_sphere":
        assert np.array_equal(e, want)
    gpu.close()


def test_insert_energy_rejects_convex_and_handles_empty(hmc, wl):
    L = [12.0, 12.0, 12.0]
    pos = wl.random_dilute(100, L, 3, 1.8, 4)
    cv = hmc.GpuParallelSweep(L, pos, model="convex", vertices=O.CUBE)
    with pytest.raises(hmc.HmcError) as e:
        cv.delta_energy_insert([[0.0, 0.0, 0.0]])
    assert e.value.code == hmc.HMC_ERR_UNSUPPORTED
    cv.close()
    hs = hmc.GpuParallelSweep(L, pos)
    e0, b0 = hs.delta_energy_insert(np.zeros((0, 3)))
    assert len(e0) == 0 and len(b0) == 0
    hs.close()
