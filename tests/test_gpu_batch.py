"""GPU parity for replica batches: every replica of a batch makes exactly the
trials of its own oracle system (counts per apply summed over the run, and
final FP64 positions bit for bit), with and without the captured hipGraph.
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


def _oracle_run(L, pos, seed, d, steps):
    ref = O.OracleSystem(L, pos, seed=seed)
    b, s = ref.positions()
    tot = [0, 0]
    for _ in range(steps):
        a, r = ref.parallel_apply(1.0, d)
        ref.increment_step()
        tot[0] += a
        tot[1] += r
    return (b, s), tuple(tot), ref.positions()


@pytest.mark.parametrize("layout", ["persistent", "tiles", "lanes"])
@pytest.mark.parametrize("use_graph", [True, False], ids=["graph", "direct"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_replicas_match_their_own_oracles(hmc, wl, precision, use_graph, layout):
    # 864 .. 2,916 bodies: at least 5 spaces per axis, so the tile layout applies
    systems = wl.batched_replicas(n_replicas=6, k_range=(6, 9), phi_range=(0.30, 0.54))
    seeds = [42 + r for r in range(6)]
    d = [0.08, 0.1, 0.12, 0.05, 0.08, 0.2]
    steps = 5
    refs = [_oracle_run(L, pos, seeds[r], d[r], steps) for r, (L, pos) in enumerate(systems)]
    batch = hmc.GpuBatch([(L, refs[r][0][0], refs[r][0][1]) for r, (L, _) in enumerate(systems)],
                         precision=precision, lds_tiles=layout != "lanes", persistent=layout == "persistent")
    want = {"persistent": 2 if precision == "mixed" else 1, "tiles": 1, "lanes": 0}[layout]
    assert batch.stats(0)["tiled"] == want
    counts = batch.run(0, seeds, 1.0, d, steps, use_graph=use_graph)
    for r in range(6):
        assert counts[r] == refs[r][1], (r, counts[r], refs[r][1])
        gb, gs = batch.positions(r)
        ob, os_ = refs[r][2]
        assert np.array_equal(gb.view(np.uint64), ob.view(np.uint64)), r
        assert np.array_equal(gs.view(np.uint64), os_.view(np.uint64)), r
    batch.close()


def test_batch_graph_replays_equal_direct_launches(hmc, wl):
    # one run of 8 steps (first step direct, then 7 replays of the captured
    # graph, with host work in between) against 8 directly launched steps
    systems = wl.batched_replicas(n_replicas=10, k_range=(4, 8))
    seeds = list(range(100, 110))
    a = hmc.GpuBatch([(L, p, None) for L, p in systems])
    b = hmc.GpuBatch([(L, p, None) for L, p in systems])
    ca = a.run(0, seeds, 1.0, 0.08, 4, use_graph=True)
    a.positions(0)
    ca2 = a.run(4, seeds, 1.0, 0.08, 4, use_graph=True)
    cb = b.run(0, seeds, 1.0, 0.08, 8, use_graph=False)
    assert [(x[0] + y[0], x[1] + y[1]) for x, y in zip(ca, ca2)] == cb
    for r in range(len(systems)):
        assert np.array_equal(a.positions(r)[0], b.positions(r)[0])
    a.close()
    b.close()


@pytest.mark.parametrize("persistent", [True, False], ids=["persistent", "percolour"])
def test_batch_config_c5_full(hmc, wl, persistent):
    # 256 replicas, N log-uniform over FCC counts 4 k^3 (k = 8..25), packing
    # fractions 0.30 .. 0.54, seeds 42 + r, d = 0.08: 10 steps, six replicas
    # checked against their own oracles, all replicas overlap-free
    systems = wl.batched_replicas()
    n_total = sum(len(p) for _, p in systems)
    seeds = [42 + r for r in range(len(systems))]
    t = time.time()
    # the microstates' stored (wrapped) bodies; sites follow as wrap(body + 0)
    batch = hmc.GpuBatch([(L, O.wrap(L, p), None) for L, p in systems], precision="mixed", persistent=persistent)
    assert batch.stats(0)["tiled"] == (2 if persistent else 1)
    counts = batch.run(0, seeds, 1.0, 0.08, 10)
    log(f"C5 ({'persistent' if persistent else 'per-colour'}): {len(systems)} replicas, N total {n_total}: "
        f"10 steps ({time.time() - t:.1f}s)")
    for r in (0, 51, 101, 153, 204, 255):
        L, pos = systems[r]
        _, tot, (ob, os_) = _oracle_run(L, pos, seeds[r], 0.08, 10)
        assert counts[r] == tot, (r, counts[r], tot)
        gb, gs = batch.positions(r)
        assert np.array_equal(gb, ob) and np.array_equal(gs, os_), r
    for r in range(0, len(systems), 37):
        _, gs = batch.positions(r)
        assert O.OracleSystem(systems[r][0], gs).total_energy() == 0.0, r
    batch.close()


def test_batch_autotune_leaves_state_and_results_unchanged(hmc, wl):
    # the autotuner times several tile layouts from the current state and
    # restores it after each; the chosen layout makes the same decisions
    systems = wl.batched_replicas(n_replicas=12, k_range=(6, 12))
    seeds = [7 + r for r in range(12)]
    a = hmc.GpuBatch([(L, p, None) for L, p in systems], precision="mixed", persistent=False)
    b = hmc.GpuBatch([(L, p, None) for L, p in systems], precision="mixed", persistent=False)
    ms, policy = a.autotune(0, seeds, 1.0, 0.08, n_steps=2)
    log(f"autotune: {ms:.3f} ms per step with policy {policy}")
    assert ms > 0.0 and policy in (1, 4, 16, 64, 256)
    assert a.stats(0)["tiled"] == 1
    assert all(a.stats(r)["cycles"] == 0 for r in range(12))
    for r in range(12):
        assert np.array_equal(a.positions(r)[0], b.positions(r)[0]), r
    ca = a.run(0, seeds, 1.0, 0.08, 4)
    cb = b.run(0, seeds, 1.0, 0.08, 4)
    assert ca == cb
    for r in range(12):
        assert np.array_equal(a.positions(r)[1], b.positions(r)[1]), r
    a.close()
    b.close()


def test_batch_autotune_needs_the_tile_layout(hmc, wl):
    # a batch on the lane-per-space pass has no tile sizes to tune
    systems = wl.batched_replicas(n_replicas=3, k_range=(6, 7))
    b = hmc.GpuBatch([(L, p, None) for L, p in systems], lds_tiles=False)
    assert b.stats(0)["tiled"] == 0
    with pytest.raises(hmc.HmcError) as e:
        b.autotune(0, [1, 2, 3], 1.0, 0.08)
    assert e.value.code == hmc.HMC_ERR_UNSUPPORTED
    b.close()
