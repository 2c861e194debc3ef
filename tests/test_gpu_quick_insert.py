"""GPU parity for QuickInsert (hmc_quick_insert_apply: hoomd-mc/src/
quick_insert.rs:94-334 with UniformIn over the periodic box) against the
oracle's restatement, in the placement loop of the reference's documentation
example (quick_insert.rs:211-262: a QuickInsert call, then a translate sweep
to relieve the overlaps), here every step with the parallel sweep.  At every
call the counts, the inserted total and the completion flag agree; every
sweep's counts agree; the bodies and sites agree bit for bit after every
call, and no sign test or candidate pair fell inside its rounding bound."""
import sys

import numpy as np
import pytest

import _oracle as O
import _pkg

pytestmark = pytest.mark.gpu

P = (1e4, 0.2, 100.0, 1.0)  # OverlapPenalty: k, maximum overlap, shoulder, delta


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hmc():
    p = _pkg.pkg()
    p.lib()
    return p


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


@pytest.mark.parametrize("dim,edge,target,model,params,d,max_steps,precision", [
    (3, 14.0, 1500, "hard_sphere", (1.0,), 0.1, 40, "fp64"),
    (3, 14.0, 1500, "hard_sphere", (1.0,), 0.1, 40, "mixed"),
    (2, 30.0, 500, "hard_sphere", (1.0,), 0.15, 40, "fp64"),
    (3, 10.0, 700, "overlap_penalty", P, 0.1, 120, "fp64"),
    (2, 24.0, 450, "overlap_penalty", P, 0.15, 120, "fp64"),
])
def test_quick_insert_matches_oracle(hmc, dim, edge, target, model, params, d, max_steps, precision):
    log(f"QuickInsert: D={dim} L={edge} target {target} {model} {precision}")
    L = [edge] * dim
    seed = 7
    start = np.zeros((1, dim))
    ref = O.OracleSystem(L, start, model=model, params=params, seed=seed)
    b, s = ref.positions()
    gpu = hmc.GpuParallelSweep(L, b, sites=s, model=model, params=params, precision=precision)
    oq = np.array([target, max(target // 8, 1), 0, 0], np.uint64)
    gq = hmc.QuickInsert(target)
    calls = 0
    for _ in range(max_steps):
        step, sub = ref.counter()
        ga, gr, used = gq.apply(gpu, step, sub, seed)
        oa, orj = ref.quick_insert(oq)
        calls += 1
        assert (ga, gr) == (oa, orj), step
        assert gq.inserted == int(oq[2]) and gq.is_complete() == bool(oq[3]), step
        assert used == ref.counter()[1] - sub, step
        assert gpu.n == ref.n, step
        gb, gs = gpu.positions()
        rb, rs = ref.positions()
        assert np.array_equal(bits(gb), bits(rb)) and np.array_equal(bits(gs), bits(rs)), step
        if gq.is_complete():
            break
        step, sub = ref.counter()
        ga, gr, used = gpu.apply(step, sub, seed, 1.0, d)
        oa, orj = ref.parallel_apply(1.0, d)
        assert (ga, gr) == (oa, orj), step
        ref.increment_step()
    log(f"  {calls} calls, inserted {gq.inserted} of {target}, complete {gq.is_complete()}, N {gpu.n}, "
        f"layout {gpu.stats()['tiled']}")
    assert gq.inserted > 0 and gq.ambiguous == 0
    if model == "hard_sphere":
        assert gq.is_complete() and gpu.n == target + 1
    gb, gs = gpu.positions()
    rb, rs = ref.positions()
    assert np.array_equal(bits(gb), bits(rb)) and np.array_equal(bits(gs), bits(rs))
    gpu.close()


def test_quick_insert_rejects_oriented_bodies(hmc):
    # UniformIn of OrientedPoint bodies also draws an orientation (rand's
    # StandardUniform for Versor), which is not restated: refused
    wl = _pkg.workloads()
    L, pos = wl.simple_cubic(n_side=4, packing_fraction=0.3)
    q0 = np.tile([1.0, 0.0, 0.0, 0.0], (len(pos), 1))
    sim = hmc.GpuParallelSweep(L, wl.wrap(L, pos), model="convex", vertices=wl.UNIT_CUBE, orientations=q0)
    with pytest.raises(hmc.HmcError):
        hmc.QuickInsert(10).apply(sim, 0, 0, 1)
    sim.close()
