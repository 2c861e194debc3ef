"""GPU parity for hmc_clone_with_boundary (Microstate::clone_with_boundary
onto scaled box edges, the trial state of QuickCompress) against the
oracle: the new context's bodies and sites equal, bit for bit, the oracle's
Hypercuboid::map_point followed by the wrap of update_body_properties; sweeps
on the clone replay the oracle's on a microstate built from those bodies;
the source context is left unchanged."""
import ctypes as C
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


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


CASES = [
    # dim, n, L, volume factor, model, params, precision
    (3, 3000, [20.0, 21.0, 22.0], 0.9, "hard_sphere", (1.0,), "mixed"),
    (3, 3000, [20.0, 21.0, 22.0], 0.97, "hard_sphere", (1.0,), "fp64"),
    (2, 2500, [70.0, 72.0], 1.1, "hard_sphere", (1.0,), "mixed"),
    (3, 2000, [18.0, 18.0, 18.0], 0.95, "lj_shifted", (1.0, 1.0, 2.5, 2.5), "mixed"),
    (3, 1500, [16.0, 16.0, 16.0], 0.9, "overlap_penalty", (1e4, 0.2, 100.0, 1.0), "fp64"),
]


@pytest.mark.parametrize("dim,n,L,v,model,params,precision", CASES)
def test_clone_matches_oracle_and_replays(hmc, wl, dim, n, L, v, model, params, precision):
    log(f"clone_with_boundary: D={dim} N={n} v={v} {model} {precision}")
    pos = wl.random_dilute(n, L, dim, 1.0, 11)
    ref = O.OracleSystem(L, pos, model=model, params=params, seed=9)
    b, s = ref.positions()
    gpu = hmc.GpuParallelSweep(L, b, sites=s, model=model, params=params, precision=precision)
    for _ in range(3):
        step, sub = ref.counter()
        gpu.apply(step, sub, 9, 1.0, 0.1)
        ref.parallel_apply(1.0, 0.1)
        ref.increment_step()
    b0, s0 = gpu.positions()
    L2 = hmc.scale_volume(L, v)
    clone = gpu.clone_with_boundary(L2)
    assert clone.box_edges == L2
    # the source is unchanged
    b1, s1 = gpu.positions()
    assert np.array_equal(bits(b0), bits(b1)) and np.array_equal(bits(s0), bits(s1))
    # bodies: map_point, then the wrap of update_body_properties (as add_body applies it)
    mapped, n_out = O.map_points(L, L2, b0)
    assert n_out == 0
    step, sub = ref.counter()
    ref2 = O.OracleSystem(L2, mapped, model=model, params=params, seed=9, step=step)
    rb, rs = ref2.positions()
    cb, cs = clone.positions()
    assert np.array_equal(bits(cb), bits(rb)) and np.array_equal(bits(cs), bits(rs))
    # sweeps on the clone replay the oracle's
    for k in range(4):
        step, sub = ref2.counter()
        ga, gr, used = clone.apply(step, sub, 9, 1.0, 0.1)
        oa, orj = ref2.parallel_apply(1.0, 0.1)
        _, sub2 = ref2.counter()
        assert (ga, gr, used) == (oa, orj, sub2 - sub), k
        ref2.increment_step()
    cb, cs = clone.positions()
    rb, rs = ref2.positions()
    assert np.array_equal(bits(cb), bits(rb)) and np.array_equal(bits(cs), bits(rs))
    if model != "hard_sphere":
        e, bound = clone.total_energy()
        assert abs(e - ref2.total_energy()) <= bound
        assert clone.stats()["ambiguous"] == 0
    clone.close()
    gpu.close()


def test_clone_convex_keeps_orientations(hmc, wl):
    L = [12.0, 12.0, 12.0]
    pos = wl.random_dilute(150, L, 3, 1.8, 4)
    q = np.random.default_rng(5).normal(size=(150, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    gpu = hmc.GpuParallelSweep(L, wl.wrap(L, pos), model="convex", vertices=O.CUBE, orientations=q)
    bq0, sq0 = gpu.orientations()
    b0, _ = gpu.positions()
    L2 = hmc.scale_volume(L, 0.9)
    clone = gpu.clone_with_boundary(L2)
    bq, sq = clone.orientations()
    assert np.array_equal(bits(bq), bits(bq0)) and np.array_equal(bits(sq), bits(sq0))
    mapped, n_out = O.map_points(L, L2, b0)
    rb, rs = O.OracleSystem(L2, mapped).positions()
    cb, cs = clone.positions()
    assert n_out == 0 and np.array_equal(bits(cb), bits(rb)) and np.array_equal(bits(cs), bits(rs))
    clone.apply_rotate(0, 0, 3, 1.0, 0.1)
    clone.close()
    gpu.close()


def test_clone_rejects_a_box_too_small_for_the_ranges(hmc, wl):
    L = [6.0, 6.0, 6.0]
    pos = wl.random_dilute(50, L, 3, 1.0, 7)
    gpu = hmc.GpuParallelSweep(L, pos, model="lj", params=(1.0, 1.0, 2.5))
    with pytest.raises(hmc.HmcError) as e:
        gpu.clone_with_boundary([4.0, 6.0, 6.0])  # r_cut 2.5 > L/2
    assert e.value.code == hmc.HMC_ERR_INVALID
    # the source still runs
    gpu.apply(0, 0, 1, 1.0, 0.1)
    gpu.close()
