"""GPU parity of the move-size tuner (ParallelSweep::tune with Translate,
tune_local.rs:21-89) and of GrandCanonical (grand_canonical.rs:145-205)
through the C ABI, against the oracle: the tuned maximum distance bit for
bit and every round's counts; insertion/removal counts and the final body
list bit for bit over GrandCanonical moves interleaved with sweeps."""
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


def _system(wl, kind):
    """(L, positions, GPU model kwargs, oracle kwargs, kT)"""
    if kind == "hs3d":
        L, pos = wl.fcc_spheres((6, 6, 6), packing_fraction=0.45)
        return L, pos, {}, {}, 1.0
    if kind == "hs2d":
        L, pos = wl.square_lattice_disks(n_side=30, packing_fraction=0.6)
        return L, pos, {}, {}, 1.0
    if kind == "lj3d":
        L, pos = wl.fcc_number_density((5, 5, 5), 0.8)
        prm = (1.0, 1.0, 2.5)
        return L, pos, {"model": "lj", "params": prm}, {"model": "lj", "params": prm}, 1.2
    L, pos = wl.fcc_spheres((5, 5, 5), packing_fraction=0.5)
    prm = (10000.0, 0.2, 100.0, 1.0)
    return L, pos, {"model": "overlap_penalty", "params": prm}, {"model": "overlap_penalty", "params": prm}, 1.0


@pytest.mark.parametrize("kind", ["hs3d", "hs2d", "lj3d", "op3d"])
def test_tune_matches_oracle(hmc, wl, kind):
    L, pos, gk, ok, kT = _system(wl, kind)
    body = wl.wrap(L, pos)
    log(f"tune {kind}: N={len(pos)}")
    ref = O.OracleSystem(L, body, step=3, seed=7, **ok)
    sim = hmc.GpuParallelSweep(L, body, **gk)
    # a few sweeps first (same counters on both)
    for _ in range(3):
        step, sub = ref.counter()
        g = sim.apply(step, sub, 7, kT, 0.1)
        assert g[:2] == ref.parallel_apply(kT, 0.1)
    step, sub = ref.counter()
    m_ref, c_ref = ref.tune_translate(kT, 0.25, 3000, 6, 0.05)
    m_gpu, c_gpu, amb = sim.tune_translate(step, sub, 7, kT, 0.25, 3000, 6, 0.05)
    log(f"  counts {c_gpu}, max move {m_gpu:.6f}, ambiguous {amb}")
    assert c_gpu == c_ref
    assert np.float64(m_gpu).view(np.uint64) == np.float64(m_ref).view(np.uint64)
    assert amb == 0
    # the state is not modified
    b0, s0 = ref.positions()
    b1, s1 = sim.positions()
    assert np.array_equal(b0.view(np.uint64), b1.view(np.uint64))
    sim.close()


def test_tune_c2_default_options(hmc, wl):
    # the C2 system (N = 1,000,188 hard spheres, packing fraction 0.50) with
    # TuneOptions::default(): target 0.2, 8,000 samples, 32 steps
    L, pos = wl.fcc_spheres((63, 63, 63), packing_fraction=0.50)
    body = wl.wrap(L, pos)
    ref = O.OracleSystem(L, body, step=0, seed=42)
    sim = hmc.GpuParallelSweep(L, body)
    m_ref, c_ref = ref.tune_translate(1.0, 0.2, 8000, 32, 0.08)
    m_gpu, c_gpu, amb = sim.tune_translate(0, 0, 42, 1.0, 0.2, 8000, 32, 0.08)
    log(f"C2 tune: max move 0.08 -> {m_gpu:.6f}, last round {c_gpu[-1]}")
    assert c_gpu == c_ref and m_gpu == m_ref and amb == 0
    sim.close()


@pytest.mark.parametrize("kind,fugacity,precision", [("hs3d", 0.5, "fp64"), ("hs3d", 0.5, "mixed"),
                                                      ("lj3d", 0.03, "fp64")])
def test_grand_canonical_with_sweeps_matches_oracle(hmc, wl, kind, fugacity, precision):
    L, pos, gk, ok, kT = _system(wl, kind)
    gk = dict(gk, precision=precision)
    body = wl.wrap(L, pos)
    if kind == "hs3d":  # a fluid at packing fraction 0.1
        L, pos = wl.fcc_spheres((4, 4, 4), packing_fraction=0.10)
    else:  # LJ at number density 0.3
        L, pos = wl.fcc_number_density((5, 5, 5), 0.3)
    body = wl.wrap(L, pos)
    ref = O.OracleSystem(L, body, seed=11, **ok)
    sim = hmc.GpuParallelSweep(L, body, **gk)
    tot = np.zeros(4, int)
    for k in range(60):
        step, sub = ref.counter()
        g = sim.grand_canonical(step, sub, 11, kT, fugacity)
        o = ref.grand_canonical(kT, fugacity)
        assert g[:4] == o, (k, g, o)
        assert g[4] == 0
        tot += o
        step, sub = ref.counter()
        a = sim.apply(step, sub, 11, kT, 0.1)
        assert a[:2] == ref.parallel_apply(kT, 0.1), k
        ref.increment_step()
        assert sim.n == ref.n
    log(f"GrandCanonical {kind}: N {len(pos)} -> {ref.n}; insert +{tot[0]}/-{tot[1]}, remove +{tot[2]}/-{tot[3]}")
    assert tot[0] > 0 and tot[2] > 0
    b0, s0 = ref.positions()
    b1, s1 = sim.positions()
    assert np.array_equal(b0.view(np.uint64), b1.view(np.uint64))
    assert np.array_equal(s0.view(np.uint64), s1.view(np.uint64))
    sim.close()


@pytest.mark.parametrize("move", ["translate", "rotate"])
def test_tune_convex_matches_oracle(hmc, wl, move):
    # 216 unit cubes, simple cubic at packing fraction 0.55, orientations
    # randomised by rotation sweeps on both sides first
    n = 6
    Lb = (n**3 / 0.55) ** (1.0 / 3.0)
    a = Lb / n
    g = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3)
    L = np.array([Lb, Lb, Lb])
    pos = (g + 0.5) * a - Lb / 2
    q = np.tile([1.0, 0.0, 0.0, 0.0], (len(pos), 1))
    ref = O.ConvexOracle(L, pos, q, O.CUBE, seed=13)
    bp, bq, sp, sq = ref.state()
    gpu = hmc.GpuParallelSweep(L, bp, sites=sp, model="convex", vertices=O.CUBE, orientations=bq,
                               site_orientations=sq)
    for _ in range(4):
        step, sub = ref.counter()
        gg = gpu.apply_rotate(step, sub, 13, 1.0, 0.5)
        assert gg[:2] == ref.apply(True, 0.5)
        ref.increment_step()
    step, sub = ref.counter()
    size = 0.05 if move == "translate" else 0.1
    m_ref, c_ref = ref.tune(move == "rotate", 0.3, 2000, 6, size)
    m_gpu, c_gpu, amb = gpu.tune(move, step, sub, 13, 1.0, 0.3, 2000, 6, size)
    log(f"tune convex {move}: {size} -> {m_gpu:.6f}, counts {c_gpu}")
    assert c_gpu == c_ref and m_gpu == m_ref and amb == 0
    gpu.close()
