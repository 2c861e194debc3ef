"""The oracle's move-size tuner (tune_local.rs:21-89, lib.rs:514-524) and
GrandCanonical (grand_canonical.rs:145-205), pinned by what their formulas
fix exactly and by the ideal-gas limit:
- a tuner round whose samples are all accepted (or all rejected) multiplies
  the move size by (1 + 1.5)/(target + 1.5) (or 1.5/(target + 1.5));
- tuning leaves the microstate and its counter unchanged;
- at fugacity 0 every removal is accepted and no insertion is;
- for small spheres at low density (an ideal gas within 0.1%),
  GrandCanonical samples a Poisson number of bodies of mean V f."""
import numpy as np
import pytest

import _oracle as O


def _lattice(n_side, a, dim):
    g = np.stack(np.meshgrid(*[np.arange(n_side)] * dim, indexing="ij"), -1).reshape(-1, dim)
    L = np.full(dim, n_side * a)
    return L, (g + 0.5) * a - L / 2


@pytest.mark.parametrize("dim", [2, 3])
def test_tune_all_accepted_scales_by_closed_form(dim):
    # small spheres far apart: every trial is accepted
    L, pos = _lattice(6, 3.0, dim)
    s = O.OracleSystem(L, pos, params=(0.5,), boundary_range=1.0, body_range=1.0, seed=5)
    m, counts = s.tune_translate(1.0, 0.3, 500, 6, 0.1)
    assert counts == [(500, 0)] * 6
    want = 0.1
    for _ in range(6):
        want *= (1.0 + 1.5) / (0.3 + 1.5)
    assert m == want


def test_tune_all_rejected_scales_by_closed_form():
    # a dense square lattice of disks touching their neighbours: a random
    # move of the given size always overlaps
    L, pos = _lattice(8, 1.0, 2)
    s = O.OracleSystem(L, pos, params=(0.999999,), boundary_range=1.0, body_range=1.0, seed=2)
    m, counts = s.tune_translate(1.0, 0.2, 300, 3, 0.2)
    # acceptance is tiny but not exactly zero: check the closed form per step
    m2 = 0.2
    for acc, rej in counts:
        assert acc + rej == 300
        m2 *= (acc / 300.0 + 1.5) / (0.2 + 1.5)
    assert m == m2
    assert sum(a for a, _ in counts) < 30


def test_tune_leaves_the_state_unchanged():
    L, pos = _lattice(6, 1.2, 3)
    s = O.OracleSystem(L, pos, seed=9)
    s.parallel_apply(1.0, 0.1)
    b0, s0 = s.positions()
    c0 = s.counter()
    m, counts = s.tune_translate(1.0, 0.2, 1000, 4, 0.1)
    b1, s1 = s.positions()
    assert np.array_equal(b0, b1) and np.array_equal(s0, s1) and s.counter() == c0
    # the same call again gives the same result (the stream is the counter's)
    assert s.tune_translate(1.0, 0.2, 1000, 4, 0.1) == (m, counts)


def test_tune_reaches_the_target():
    L, pos = _lattice(8, 1.1, 3)
    s = O.OracleSystem(L, pos, seed=3)
    m, counts = s.tune_translate(1.0, 0.3, 4000, 24, 0.01)
    acc, rej = counts[-1]
    assert abs(acc / (acc + rej) - 0.3) < 0.05
    assert m > 0.01


def test_gc_zero_fugacity_only_removes():
    L, pos = _lattice(4, 1.5, 3)
    s = O.OracleSystem(L, pos, seed=11)
    n0 = s.n
    tot = np.zeros(4, int)
    for _ in range(40):
        c = s.grand_canonical(1.0, 0.0)
        tot += c
        s.increment_step()
    assert tot[0] == 0 and tot[3] == 0  # no insertion accepted, no removal rejected
    assert s.n == n0 - tot[2] and tot[1] + tot[2] == 40


def test_gc_ideal_gas_is_poisson():
    # spheres of diameter 0.2 in a box of volume 1000: at 20 bodies the
    # excluded volume shifts <n> by under 0.1%; <n> = V f = 20
    L = np.array([10.0, 10.0, 10.0])
    s = O.OracleSystem(L, np.zeros((0, 3)), params=(0.2,), boundary_range=1.0, body_range=1.0, seed=1)
    ns = []
    for k in range(30000):
        s.grand_canonical(1.0, 0.02)
        s.increment_step()
        if k >= 2000:
            ns.append(s.n)
    ns = np.array(ns, float)
    assert abs(ns.mean() - 20.0) < 1.0
    assert abs(ns.var() - 20.0) < 6.0


def test_convex_tune_closed_forms():
    # unit cubes 3 apart (circumsphere diameter 1.73): no move of these sizes
    # makes an overlap, so every sample is accepted; a rotation size stops at
    # pi/2 (rotate/versor.rs:125-137)
    L, pos = _lattice(4, 3.0, 3)
    q = np.tile([1.0, 0.0, 0.0, 0.0], (len(pos), 1))
    ref = O.ConvexOracle(L, pos, q, O.CUBE, seed=3)
    m, counts = ref.tune(True, 0.2, 200, 6, 1.0)
    assert counts == [(200, 0)] * 6
    want = 1.0
    for _ in range(6):
        want = min(want * (2.5 / 1.7), np.pi / 2.0)
    assert m == want == np.pi / 2.0
    m, counts = ref.tune(False, 0.2, 200, 4, 0.1)
    want = 0.1
    for _ in range(4):
        want *= 2.5 / 1.7
    assert counts == [(200, 0)] * 4 and m == want
