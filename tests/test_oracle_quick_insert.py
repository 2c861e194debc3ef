"""The oracle's QuickInsert (hoomd-mc/src/quick_insert.rs:286-334 with a
UniformIn distribution over a periodic box, uniform_in.rs:99-110) against a
brute-force restatement: the candidate positions are drawn here from the
counter stream with a Python restatement of rand's Uniform, and each
candidate's insertion energy is decided from all-pairs minimum-image
distances in numpy -- no cell list, no ghosts, no C++ sampling.  Candidates
within 1e-9 of a decision threshold would make the comparison depend on
rounding; the assertions check that none occur.

Also the reference's own unit test (quick_insert.rs:338-401): target 10 in a
box of edge 6 starting from one body, complete after the inserts, total
energy 0 and 11 bodies (here on a periodic box with HardSphere, the sweep's
Hamiltonian; the reference's test uses a closed box with a Boxcar of the same
diameter).
"""
import math

import numpy as np
import pytest

import _oracle as O


def _wrap(x, L):
    # Periodic<Hypercuboid>::wrap: lambda = x / L, minus round (half away
    # from zero); 0.5 maps to -0.5
    lam = x / L
    r = math.floor(abs(lam))
    if abs(lam) - r >= 0.5:
        r += 1
    lam = lam - math.copysign(r, lam)
    if lam == 0.5:
        lam = -0.5
    return lam * L


def _min_image_r(t, sites, L):
    d = t[None, :] - sites
    d = d - L[None, :] * np.round(d / L[None, :])
    return np.sqrt((d * d).sum(axis=1))


def _brute_quick_insert(bodies, L, qi, step, sub, seed, model):
    """Returns (new bodies, accepted, rejected, closest threshold distance)."""
    D = len(L)
    target, allowed, inserted, complete = (int(x) for x in qi)
    if complete:
        return bodies, 0, 0, np.inf
    sites = np.array([[_wrap(_wrap(b[a], L[a]) + 0.0, L[a]) for a in range(D)] for b in bodies])

    def pair_decisions(t, pts):
        if len(pts) == 0:
            return False, False, np.inf
        r = _min_image_r(t, pts, L)
        if model == "hard_sphere":
            return bool((r < 1.0).any()), False, float(np.abs(r - 1.0).min())
        delta, dmax = 1.0, 0.2
        infinite = bool((r < delta - dmax).any())
        near = min(float(np.abs(r - delta).min()), float(np.abs(r - (delta - dmax)).min()))
        return infinite, bool((r < delta).any()), near

    # total energy <= 0: no pair closer than the threshold
    energy_zero = True
    margin = np.inf
    for i in range(len(sites)):
        inf_, ovl, near = pair_decisions(sites[i], np.delete(sites, i, axis=0))
        margin = min(margin, near)
        energy_zero &= not (inf_ or ovl)
    if energy_zero and inserted >= target:
        qi[3] = 1
        return bodies, 0, 0, margin
    acc = rej = 0
    out = [np.array(b, float) for b in bodies]
    if energy_zero:
        bits = O.counter_stream(step, sub, seed, 0, target * D)
        overl = 0
        for k in range(target):
            p = np.array([O.uniform_f64(bits[k * D + a], -L[a] / 2.0, L[a] / 2.0) for a in range(D)])
            t = np.array([_wrap(p[a] + 0.0, L[a]) for a in range(D)])
            inf_, ovl, near = pair_decisions(t, sites)
            margin = min(margin, near)
            if not inf_:
                b = np.array([_wrap(p[a], L[a]) for a in range(D)])
                out.append(b)
                sites = np.vstack([sites, [[_wrap(b[a] + 0.0, L[a]) for a in range(D)]]]) if len(sites) else \
                    np.array([[_wrap(b[a] + 0.0, L[a]) for a in range(D)]])
                acc += 1
                inserted += 1
                overl += 1 if ovl else 0
                if inserted == target or overl >= allowed:
                    break
            else:
                rej += 1
    qi[2] = inserted
    return out, acc, rej, margin


@pytest.mark.parametrize("dim,edge,target,model,seed", [
    (3, 8.0, 40, "hard_sphere", 5),
    (2, 12.0, 90, "hard_sphere", 11),
    (2, 10.0, 60, "overlap_penalty", 3),
    (3, 7.0, 50, "overlap_penalty", 8),
])
def test_quick_insert_matches_brute_force(dim, edge, target, model, seed):
    L = np.full(dim, edge)
    start = np.zeros((1, dim))
    # OverlapPenalty: k 10000, maximum overlap 0.2, shoulder 100, delta 1
    params = (1.0,) if model == "hard_sphere" else (10000.0, 0.2, 100.0, 1.0)
    ref = O.OracleSystem(L, start, model=model, params=params, seed=seed)
    qi = np.array([target, max(target // 8, 1), 0, 0], np.uint64)
    qb = qi.copy()
    bodies = [b for b in start]
    for call in range(6):
        step, sub = ref.counter()
        acc, rej = ref.quick_insert(qi)
        bodies, bacc, brej, margin = _brute_quick_insert(bodies, L, qb, step, sub, seed, model)
        assert margin > 1e-9
        assert (acc, rej) == (bacc, brej), call
        assert np.array_equal(qi, qb), (call, qi, qb)
        b, _ = ref.positions()
        assert np.array_equal(b, np.array(bodies)), call
        if model == "overlap_penalty" and not qi[3]:
            # the reference's use: translate sweeps relieve the overlaps
            for _ in range(3):
                ref.parallel_apply(1.0, 0.1)
            bodies = [x for x in ref.positions()[0]]
        ref.increment_step()
    assert qi[2] > 0


def test_reference_unit_test_hard_spheres():
    # quick_insert.rs:338-401: target 10, box edge 6, one body at the origin,
    # at most 100 applies; complete, 10 inserted, 11 bodies, energy 0
    L = np.array([6.0, 6.0])
    ref = O.OracleSystem(L, np.zeros((1, 2)), model="hard_sphere", params=(1.0,), seed=0)
    qi = np.array([10, max(10 // 8, 1), 0, 0], np.uint64)
    assert qi[1] == 1
    for _ in range(100):
        ref.quick_insert(qi)
        if qi[3]:
            break
    ref.parallel_apply(1.0, 0.1)
    assert qi[2] == 10 and qi[3] == 1
    assert ref.n == 11
    assert ref.total_energy() == 0.0


def test_complete_quick_insert_does_nothing():
    L = np.array([6.0, 6.0])
    ref = O.OracleSystem(L, np.zeros((1, 2)), model="hard_sphere", params=(1.0,), seed=2)
    qi = np.array([3, 1, 3, 1], np.uint64)
    _, sub = ref.counter()
    assert ref.quick_insert(qi) == (0, 0)
    assert ref.n == 1 and ref.counter()[1] == sub  # returns before increment_substep
