"""The oracle's overlap-penalty pair energy (oracle/mc.hpp PairModel kind 4:
Isotropic { Expanded { delta, OverlapPenalty }, r_cut: delta }, the
placement Hamiltonian of benchmarks/src/mc/hard_sphere.rs:192-198).

Pinned by the reference's own known answers (overlap_penalty.rs tests
`shoulder`, `core`, `harmonic`) evaluated on a two-body system, by the
early return of delta_energy_one on an infinite final energy
(pairwise_cutoff.rs:919-941), and by an independent all-pairs sum."""
import math

import numpy as np
import pytest

import _oracle as O
import _pkg

wl = _pkg.workloads()
P = (1e4, 0.2, 100.0, 1.0)  # OverlapPenalty::default(), delta 1


def two_body_energy(params, x, dim=3):
    """Total energy of two bodies at distance delta + x (the box edge is a
    power of two, so that wrapping, (r / L - round) L, keeps r; sqrt(r^2) = r
    exactly in IEEE arithmetic, and r - delta = x exactly here)."""
    delta = params[3]
    r = delta + x
    assert r - delta == x
    L = [8.0] * dim
    pos = np.zeros((2, dim))
    pos[1, 0] = r
    s = O.OracleSystem(L, pos, model="overlap_penalty", params=params)
    return s.total_energy()


def test_shoulder():
    # overlap_penalty.rs `shoulder`: k 0, eps 15, d_max 1
    p = (0.0, 1.0, 15.0, 1.0)
    assert two_body_energy(p, 1.0) == 0.0
    assert two_body_energy(p, 0.0) == 0.0
    assert two_body_energy(p, -2.0**-52) == 15.0  # just below contact
    assert two_body_energy(p, -0.5) == 15.0


def test_core():
    # overlap_penalty.rs `core`: k 0, eps 0, d_max 0.5
    p = (0.0, 0.5, 0.0, 1.0)
    assert two_body_energy(p, 1.0) == 0.0
    assert two_body_energy(p, 0.0) == 0.0
    assert two_body_energy(p, -2.0**-52) == 0.0
    assert two_body_energy(p, -0.5) == 0.0
    assert two_body_energy(p, math.nextafter(-0.5, -1.0)) == math.inf
    assert two_body_energy(p, -0.75) == math.inf


def test_harmonic():
    # overlap_penalty.rs `harmonic`: k 6, eps 0, d_max 10 (here at delta 1
    # and 2: the expansion shifts the argument only)
    for delta in (1.0, 2.0):
        p = (6.0, 10.0, 0.0, delta)
        assert two_body_energy(p, 1.0) == 0.0
        assert two_body_energy(p, 0.0) == 0.0
        assert two_body_energy(p, -0.125) == 0.5 * 6.0 * 0.125**2
    # the default parameters (k 10^4, d_max 0.2, eps 100) in 2D
    assert two_body_energy((1e4, 0.2, 100.0, 1.0), -0.125, dim=2) == 0.5 * 1e4 * 0.125**2 + 100.0


def test_delta_energy_one_infinite_final_and_initial():
    p = (1e4, 0.2, 100.0, 1.0)
    L = [8.0, 8.0, 8.0]  # a power of two: wrapping keeps the positions
    pos = np.array([[0.0, 0.0, 0.0], [0.9, 0.0, 0.0], [3.0, 0.0, 0.0]])
    s = O.OracleSystem(L, pos, model="overlap_penalty", params=p)
    e1 = 0.5 * 1e4 * 0.1**2 + 100.0
    # body 1 from 0.9 (penalised) to 0.5 (beyond d_max): infinite
    assert s.delta_energy_one(1, [0.5, 0.0, 0.0]) == math.inf
    # from 0.9 to 2.0 (free): -E(0.9)
    assert s.delta_energy_one(1, [2.0, 0.0, 0.0]) == 0.0 - (0.5 * 1e4 * (0.9 - 1.0) ** 2 + 100.0)
    assert abs(s.total_energy() - e1) < 1e-9
    # a body that starts beyond d_max: any finite final state is dH = -inf
    deep = np.array([[0.0, 0.0, 0.0], [0.5, 0.0, 0.0]])
    s2 = O.OracleSystem(L, deep, model="overlap_penalty", params=p)
    assert s2.total_energy() == math.inf
    assert s2.delta_energy_one(1, [2.0, 0.0, 0.0]) == -math.inf


@pytest.mark.parametrize("dim,n,Lx,mind,seed", [(3, 500, 9.0, 0.85, 1), (2, 600, 28.0, 0.85, 2)])
def test_total_energy_matches_all_pairs(dim, n, Lx, mind, seed):
    p = (1e4, 0.2, 100.0, 1.0)
    L = np.array([Lx] * dim)
    pos = wl.random_dilute(n, L, dim, mind, seed)
    d = pos[:, None, :] - pos[None, :, :]
    d -= L * np.round(d / L)
    r = np.sqrt((d * d).sum(-1))
    iu = np.triu_indices(n, 1)
    r = r[iu]
    x = r[r < 1.0] - 1.0
    assert x.size > 0 and (x >= -0.2).all()
    want = float(np.sum(0.5 * 1e4 * x * x + 100.0))
    got = O.OracleSystem(L, pos, model="overlap_penalty", params=p).total_energy()
    assert abs(got - want) <= 1e-12 * want


def test_parallel_sweep_relaxes_overlaps():
    # what QuickInsert relies on: sweeps under the penalty lower the energy
    p = (1e4, 0.2, 100.0, 1.0)
    L = [9.0, 9.0, 9.0]
    pos = wl.random_dilute(400, L, 3, 0.85, 5)
    s = O.OracleSystem(L, pos, model="overlap_penalty", params=p, seed=3)
    e0 = s.total_energy()
    for _ in range(30):
        s.parallel_apply(1.0, 0.1)
        s.increment_step()
    assert s.total_energy() < 0.2 * e0


def test_delta_energy_insert_known_answers():
    # pairwise_cutoff.rs:990-1012 (hard circle, diameter 1, a body at the
    # origin): inserting at (0.4, 0) is infinite, at (1.5, 0) zero
    h = O.OracleSystem([8.0, 8.0], np.zeros((1, 2)))
    assert h.delta_energy_insert([0.4, 0.0]) == math.inf
    assert h.delta_energy_insert([1.5, 0.0]) == 0.0
    # the overlap penalty: one penalised pair, then beyond the maximum overlap
    s = O.OracleSystem([8.0] * 3, np.array([[0.0, 0.0, 0.0], [3.0, 0.0, 0.0]]), model="overlap_penalty", params=P)
    x = 0.875 - 1.0
    assert s.delta_energy_insert([0.875, 0.0, 0.0]) == 0.5 * 1e4 * (x * x) + 100.0
    assert s.delta_energy_insert([0.5, 0.0, 0.0]) == math.inf
    assert s.delta_energy_insert([1.5, 0.0, 0.0]) == 0.0
