"""Workload generators: the stored-position convention (wrap) against the
oracle's Periodic::wrap, and the sizes of the configurations."""
import numpy as np

import _oracle as O
import _pkg


def test_wrap_matches_oracle_bit_for_bit():
    wl = _pkg.workloads()
    rng = np.random.default_rng(7)
    for L in (np.array([3.7, 5.0, 10.25]), np.array([101.55549310303758] * 3), np.array([135.58, 135.58])):
        d = len(L)
        x = np.concatenate([rng.uniform(-3 * L.max(), 3 * L.max(), (5000, d)),
                            (rng.integers(-8, 8, (500, d)) + 0.5) * L,  # exact halves: -1/2, never +1/2
                            rng.integers(-8, 8, (500, d)) * L, -0.0 * np.ones((1, d))])
        a, b = wl.wrap(L, x), O.wrap(L, x)
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_configuration_sizes():
    wl = _pkg.workloads()
    assert len(wl.fcc_spheres()[1]) == 1_000_188
    assert len(wl.square_lattice_disks()[1]) == 16_384
    L, pos = wl.simple_cubic()
    assert len(pos) == 262_144 and abs(L[0] - 78.11) < 0.01
    assert np.array_equal(wl.UNIT_CUBE, O.CUBE)
