"""Scale distribution function (SDF): the oracle's histogram (oracle/sdf.hpp)
pinned by its all-pairs neighbour search, closed forms (two bodies, a perfect
FCC lattice, a square lattice of disks), overlap counting, and the hard-sphere equation of state
(Carnahan-Starling) through the package's beta P / rho extrapolation on an
oracle trajectory.  The reference has no pressure compute; the SDF follows
HOOMD-blue's hpmc compute.SDF definition."""
import math

import numpy as np
import pytest

import _oracle as O
import _pkg

wl = _pkg.workloads()

XMAX, DX = 0.02, 1e-4
NBINS = int(math.ceil(XMAX / DX))


def _bin_of(r, sigma=1.0):
    return int(math.floor((1.0 - sigma / r) / DX))


@pytest.mark.parametrize("dim,n,L,mind,seed", [
    (3, 600, [11.0, 12.0, 13.0], 1.0, 1),
    (3, 400, [9.0, 9.0, 9.0], 0.97, 2),      # with overlaps
    (3, 300, [5.5, 40.0, 6.0], 1.0, 3),      # an axis with fewer than 3 cells of r_max: all pairs
    (2, 800, [40.0, 44.0], 1.0, 4),
    (2, 500, [30.0, 30.0], 0.95, 5),
])
def test_cell_search_matches_all_pairs(dim, n, L, mind, seed):
    pos = wl.random_dilute(n, L, dim, mind, seed)
    h_cells, ov_cells = O.sdf(L, pos, 1.0, XMAX, DX, NBINS)
    h_all, ov_all = O.sdf(L, pos, 1.0, XMAX, DX, NBINS, all_pairs=True)
    assert np.array_equal(h_cells, h_all) and ov_cells == ov_all
    if mind >= 1.0:
        assert ov_all == 0


def test_all_pairs_matches_numpy_brute_force():
    # minimum over j of the minimum-image distance, written independently
    L = np.array([8.0, 9.0, 10.0])
    pos = wl.random_dilute(300, L, 3, 1.0, 11)
    d = pos[:, None, :] - pos[None, :, :]
    d -= L * np.round(d / L)
    r2 = (d * d).sum(-1)
    np.fill_diagonal(r2, np.inf)
    x = 1.0 - 1.0 / np.sqrt(r2.min(1))
    want = np.zeros(NBINS, np.uint64)
    for xi in x[x < XMAX]:
        want[int(math.floor(xi / DX))] += 1
    got, ov = O.sdf(L, pos, 1.0, XMAX, DX, NBINS, all_pairs=True)
    # the two may differ only for x within rounding of a bin edge (none here)
    edge = np.abs(x / DX - np.round(x / DX))
    assert edge[x < XMAX].min() > 1e-6
    assert np.array_equal(got, want) and ov == 0


@pytest.mark.parametrize("dim", [2, 3])
def test_two_bodies_closed_form(dim):
    # x = 1 - sigma / r for the only pair, inside and across the boundary
    r = 1.0 / (1.0 - 0.00505)  # x = 0.00505: the middle of bin 50
    L = [10.0] * dim
    a = np.zeros((2, dim))
    a[1, 0] = r
    b = np.zeros((2, dim))
    b[0, 0] = -5.0 + 0.25
    b[1, 0] = 5.0 - (r - 0.25)
    for pos in (a, b):
        h, ov = O.sdf(L, pos, 1.0, XMAX, DX, NBINS)
        assert ov == 0 and h.sum() == 2 and h[50] == 2 == h[_bin_of(r)]
    # sigma scales x: the same pair with sigma 0.995 sits at x = 1 - 0.995 (1 - 0.00505)
    h, _ = O.sdf(L, a, 0.995, XMAX, DX, NBINS)
    assert h[_bin_of(r, 0.995)] == 2
    # beyond x_max nothing is counted; overlapping pairs are counted apart
    far = a.copy()
    far[1, 0] = 1.03
    assert O.sdf(L, far, 1.0, XMAX, DX, NBINS)[0].sum() == 0
    near = a.copy()
    near[1, 0] = 0.99
    h, ov = O.sdf(L, near, 1.0, XMAX, DX, NBINS)
    assert ov == 2 and h.sum() == 0


def test_fcc_lattice_closed_form():
    # every body of a perfect FCC lattice has its nearest neighbours at a / sqrt(2)
    for phi in (0.70, 0.71, 0.72):
        L, pos = wl.fcc_spheres(n_cells=(4, 4, 4), packing_fraction=phi)
        a = L[0] / 4
        x = 1.0 - math.sqrt(2.0) / a
        k = x / DX
        if abs(k - round(k)) < 1e-3:
            continue  # too close to a bin edge for this check
        h, ov = O.sdf(L, pos, 1.0, XMAX, DX, NBINS)
        assert ov == 0 and h.sum() == len(pos) and h[int(math.floor(k))] == len(pos)
        return
    pytest.fail("no packing fraction away from a bin edge")


def test_square_lattice_closed_form():
    # every disk of a square lattice has its nearest neighbours at the
    # lattice constant a, so x = 1 - sigma / a for all of them
    for phi in (0.77, 0.76, 0.75):
        L, pos = wl.square_lattice_disks(n_side=12, packing_fraction=phi)
        a = L[0] / 12
        k = (1.0 - 1.0 / a) / DX
        if not 0.0 <= k < NBINS or abs(k - round(k)) < 1e-3:
            continue
        h, ov = O.sdf(L, pos, 1.0, XMAX, DX, NBINS)
        assert ov == 0 and h.sum() == len(pos) and h[int(math.floor(k))] == len(pos)
        return
    pytest.fail("no packing fraction inside the histogram and away from a bin edge")


def test_empty_and_single_body():
    h, ov = O.sdf([10.0, 10.0, 10.0], np.zeros((0, 3)), 1.0, XMAX, DX, NBINS)
    assert h.sum() == 0 and ov == 0
    h, ov = O.sdf([10.0, 10.0, 10.0], np.zeros((1, 3)), 1.0, XMAX, DX, NBINS)
    assert h.sum() == 0 and ov == 0


def carnahan_starling(phi):
    return (1 + phi + phi * phi - phi**3) / (1 - phi) ** 3


def test_hard_sphere_pressure_matches_carnahan_starling():
    # fluid at packing fraction 0.30 (Z_CS = 3.974), melted from FCC by the
    # oracle's ParallelSweep, then SDF counts summed over frames
    phi = 0.30
    L, pos = wl.fcc_spheres(n_cells=(8, 8, 8), packing_fraction=phi)
    s = O.OracleSystem(L, pos, seed=7)
    for _ in range(150):
        s.parallel_apply(1.0, 0.3)
        s.increment_step()
    hist = np.zeros(NBINS, np.uint64)
    frames = 300
    for _ in range(frames):
        for _ in range(2):
            s.parallel_apply(1.0, 0.3)
            s.increment_step()
        _b, sites = s.positions()
        h, ov = O.sdf(L, sites, 1.0, XMAX, DX, NBINS)
        assert ov == 0
        hist += h
    z = _pkg.pkg().sdf_beta_p_over_rho(hist, len(pos), DX, 3, n_frames=frames)
    assert abs(z / carnahan_starling(phi) - 1.0) < 0.04, z
