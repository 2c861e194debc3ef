"""hmc_total_energy_noalloc (the allocation-free total energy) against
hmc_total_energy through the C ABI, on two calls in a row, the second on a
smaller system: for hard spheres (zero, then an overlap: infinite) identical
bit for bit; for an isotropic model (finite sums) within the sum of the two
returned summation-order bounds, since the order of the bodies within a cell
comes from the binning's atomic counters and may differ between any two
calls (the bounds agree to 1e-12 relative)."""
import sys

import numpy as np
import pytest

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


def _same(a, b):
    return np.array_equal(np.array(a).view(np.uint64), np.array(b).view(np.uint64))


def test_noalloc_total_energy_matches_lj(hmc, wl):
    prm = (1.0, 1.0, 2.5, 2.5)
    for cells in ((16, 16, 16), (6, 6, 6)):  # 16,384 bodies, then 864
        L, pos = wl.fcc_number_density(cells, 0.8)
        log(f"noalloc total energy: shifted LJ N={len(pos)}")
        g = hmc.GpuParallelSweep(L, wl.wrap(L, pos), model="lj_shifted", params=prm, precision="mixed")
        g.run(0, 3, 1.2, 0.1, 4)
        want = g.total_energy()
        got = g.total_energy_noalloc()
        assert abs(got[0] - want[0]) <= got[1] + want[1], (got, want)
        assert abs(got[1] - want[1]) <= 1e-12 * want[1]
        assert np.isfinite(got[0]) and got[0] < 0.0 and got[1] > 0.0
        # the scratch keeps the last call's partials: a second call agrees too
        again = g.total_energy_noalloc()
        assert abs(again[0] - want[0]) <= again[1] + want[1] and abs(again[1] - want[1]) <= 1e-12 * want[1]
        g.close()


def test_noalloc_total_energy_matches_hard_spheres(hmc, wl):
    L, pos = wl.fcc_spheres(n_cells=(12, 12, 12), packing_fraction=0.45)
    log(f"noalloc total energy: hard spheres N={len(pos)}")
    g = hmc.GpuParallelSweep(L, wl.wrap(L, pos), precision="mixed")
    g.run(0, 3, 1.0, 0.08, 4)
    want = g.total_energy()
    assert want == (0.0, 0.0)
    assert _same(g.total_energy_noalloc(), want)
    g.close()
    # a smaller system with one overlapping pair: infinite
    L, pos = wl.fcc_spheres(n_cells=(4, 4, 4), packing_fraction=0.45)
    pos = pos.copy()
    pos[1] = pos[0] + np.array([0.5, 0.0, 0.0])
    log(f"noalloc total energy: hard spheres with an overlap N={len(pos)}")
    g = hmc.GpuParallelSweep(L, wl.wrap(L, pos), precision="fp64")
    want = g.total_energy()
    got = g.total_energy_noalloc()
    assert np.isinf(want[0]) and _same(got, want)
    g.close()
