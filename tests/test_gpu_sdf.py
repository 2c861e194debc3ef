"""GPU parity for the scale distribution function (hmc_sdf, hmc_batch_sdf):
the device histogram equals the oracle's (oracle/sdf.hpp) count for count on
the same site positions, at the hard-sphere bench size, in 2D, for every
replica of a batch, and the pressure it gives matches the Carnahan-Starling
equation of state."""
import sys

import numpy as np
import pytest

import _oracle as O
import _pkg

pytestmark = pytest.mark.gpu

XMAX, DX = 0.02, 1e-4
NBINS = 200


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


def test_sdf_full_size_spheres_match_oracle(hmc, wl):
    L, pos = wl.fcc_spheres()
    log(f"sdf: hard spheres N={len(pos)}")
    sim = hmc.GpuParallelSweep(L, wl.wrap(L, pos), precision="mixed")
    h0, ov0 = sim.sdf(XMAX, DX)
    _b, s = sim.positions()
    want, wov = O.sdf(L, s, 1.0, XMAX, DX, NBINS)
    assert np.array_equal(h0, want) and ov0 == wov == 0
    sim.run(0, 42, 1.0, 0.08, 8)
    h, ov = sim.sdf(XMAX, DX)
    _b, s = sim.positions()
    want, wov = O.sdf(L, s, 1.0, XMAX, DX, NBINS)
    log(f"  counted {int(h.sum())} of {len(pos)} bodies within x < {XMAX}")
    assert np.array_equal(h, want) and ov == wov == 0
    assert h.sum() > len(pos) // 10
    # the SDF leaves the sweep's replay untouched: the next applies continue as before
    a1 = sim.run(8, 42, 1.0, 0.08, 2)
    sim.close()
    ref = hmc.GpuParallelSweep(L, wl.wrap(L, pos), precision="mixed")
    ref.run(0, 42, 1.0, 0.08, 8)
    a2 = ref.run(8, 42, 1.0, 0.08, 2)
    assert a1 == a2
    ref.close()


def test_sdf_disks_match_oracle(hmc, wl):
    L, pos = wl.square_lattice_disks()
    sim = hmc.GpuParallelSweep(L, wl.wrap(L, pos), precision="fp64")
    sim.run(0, 42, 1.0, 0.15, 20)
    h, ov = sim.sdf(XMAX, DX)
    _b, s = sim.positions()
    want, wov = O.sdf(L, s, 1.0, XMAX, DX, NBINS)
    assert np.array_equal(h, want) and ov == wov == 0 and h.sum() > 0


def test_sdf_random_overlapping_configuration(hmc, wl):
    # a state with overlaps (set directly): the overlap counts agree too
    L = [12.0, 12.0, 12.0]
    pos = wl.random_dilute(900, L, 3, 0.97, 3)
    sim = hmc.GpuParallelSweep(L, pos, precision="fp64")
    h, ov = sim.sdf(XMAX, DX)
    want, wov = O.sdf(L, pos, 1.0, XMAX, DX, NBINS)
    assert np.array_equal(h, want) and ov == wov and ov > 0


def test_batch_sdf_matches_oracle_per_replica(hmc, wl):
    systems = wl.batched_replicas(n_replicas=8, k_range=(3, 7), phi_range=(0.30, 0.54))
    batch = hmc.GpuBatch([(L, wl.wrap(L, p), None) for L, p in systems], precision="mixed")
    batch.run(0, [42 + r for r in range(8)], 1.0, 0.08, 10)
    hist, ov = batch.sdf(XMAX, DX)
    for r, (L, _p) in enumerate(systems):
        _b, s = batch.positions(r)
        want, wov = O.sdf(L, s, 1.0, XMAX, DX, NBINS)
        assert np.array_equal(hist[r], want) and int(ov[r]) == wov == 0, r
    batch.close()


def test_sdf_pressure_matches_carnahan_starling(hmc, wl):
    phi = 0.30
    z_cs = (1 + phi + phi * phi - phi**3) / (1 - phi) ** 3
    L, pos = wl.fcc_spheres(n_cells=(10, 10, 10), packing_fraction=phi)
    sim = hmc.GpuParallelSweep(L, wl.wrap(L, pos), precision="mixed")
    sim.run(0, 5, 1.0, 0.3, 200)
    hist = np.zeros(NBINS, np.uint64)
    frames = 300
    for f in range(frames):
        sim.run(200 + 2 * f, 5, 1.0, 0.3, 2)
        h, ov = sim.sdf(XMAX, DX)
        assert ov == 0
        hist += h
    z = hmc.sdf_beta_p_over_rho(hist, len(pos), DX, 3, n_frames=frames)
    log(f"  beta P / rho = {z:.4f} (Carnahan-Starling {z_cs:.4f})")
    assert abs(z / z_cs - 1.0) < 0.03
    sim.close()


def test_sdf_rejects_unsupported_and_invalid(hmc, wl):
    L = [12.0, 12.0, 12.0]
    pos = wl.random_dilute(100, L, 3, 1.0, 4)
    sim = hmc.GpuParallelSweep(L, pos, precision="fp64")
    with pytest.raises(hmc.HmcError) as e:
        sim.sdf(0.6, DX)
    assert e.value.code == hmc.HMC_ERR_INVALID
    with pytest.raises(hmc.HmcError) as e:
        sim.sdf(XMAX, 0.0, nbins=10)
    assert e.value.code == hmc.HMC_ERR_INVALID
    sim.close()
    lj = hmc.GpuParallelSweep(L, pos, model="lj_shifted", precision="fp64")
    with pytest.raises(hmc.HmcError) as e:
        lj.sdf(XMAX, DX)
    assert e.value.code == hmc.HMC_ERR_UNSUPPORTED
    lj.close()
