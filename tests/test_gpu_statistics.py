"""Statistical oracles for the GPU sweep (equilibrium averages, not replay).

* LJ: the mean energy per body of the GPU ParallelSweep equals the oracle's
  serial Sweep (hoomd-mc/src/sweep.rs:159-224) -- a different algorithm with
  the same equilibrium distribution -- within block-averaged error bars.
* Hard spheres: beta P / rho from the SDF of GPU trajectories at three fluid
  packing fractions agrees with the Carnahan-Starling equation of state and
  with the SDF of an oracle Sweep trajectory of the same system, within the
  error bars (plus 1.5% for the equation of state, whose own accuracy and
  the SDF's finite-x extrapolation are of that order).
"""
import sys
import time

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


def block_stats(x, nblocks=20):
    x = np.asarray(x, np.float64)
    m = len(x) // nblocks
    b = x[: m * nblocks].reshape(nblocks, m).mean(axis=1)
    return float(x.mean()), float(b.std(ddof=1) / np.sqrt(nblocks))


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_lj_mean_energy_matches_serial_sweep(hmc, wl, precision):
    # 864 shifted-LJ bodies, FCC 6^3 at number density 0.8, kT 1.2, d 0.1:
    # 300 applies of equilibration, then 1,200 samples of U / N (one apply
    # of either sweep is N trials)
    L, pos = wl.fcc_number_density((6, 6, 6), 0.8)
    prm = (1.0, 1.0, 2.5, 2.5)
    t = time.time()
    ref = O.OracleSystem(L, pos, model="lj_shifted", params=prm, seed=11)
    b, s = ref.positions()
    gpu = hmc.GpuParallelSweep(L, b, sites=s, model="lj_shifted", params=prm, precision=precision)
    ug, uo = [], []
    for k in range(1500):
        gpu.apply(k, 0, 29, 1.2, 0.1)
        ref.sweep_apply(1.2, 0.1)
        ref.increment_step()
        if k >= 300:
            e, bound = gpu.total_energy()
            ug.append(e / len(pos))
            uo.append(ref.total_energy() / len(pos))
    gpu.close()
    (mg, sg), (mo, so) = block_stats(ug), block_stats(uo)
    z = abs(mg - mo) / np.hypot(sg, so)
    log(f"LJ N=864 {precision}: GPU ParallelSweep U/N = {mg:.4f} +- {sg:.4f}, oracle Sweep {mo:.4f} +- {so:.4f} "
        f"({z:.2f} sigma, {time.time() - t:.1f}s)")
    assert z <= 4.0
    assert -6.5 < mg < -3.5


@pytest.mark.parametrize("phi", [0.20, 0.30, 0.40])
def test_sdf_pressure_three_packing_fractions(hmc, wl, phi):
    z_cs = (1 + phi + phi * phi - phi**3) / (1 - phi) ** 3
    n_frames = 400
    # GPU: N = 4,000 spheres melted from FCC, 400 frames 2 applies apart
    L, pos = wl.fcc_spheres(n_cells=(10, 10, 10), packing_fraction=phi)
    d = 0.3 if phi < 0.35 else 0.15
    sim = hmc.GpuParallelSweep(L, wl.wrap(L, pos), precision="mixed")
    sim.run(0, 5, 1.0, d, 400)
    zg = []
    for f in range(n_frames):
        sim.run(400 + 2 * f, 5, 1.0, d, 2)
        h, ov = sim.sdf(XMAX, DX)
        assert ov == 0
        zg.append(hmc.sdf_beta_p_over_rho(h, len(pos), DX, 3))
    sim.close()
    # oracle: serial Sweep of N = 500 spheres (FCC 5^3), same sampling
    L2, pos2 = wl.fcc_spheres(n_cells=(5, 5, 5), packing_fraction=phi)
    ref = O.OracleSystem(L2, pos2, seed=9)
    for _ in range(400):
        ref.sweep_apply(1.0, d)
        ref.increment_step()
    zo = []
    for f in range(n_frames):
        for _ in range(2):
            ref.sweep_apply(1.0, d)
            ref.increment_step()
        _b, s2 = ref.positions()
        h2, ov2 = O.sdf(L2, s2, 1.0, XMAX, DX, NBINS)
        assert ov2 == 0
        zo.append(hmc.sdf_beta_p_over_rho(h2, len(pos2), DX, 3))
    (mg, sg), (mo, so) = block_stats(zg), block_stats(zo)
    log(f"phi {phi}: beta P / rho GPU {mg:.4f} +- {sg:.4f}, oracle Sweep {mo:.4f} +- {so:.4f}, "
        f"Carnahan-Starling {z_cs:.4f}")
    assert abs(mg - mo) <= 4.0 * np.hypot(sg, so)
    assert abs(mg - z_cs) <= 4.0 * sg + 0.015 * z_cs
    assert abs(mo - z_cs) <= 4.0 * so + 0.015 * z_cs
