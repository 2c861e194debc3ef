"""Statistical checks of the oracle's sweeps (CPU).

* The reference's harmonic-oscillator test (hoomd-mc/src/sweep.rs:420-461):
  one point body in External(Harmonic) with K = 2 about (1, -2), Translate
  with maximum distance 0.1, 200,000 sweeps; the mean position is the
  centre and the mean energy is kT (two quadratic degrees of freedom), each
  within the reference's tolerance 0.3 -- and, with the statistics in hand,
  within five block-averaged standard errors.
* Serial Sweep and ParallelSweep sample the same distribution: the mean LJ
  energy per body of a small system agrees between them within the
  block-averaged error bars.
"""
import numpy as np
import pytest

import _oracle as O
import _pkg


@pytest.fixture(scope="module")
def wl():
    return _pkg.workloads()


def block_stats(x, nblocks=20):
    """Mean and standard error of the mean from nblocks block averages."""
    x = np.asarray(x, np.float64)
    m = len(x) // nblocks
    b = x[: m * nblocks].reshape(nblocks, m).mean(axis=1)
    return float(x.mean()), float(b.std(ddof=1) / np.sqrt(nblocks))


@pytest.mark.parametrize("parallel", [False, True], ids=["sweep", "parallel_sweep"])
@pytest.mark.parametrize("kt", [1.0, 2.5])
def test_harmonic_oscillator(kt, parallel):
    K = 2.0
    origin = np.array([1.0, -2.0])
    # a periodic box far larger than the excursions (the reference uses a
    # closed boundary; no body comes within reach of the edges here)
    L = np.array([20.0, 20.0])
    ref = O.OracleSystem(L, origin[None, :], seed=7)
    ref.set_external_harmonic(K, origin)
    n_steps, nblocks = 200_000, 20
    m = n_steps // nblocks
    pos_blocks, e_blocks = [], []
    for _ in range(nblocks):
        ps, es = ref.run_accumulate(parallel, kt, 0.1, m)
        pos_blocks.append(ps / m)
        e_blocks.append(es / m)
    pos_blocks = np.array(pos_blocks)
    e_blocks = np.array(e_blocks)
    pos = pos_blocks.mean(axis=0)
    e = e_blocks.mean()
    se_pos = pos_blocks.std(axis=0, ddof=1) / np.sqrt(nblocks)
    se_e = e_blocks.std(ddof=1) / np.sqrt(nblocks)
    # the reference's bar
    assert abs(pos[0] - origin[0]) <= 0.3 and abs(pos[1] - origin[1]) <= 0.3, pos
    assert abs(e - kt) <= 0.3, e
    # and the tighter one its statistics allow
    assert (np.abs(pos - origin) <= 5.0 * se_pos + 1e-3).all(), (pos, se_pos)
    assert abs(e - kt) <= 5.0 * se_e + 1e-3, (e, se_e)


def test_serial_and_parallel_sweep_agree_on_lj_energy(wl):
    # 864 shifted-LJ bodies (FCC 6^3 at number density 0.8 -- a liquid at
    # kT = 1.2 after melting), r_cut 2.5: 300 applies of equilibration, then
    # 1,200 samples of U / N, one per apply (an apply of either sweep is N
    # trials)
    L, pos = wl.fcc_number_density((6, 6, 6), 0.8)
    prm = (1.0, 1.0, 2.5, 2.5)
    means = {}
    for parallel in (False, True):
        ref = O.OracleSystem(L, pos, model="lj_shifted", params=prm, seed=11)
        u = []
        for k in range(1500):
            if parallel:
                ref.parallel_apply(1.2, 0.1)
            else:
                ref.sweep_apply(1.2, 0.1)
            ref.increment_step()
            if k >= 300:
                u.append(ref.total_energy() / len(pos))
        means[parallel] = block_stats(u)
    (ms, ss), (mp, sp) = means[False], means[True]
    assert abs(ms - mp) <= 4.0 * np.hypot(ss, sp), (means, abs(ms - mp) / np.hypot(ss, sp))
    # a liquid's energy, far from the FCC lattice's
    assert -6.5 < ms < -3.5, ms
