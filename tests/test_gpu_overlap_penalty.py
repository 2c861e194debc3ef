"""GPU parity for Metropolis sweeps under the overlap penalty
(PairwiseCutoff<Isotropic<Expanded<OverlapPenalty>>> with r_cut = delta,
the Hamiltonian the reference's placement runs its sweeps with) through the
C ABI, against the oracle.

Bar, as for Lennard-Jones: per-apply counts, sorted per-trial traces and
final FP64 positions bit for bit, no ambiguous decision, and total energies
within the summation-order bound.  The configurations hold penalised
overlaps, and some pairs beyond the maximum allowed overlap (infinite
initial energy: every finite trial is accepted, dH = -inf), so that both
infinite branches of delta_energy_one (pairwise_cutoff.rs:919-941) run.
"""
import ctypes as C
import math
import sys

import numpy as np
import pytest

import _oracle as O
import _pkg

pytestmark = pytest.mark.gpu

DEFAULT = (1e4, 0.2, 100.0, 1.0)  # OverlapPenalty::default(), delta 1
SOFT = (50.0, 0.3, 2.0, 1.5)      # finite dH of order kT: Metropolis draws decide


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


def _replay(hmc, L, pos, params, kT, d, n_applies, seed=42):
    ref = O.OracleSystem(L, pos, model="overlap_penalty", params=params, seed=seed)
    b, s = ref.positions()
    gpu = hmc.GpuParallelSweep(L, b, sites=s, model="overlap_penalty", params=params)
    e_gpu, bound = gpu.total_energy()
    e_ref = ref.total_energy()
    assert (e_gpu == e_ref == math.inf) or abs(e_gpu - e_ref) <= bound, (e_gpu, e_ref, bound)
    cap = 64 * len(pos) + 1024
    gpu.trace(True, capacity=cap)
    ref.trace(True)
    total_acc = 0
    for k in range(n_applies):
        step, sub = ref.counter()
        ga, gr, used = gpu.apply(step, sub, seed, kT, d)
        oa, orj = ref.parallel_apply(kT, d)
        _, sub2 = ref.counter()
        assert (ga, gr, used) == (oa, orj, sub2 - sub), (k, ga, gr, used, oa, orj, sub2 - sub)
        total_acc += ga
        gt = gpu.read_trace(capacity=cap)
        ot = ref.read_trace()
        gt = gt[gt["status"] != 2]
        ot = ot[ot["status"] != 2]
        key = ["substep", "color", "space", "body", "status"]
        gt = gt[np.lexsort((gt["space"], gt["color"], gt["substep"]))][key]
        ot = ot[np.lexsort((ot["space"], ot["color"], ot["substep"]))][key]
        assert np.array_equal(gt, ot), k
        ref.increment_step()
    gb, gs = gpu.positions()
    ob, os_ = ref.positions()
    assert np.array_equal(gb.view(np.uint64), ob.view(np.uint64))
    assert np.array_equal(gs.view(np.uint64), os_.view(np.uint64))
    e_gpu, bound = gpu.total_energy()
    e_ref = ref.total_energy()
    assert (e_gpu == e_ref == math.inf) or abs(e_gpu - e_ref) <= bound, (e_gpu, e_ref, bound)
    st = gpu.stats()
    gpu.close()
    assert st["ambiguous"] == 0, st
    return total_acc, e_ref


@pytest.mark.parametrize("dim,n,Lx,mind,params,seed", [
    (3, 1500, 13.0, 0.85, DEFAULT, 1),
    (2, 1200, 40.0, 0.85, DEFAULT, 2),
    (3, 1200, 16.0, 1.25, SOFT, 3),
    (2, 1000, 56.0, 1.25, SOFT, 4),
])
def test_overlap_penalty_matches_oracle(hmc, wl, dim, n, Lx, mind, params, seed):
    L = [Lx] * dim
    pos = wl.random_dilute(n, L, dim, mind, seed)
    log(f"overlap penalty: D={dim} N={n} params={params}")
    acc, e = _replay(hmc, L, pos, params, 1.0, 0.1, 12)
    assert acc > 0 and math.isfinite(e)


def test_infinite_initial_energy_matches_oracle(hmc, wl):
    # pairs closer than delta - d_max: the bodies in them accept any finite
    # trial (dH = -inf) and the total energy is infinite until they separate
    L = [12.0, 12.0, 12.0]
    pos = wl.random_dilute(900, L, 3, 0.7, 5)
    d = pos[:, None, :] - pos[None, :, :]
    d -= np.asarray(L) * np.round(d / np.asarray(L))
    r = np.sqrt((d * d).sum(-1))
    np.fill_diagonal(r, np.inf)
    assert (r < 0.8).any()
    log("overlap penalty: infinite initial energy, N=900")
    _replay(hmc, L, pos, DEFAULT, 1.0, 0.15, 10)


def test_overlap_penalty_rejects_mixed_precision(hmc, wl):
    L = [12.0, 12.0, 12.0]
    pos = wl.random_dilute(100, L, 3, 1.0, 6)
    with pytest.raises(hmc.HmcError) as e:
        hmc.GpuParallelSweep(L, pos, model="overlap_penalty", precision="mixed")
    assert e.value.code == hmc.HMC_ERR_UNSUPPORTED
