"""GPU parity for QuickCompress (hmc_quick_compress_apply over a device
context: hoomd-mc/src/quick_compress.rs:19-444) against the oracle's
restatement, in the placement loop of benchmarks/src/place.rs:31-118 (one
QuickCompress call every k steps, a translate sweep every step, here the
parallel sweep).  At every call the state, the target flag and the box edges
agree; every sweep's counts agree; the final bodies and sites agree bit for
bit, and no energy comparison fell inside the summation-order bound."""
import ctypes as C
import sys

import numpy as np
import pytest

import _oracle as O
import _pkg

pytestmark = pytest.mark.gpu

P = (1e4, 0.2, 100.0, 1.0)


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


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


@pytest.mark.parametrize("dim,n,rho0,rho_t,model,params,every,d,max_steps,precision", [
    (3, 400, 0.42, 0.6, "overlap_penalty", P, 20, 0.1, 3000, "fp64"),
    (2, 300, 0.30, 0.45, "hard_sphere", (1.0,), 5, 0.15, 600, "fp64"),
    (2, 300, 0.30, 0.45, "hard_sphere", (1.0,), 5, 0.15, 600, "mixed"),  # persistent apply
    (3, 500, 0.30, 0.25, "hard_sphere", (1.0,), 10, 0.1, 400, "mixed"),  # expansion
    (3, 500, 0.55, 0.70, "lj_shifted", (1.0, 1.0, 2.5, 2.5), 10, 0.1, 300, "mixed"),  # negative energies
])
def test_quick_compress_matches_oracle(hmc, wl, dim, n, rho0, rho_t, model, params, every, d, max_steps,
                                       precision):
    log(f"QuickCompress: D={dim} N={n} {rho0} -> {rho_t} {model} {precision}")
    L = [(n / rho0) ** (1.0 / dim)] * dim
    vt = n / rho_t
    seed = 5
    ref = O.OracleSystem(L, wl.random_dilute(n, L, dim, 1.0, 3), model=model, params=params, seed=seed)
    b, s = ref.positions()
    gpu = hmc.GpuParallelSweep(L, b, sites=s, model=model, params=params, precision=precision)
    oqc = O.OracleQuickCompress(vt)
    gqc = hmc.QuickCompress(vt)
    compressions = 0
    for _ in range(max_steps):
        step, sub = ref.counter()
        if step % every == 0:
            gqc.apply(gpu, step, sub, seed)
            ref, L2 = oqc.apply(ref, L, seed, model, params)
            compressions += L2 != L
            L = L2
            assert gqc.state == oqc.state and bool(gqc.st.is_final) == oqc.is_final, step
            assert gpu.box_edges == L, step
            if oqc.state == "complete":
                break
        step, sub = ref.counter()
        ga, gr, used = gpu.apply(step, sub, seed, 1.0, d)
        oa, orj = ref.parallel_apply(1.0, d)
        assert (ga, gr) == (oa, orj), step
        ref.increment_step()
    log(f"  {compressions} box changes, state {oqc.state}, density {n / np.prod(L):.4f}, "
        f"layout {gpu.stats()['tiled']}")
    assert compressions > 0 and gqc.ambiguous == 0
    gb, gs = gpu.positions()
    rb, rs = ref.positions()
    assert np.array_equal(bits(gb), bits(rb)) and np.array_equal(bits(gs), bits(rs))
    if model == "overlap_penalty":
        assert oqc.state == "complete"
    gpu.close()
