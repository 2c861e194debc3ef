"""The device's restatements of glibc's log, exp, sincos and cbrt against the
host C library (libm.so.6, the library Rust's std calls on Linux), bit for
bit, on 10^8 arguments per function drawn from the ranges the sweep uses:

* log: Open01 draws (m + 1/2) 2^-52 in (0, 1), the ziggurat tail's inputs
  (rand_distr 0.6 StandardNormal zero_case);
* exp: -x^2/2 for x below the ziggurat's base layer (the wedge test), and
  -dH/kT down to the underflow range (Metropolis, parallel_sweep.rs:267-273);
* sincos: half the rotation angle of VersorDisplacement (versor.rs:58-61),
  both as its sampler produces it and uniformly over [0, pi/2];
* cbrt: the 3D ball sampler's cbrt(u), u = k 2^-53 (distribution.rs:63-66).
"""
import sys

import numpy as np
import pytest

import _oracle as O
import _pkg

pytestmark = pytest.mark.gpu
CHUNK = 1 << 23
TOTAL = 100_000_000


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hmc():
    p = _pkg.pkg()
    p.lib()
    return p


def _check(hmc, kind, draw, total=TOTAL, seed=0):
    rng = np.random.default_rng(seed)
    done = 0
    while done < total:
        x = draw(rng, min(CHUNK, total - done))
        got = hmc.debug_math(kind, x)
        want = O.libm_eval(kind, x)
        bad = np.flatnonzero(got.view(np.uint64) != want.view(np.uint64))
        assert bad.size == 0, (kind, x[bad[:4]].tolist(), got[bad[:4]].tolist(), want[bad[:4]].tolist())
        done += x.size
    log(f"{kind}: {done} arguments bit-identical to the host library")


def test_log_on_open01_draws(hmc):
    def draw(rng, n):
        m = rng.integers(0, 1 << 52, size=n, dtype=np.uint64)
        return (m.astype(np.float64) + 0.5) * 2.0**-52
    _check(hmc, "log", draw)
    edge = np.array([2.0**-53, 0.5 * 2.0**-52, 1.0 - 2.0**-53, 0.9375, 0.9375 - 2.0**-53, 1.0, 0.5, 0.25,
                     1.0 + 2.0**-52, 1.0647, 1.375, 0.6875, 1e-300, 5e-324, 2.0**-1022, 1e300, 2.0])
    assert np.array_equal(hmc.debug_math("log", edge).view(np.uint64), O.libm_eval("log", edge).view(np.uint64))


def test_exp_wedge_and_metropolis_ranges(hmc):
    x0 = 3.910757959537090045  # the ziggurat's widest layer

    def wedge(rng, n):
        x = rng.uniform(0.0, x0, size=n)
        return -x * x / 2.0

    def metropolis(rng, n):
        # -dH / kT over many scales, down past the underflow to zero
        return -np.exp(rng.uniform(np.log(1e-12), np.log(760.0), size=n))
    _check(hmc, "exp", wedge, seed=1)
    _check(hmc, "exp", metropolis, total=20_000_000, seed=2)
    edge = np.array([0.0, -0.0, -1e-300, -2.0**-60, -745.1332191019411, -745.1332191019412, -745.2,
                     -708.3964185322641, -708.4, -1e10, -np.inf, 2.0**-60, 1.0, 700.0])
    assert np.array_equal(hmc.debug_math("exp", edge).view(np.uint64), O.libm_eval("exp", edge).view(np.uint64))


@pytest.mark.parametrize("fn", ["sin", "cos"])
def test_sincos_of_half_rotation_angles(hmc, fn):
    def sampler(rng, n):
        # 0.5 |s| for s ~ Normal(0, sigma)^3, sigma over the move sizes in use
        sigma = rng.choice([0.05, 0.1, 0.3, 1.0], size=(n, 1))
        th = np.linalg.norm(rng.normal(size=(n, 3)) * sigma, axis=1)
        return 0.5 * th[th <= np.pi]

    def uniform(rng, n):
        return rng.uniform(0.0, np.pi / 2, size=n)
    _check(hmc, fn, sampler, total=TOTAL // 2, seed=3)
    _check(hmc, fn, uniform, total=TOTAL // 2, seed=4)
    edge = np.array([0.0, 2.0**-30, 2.0**-27, 2.0**-26, 0.126, 0.125999, 0.85546875, 0.8554687, 1.0,
                     np.pi / 2, 2.4262])
    assert np.array_equal(hmc.debug_math(fn, edge).view(np.uint64), O.libm_eval(fn, edge).view(np.uint64))


def test_cbrt_of_ball_radii(hmc):
    def draw(rng, n):
        return rng.integers(0, 1 << 53, size=n, dtype=np.uint64).astype(np.float64) * 2.0**-53
    _check(hmc, "cbrt", draw, total=20_000_000, seed=5)
