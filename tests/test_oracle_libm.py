"""The oracle's correctly rounded log/exp (45-digit decimal vectors) and sin/cos
(60-digit Taylor vectors), and
the size of the glibc boundary those functions stand in for."""
import ctypes as C
import json
import math
import pathlib

import numpy as np
import pytest

import _oracle as O

GOLDEN = pathlib.Path(__file__).with_name("golden") / "libm_cr_vectors.json"


def _lib():
    L = O.lib()
    L.oracle_cr_log.argtypes = [C.c_double]
    L.oracle_cr_log.restype = C.c_double
    L.oracle_cr_exp.argtypes = [C.c_double]
    L.oracle_cr_exp.restype = C.c_double
    L.oracle_cr_sin.argtypes = [C.c_double]
    L.oracle_cr_sin.restype = C.c_double
    L.oracle_cr_cos.argtypes = [C.c_double]
    L.oracle_cr_cos.restype = C.c_double
    L.oracle_set_libm.argtypes = [C.c_int]
    L.oracle_libm_counters.argtypes = [O.U64_P]
    return L


def test_cr_log_matches_decimal():
    L = _lib()
    d = json.loads(GOLDEN.read_text())
    for xh, wh in d["log"]:
        x, want = float.fromhex(xh), float.fromhex(wh)
        assert L.oracle_cr_log(x) == want, (xh, wh)


def test_cr_exp_matches_decimal():
    L = _lib()
    d = json.loads(GOLDEN.read_text())
    for yh, wh in d["exp"]:
        y, want = float.fromhex(yh), float.fromhex(wh)
        assert L.oracle_cr_exp(y) == want, (yh, wh)


@pytest.mark.parametrize("fn", ["sin", "cos"])
def test_cr_sin_cos_match_taylor_at_60_digits(fn):
    # the rotation sampler's sin/cos of half the angle, on [0, pi/2]
    L = _lib()
    f = getattr(L, f"oracle_cr_{fn}")
    d = json.loads(GOLDEN.read_text())
    for xh, wh in d[fn]:
        x, want = float.fromhex(xh), float.fromhex(wh)
        assert f(x) == want, (fn, xh, wh)


def test_glibc_log_is_not_always_correctly_rounded():
    # The reason the oracle has two libm modes: on these inputs the host
    # library and the correctly rounded value differ now and then.
    L = _lib()
    d = json.loads(GOLDEN.read_text())
    diff = sum(math.log(float.fromhex(x)) != float.fromhex(w) for x, w in d["log"])
    assert diff < 0.01 * len(d["log"])


def test_libm_modes_and_counters():
    L = _lib()
    cnt = np.zeros(4, np.uint64)
    O.lib().oracle_libm_reset()
    z_glibc = O.sample(1, 2, 3, 4, "normal", 200000)
    L.oracle_libm_counters(O._p(cnt, O.U64_P))
    assert cnt[0] > 0 and cnt[2] > 0  # tails and wedges were sampled
    L.oracle_set_libm(1)
    try:
        z_cr = O.sample(1, 2, 3, 4, "normal", 200000)
    finally:
        L.oracle_set_libm(0)
    # the two modes differ at most where a divergence was counted
    assert np.sum(z_glibc != z_cr) <= int(cnt[1] + cnt[3]) // 2 + 1
