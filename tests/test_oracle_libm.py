"""The glibc data tables the device restates (glibc_tables.inc) and the
oracle's C library calls.

* exp and sincos tables: regenerated here from their definitions
  (2^(k/128) = H (1 + T); sin and cos of k/128 as hi + lo) and compared
  entry for entry.
* log table: every (invc, logc) pair satisfies the constraints of glibc's
  construction -- logc a multiple of 2^-43 within 2^-66 of log(1/invc), and
  1/invc inside its subinterval of [0x1.6p-1, 0x1.6p0) -- checked with
  60-digit decimal logs.  Which candidate glibc chose among the ones that
  satisfy them is pinned by the results: the device's log equals the host
  library's on 10^8 arguments (test_gpu_libm.py).
* the oracle's sincos is the host library's sincos, and differs from the
  host's separate sin / cos (FMA builds) on some arguments: the rotation
  sampler's choice of call is visible in its results.
"""
import importlib.util
import pathlib
import re
from decimal import Decimal, getcontext

import numpy as np

import _oracle as O

CSRC = pathlib.Path(__file__).resolve().parent.parent / "hoomd-rs_amd" / "csrc"


def _gen():
    spec = importlib.util.spec_from_file_location("gen_glibc_tables", CSRC / "gen_glibc_tables.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _table(name):
    text = (CSRC / "glibc_tables.inc").read_text()
    body = text[text.index(name):]
    body = body[body.index("{") + 1: body.index("};")]
    rows = re.findall(r"\{([^{}]*)\}", body)
    return [[v.strip() for v in r.split(",") if v.strip()] for r in rows]


def test_exp_table_regenerates():
    want = _gen().exp_table()
    got = [(int(a.rstrip("ul"), 16), int(b.rstrip("ul"), 16)) for a, b in _table("kGlibcExpTab")]
    assert got == want
    # e_exp_data.c's first entries, as the library encodes them
    assert got[1] == (0x3c9b3b4f1a88bf6e, 0x3feff63da9fb3335)


def test_sincos_table_regenerates():
    want = _gen().sincos_table()
    got = [tuple(float.fromhex(v) for v in r) for r in _table("kGlibcSinCosTab")]
    assert len(got) == 111
    assert got == want


def test_log_table_satisfies_its_construction():
    getcontext().prec = 60
    rows = [tuple(float.fromhex(v) for v in r) for r in _table("kGlibcLogTab")]
    assert len(rows) == 128
    off = 0x3FE6000000000000
    for i, (invc, logc) in enumerate(rows):
        lo = np.array([off + (i << 45)], np.uint64).view(np.float64)[0]
        hi = np.array([off + ((i + 1) << 45)], np.uint64).view(np.float64)[0]
        c = 1 / Decimal(invc)
        assert Decimal(lo) <= c < Decimal(hi), i
        assert (Decimal(logc) * Decimal(2) ** 43) % 1 == 0, i  # on the 2^-43 grid
        assert abs(-Decimal(invc).ln() - Decimal(logc)) < Decimal(2) ** -66, i
    # e_log_data.c's first two entries
    assert rows[0] == (float.fromhex("0x1.734f0c3e0de9fp+0"), float.fromhex("-0x1.7cc7f79e69000p-2"))
    assert rows[1] == (float.fromhex("0x1.713786a2ce91fp+0"), float.fromhex("-0x1.76feec20d0000p-2"))


def test_oracle_sincos_is_the_library_sincos():
    import ctypes as C
    libm = C.CDLL("libm.so.6")
    libm.sincos.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    libm.sin.argtypes = libm.cos.argtypes = [C.c_double]
    libm.sin.restype = libm.cos.restype = C.c_double
    x = np.random.default_rng(1).uniform(0.0, np.pi / 2, 20000)
    s, c = O.libm_eval("sin", x), O.libm_eval("cos", x)
    differ = 0
    for k in range(len(x)):
        ss, cc = C.c_double(), C.c_double()
        libm.sincos(float(x[k]), C.byref(ss), C.byref(cc))
        assert (ss.value, cc.value) == (s[k], c[k])
        differ += (libm.sin(float(x[k])), libm.cos(float(x[k]))) != (s[k], c[k])
    # the separate functions are other builds: they round otherwise now and then
    assert differ > 0


def test_normal_sampler_reaches_tail_and_wedge():
    # the ziggurat's log (tail) and exp (wedge) are both exercised by the
    # streams the replay tests use
    z = O.sample(1, 2, 3, 4, "normal", 400000)
    R = 3.6541528853610088
    assert np.sum(np.abs(z) > R) > 0
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01
