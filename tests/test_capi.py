"""CPU checks of the product library's C ABI: it loads, exports every entry
point include/hmc.h declares, and rejects unsupported or invalid
configurations before touching a device."""
import ctypes as C
import pathlib
import re

import pytest

import _pkg as _pkg_mod

ROOT = pathlib.Path(__file__).resolve().parent.parent


def _pkg():
    return _pkg_mod.pkg()


def declared_symbols():
    text = (ROOT / "include" / "hmc.h").read_text()
    return sorted(set(re.findall(r"\b(hmc_[a-z_]+)\s*\(", text)))


def test_header_declares_package_exports():
    assert declared_symbols() == sorted(_pkg().EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = _pkg().lib()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def _cfg(**kw):
    pkg = _pkg()
    c = pkg.Config()
    c.dim = kw.get("dim", 3)
    c.n_bodies = kw.get("n", 10)
    for a, L in enumerate(kw.get("L", (10.0, 10.0, 10.0))):
        c.box_edges[a] = L
    c.boundary_range = kw.get("boundary_range", 1.0)
    c.body_interaction_range = kw.get("body_range", 1.0)
    c.spatial_cell_width = kw.get("cell_width", 1.0)
    c.model = kw.get("model", 0)
    c.params[0] = kw.get("diameter", 1.0)
    for i, v in enumerate(kw.get("params", ())):
        c.params[i] = v
    c.precision = kw.get("precision", 0)
    c.device = 0
    return c


@pytest.mark.parametrize("kw,code", [
    (dict(dim=4), -2),                                 # only 2D/3D cuboids are offloaded
    (dict(model=7), -2),                               # user-generic interactions keep the CPU path
    (dict(precision=5), -1),
    (dict(boundary_range=6.0), -1),                    # Periodic::new range check
    (dict(diameter=2.0), -1),                          # pair range beyond the boundary range
    (dict(body_range=0.5), -1),                        # checkerboard narrower than the interaction
    (dict(L=(1.5, 10.0, 10.0), boundary_range=0.7, diameter=0.7, body_range=0.8), -1),  # < 2 spaces
    (dict(cell_width=0.0), -1),                        # the microstate's VecCell width is required
    (dict(model=4, params=(1e4, 0.2, 100.0, 1.0), precision=1), -2),  # overlap penalty: FP64 only
    (dict(model=4, params=(-1.0, 0.2, 100.0, 1.0)), -1),              # k < 0
    (dict(model=4, params=(1e4, 0.2, float("nan"), 1.0)), -1),        # epsilon_shoulder not finite
    (dict(model=4, params=(1e4, 0.2, 100.0, 1.5)), -1),               # delta = r_cut beyond the ranges
])
def test_create_rejects(kw, code):
    pkg = _pkg()
    h = C.c_void_p()
    assert pkg.lib().hmc_create(C.byref(_cfg(**kw)), C.byref(h)) == code
    assert not h.value
