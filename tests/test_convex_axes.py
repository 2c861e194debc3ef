"""The separating-axis data that the mixed-precision convex pass builds on the
host from a polytope's vertices (hmc_debug_convex_axes): face axes, edge
directions and scanned vertices of known polytopes, against what their
geometry fixes (a cube has 3 face axes and 3 edge directions, an octahedron
4 and 6, a tetrahedron 4 and 6), and the cases that switch the FP32 tests
off (flat and many-faced vertex sets)."""
import math

import numpy as np
import pytest

import _oracle as O
import _pkg

OCTA = np.array([[-0.5, -0.5, 0.0], [0.5, -0.5, 0.0], [0.5, 0.5, 0.0], [-0.5, 0.5, 0.0],
                 [0.0, 0.0, 1.0 / math.sqrt(2.0)], [0.0, 0.0, -1.0 / math.sqrt(2.0)]])
TETRA = np.array([[1.0, 1.0, 1.0], [1.0, -1.0, -1.0], [-1.0, 1.0, -1.0], [-1.0, -1.0, 1.0]]) * 0.5


@pytest.fixture(scope="module")
def hmc():
    p = _pkg.pkg()
    p.lib()
    return p


def _same_axes(got, want):
    """got, want: unit vectors, each set up to sign and order."""
    got = np.asarray(got, float)
    want = np.asarray(want, float)
    want = want / np.linalg.norm(want, axis=1)[:, None]
    assert len(got) == len(want)
    for w in want:
        assert np.max(np.abs(got @ w)) > 1.0 - 1e-6, w


def _check_slabs(ax, verts):
    for f in ax["faces"]:
        n, lo, hi = f[:3], f[3], f[4]
        assert abs(np.linalg.norm(n) - 1.0) < 1e-6
        p = verts @ n
        assert abs(lo - p.min()) < 1e-6 and abs(hi - p.max()) < 1e-6


def test_cube(hmc):
    ax = hmc.convex_axes(O.CUBE)
    assert (ax["n_faces"], ax["n_edges"], ax["n_scanned"], ax["sym"], ax["use"], ax["use_edges"]) == \
        (3, 3, 4, True, True, True)
    _same_axes(ax["faces"][:, :3], np.eye(3))
    _same_axes(ax["edges"], np.eye(3))
    _check_slabs(ax, O.CUBE)


def test_octahedron(hmc):
    ax = hmc.convex_axes(OCTA)
    assert (ax["n_faces"], ax["n_edges"], ax["n_scanned"], ax["sym"], ax["use"]) == (4, 6, 3, True, True)
    # edges: the 4 of the square and the 8 to the tips, in 6 directions
    h = 1.0 / math.sqrt(2.0)
    _same_axes(ax["edges"], [[1, 0, 0], [0, 1, 0], [0.5, 0.5, h], [-0.5, 0.5, h], [0.5, -0.5, h],
                             [-0.5, -0.5, h]])
    _check_slabs(ax, OCTA)


def test_tetrahedron(hmc):
    ax = hmc.convex_axes(TETRA)
    assert (ax["n_faces"], ax["n_edges"], ax["n_scanned"], ax["sym"], ax["use"]) == (4, 6, 4, False, True)
    _same_axes(ax["faces"][:, :3], [[1, 1, -1], [1, -1, 1], [-1, 1, 1], [1, 1, 1]])
    _same_axes(ax["edges"], [[0, 1, 1], [0, 1, -1], [1, 0, 1], [1, 0, -1], [1, 1, 0], [1, -1, 0]])
    _check_slabs(ax, TETRA)


def test_interior_and_duplicate_vertices_change_nothing(hmc):
    extra = np.vstack([O.CUBE, [[0.0, 0.0, 0.0], [0.1, 0.2, -0.1], [0.5, 0.5, 0.5]]])
    ax = hmc.convex_axes(extra)
    assert (ax["n_faces"], ax["n_edges"], ax["use"], ax["sym"]) == (3, 3, True, False)
    _check_slabs(ax, extra)


def test_flat_and_many_faced_sets_turn_the_fp32_tests_off(hmc):
    square = np.array([[-0.5, -0.5, 0.0], [0.5, -0.5, 0.0], [0.5, 0.5, 0.0], [-0.5, 0.5, 0.0]])
    assert not hmc.convex_axes(square)["use"]
    assert not hmc.convex_axes(np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]]))["use"]
    rng = np.random.default_rng(1)
    ball = rng.normal(size=(32, 3))
    ball /= np.linalg.norm(ball, axis=1)[:, None]
    ax = hmc.convex_axes(ball)
    assert ax["n_faces"] > 16 and not ax["use"]
