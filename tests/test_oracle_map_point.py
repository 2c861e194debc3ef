"""The oracle's Hypercuboid::map_point (oracle_map_points) and the package's
scale_volume, against the reference's own tests
(hoomd-geometry/src/shape/cuboid.rs `test_scale_volume`, `test_map_basic`,
`test_map_corner`)."""
import math

import numpy as np

import _oracle as O
import _pkg


def test_scale_volume_known_answers():
    sv = _pkg.pkg().scale_volume
    assert sv([6.0, 10.0], 4.0) == [12.0, 20.0]
    assert sv([6.0, 10.0], 0.25) == [3.0, 5.0]


def test_map_basic_known_answers():
    a, b = [6.0, 10.0], [12.0, 5.0]

    def m(src, dst, pt):
        out, n_out = O.map_points(src, dst, [pt])
        return None if n_out else tuple(out[0])

    assert m(a, b, [0.0, 0.0]) == (0.0, 0.0)
    assert m(b, a, [0.0, 0.0]) == (0.0, 0.0)
    assert m(a, b, [100.0, 0.0]) is None
    assert m(b, a, [0.0, -200.0]) is None
    assert m(a, b, [2.0, 1.0]) == (4.0, 0.5)
    assert m(b, a, [-4.0, 0.5]) == (-2.0, 1.0)
    assert m(a, b, [-3.0, -5.0]) == (-6.0, -2.5)
    assert m(b, a, [-6.0, -2.5]) == (-3.0, -5.0)


def test_map_corner_stays_inside():
    # points on the boundary of one box map inside the other (the clamp)
    rng = np.random.default_rng(3)
    ab = rng.uniform(1.0, 1000.0, size=(4096, 2))
    hits = 0
    for a, b in ab:
        pts = [[-a / 2, -a / 2], [math.nextafter(a / 2, 0.0), math.nextafter(a / 2, 0.0)]]
        out, n_out = O.map_points([a, a], [b, b], pts)
        assert n_out == 0
        assert ((-b / 2 <= out) & (out < b / 2)).all(), (a, b, out)
        # without the clamp the scaled corner would fall outside now and then
        hits += int((b / a) * math.nextafter(a / 2, 0.0) >= b / 2)
    assert hits > 0
