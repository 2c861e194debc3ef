"""Oracle checks for hard convex polyhedra (oracle/convex.hpp).

Pins: the reference's own overlap tables for the cube and octahedron
(hoomd-geometry/src/shape/convex_polytope.rs:652-880, with the test_overlap /
assert_symmetric_overlap helpers at :531-561), an independent separating-axis
test for cube pairs away from contact, closed forms of the quaternion
algebra, and the invariants of the rotation sampler and of the sweep.
"""
import math

import numpy as np
import pytest

import _oracle as O

OCTA = np.array([[-0.5, -0.5, 0.0], [0.5, -0.5, 0.0], [0.5, 0.5, 0.0], [-0.5, 0.5, 0.0],
                 [0.0, 0.0, 1.0 / math.sqrt(2.0)], [0.0, 0.0, -1.0 / math.sqrt(2.0)]])
IDENT = np.array([1.0, 0.0, 0.0, 0.0])


def axis_angle(axis, angle):
    """Versor::from_axis_angle (hoomd-vector/src/quaternion.rs:560-567)."""
    s = math.sin(angle / 2.0)
    return np.array([math.cos(angle / 2.0), axis[0] * s, axis[1] * s, axis[2] * s])


def conj(q):
    return np.array([q[0], -q[1], -q[2], -q[3]])


def local_overlap(verts, r_ab, q_a, q_b):
    """test_overlap (convex_polytope.rs:531-546): v_ij = q_a^-1 r_ab,
    o_ij = q_b * q_a^-1, then Convex::intersects_at (bounding spheres, then
    collide3d)."""
    inv = conj(q_a)
    v = O.quat_op(1, inv, np.r_[0.0, r_ab])[1:]
    o = O.quat_op(0, q_b, inv)
    R = math.sqrt(max(float(np.dot(x, x)) for x in verts))
    if float(np.dot(v, v)) >= (R + R) ** 2:
        return False
    return O.convex_collide(verts, v, o)


def symmetric(verts, r_ab, q_a, q_b, expected):
    r_ab = np.asarray(r_ab, float)
    assert local_overlap(verts, r_ab, q_a, q_b) == expected, (r_ab, expected)
    assert local_overlap(verts, -r_ab, q_b, q_a) == expected, (-r_ab, expected)


def test_overlap_octahedron_no_rot():  # convex_polytope.rs:683-782
    q = IDENT
    for r, e in [([0, 0, 0], True), ([10, 0, 0], False), ([1.1, 0, 0], False), ([0, 1.1, 0], False),
                 ([1.1, 0.2, 0], False), ([-1.1, 0.2, 0], False), ([-0.2, 1.1, 0], False),
                 ([-0.2, -1.1, 0], False), ([0.9, 0.2, 0], True), ([-0.9, 0.2, 0], True),
                 ([-0.2, 0.9, 0], True), ([-0.2, -0.9, 0], True), ([1.0, 0.2, 0], True)]:
        symmetric(OCTA, r, q, q, e)


def test_overlap_cube_no_rot():  # convex_polytope.rs:784-809
    q = IDENT
    for r, e in [([0, 0, 0], True), ([10, 0, 0], False), ([1.1, 0, 0], False), ([0, 1.1, 0], False),
                 ([0, 0, 1.1], False), ([1.1, 0.2, 0], False), ([-1.1, 0.2, 0], False),
                 ([-0.2, 1.1, 0], False), ([-0.2, -1.1, 0], False), ([0.9, 0.2, 0], True),
                 ([-0.9, 0.2, 0], True), ([-0.2, 0.9, 0], True), ([-0.2, -0.9, 0], True),
                 ([0.2, 0, 0], True), ([0.2, 0.00001, 0.00001], True), ([0.1, 0.2, 0.1], True),
                 ([1.0, 0.2, 0], True)]:
        symmetric(O.CUBE, r, q, q, e)


def test_overlap_cube_rot1():  # convex_polytope.rs:811-837
    qa, qb = IDENT, axis_angle([0, 0, 1], math.pi / 4.0)
    for r, e in [([10, 0, 0], False), ([1.3, 0, 0], False), ([-1.3, 0, 0], False), ([0, 1.3, 0], False),
                 ([0, -1.3, 0], False), ([1.3, 0.2, 0], False), ([-1.3, 0.2, 0], False),
                 ([-0.2, 1.3, 0], False), ([-0.2, -1.3, 0], False), ([1.2, 0.2, 0], True),
                 ([-1.2, 0.2, 0], True), ([-0.2, 1.2, 0], True), ([-0.2, -1.2, 0], True)]:
        symmetric(O.CUBE, r, qa, qb, e)


def test_overlap_cube_rot3():  # convex_polytope.rs:839-879
    q1 = axis_angle([1, 0, 0], math.pi / 4.0)
    q2 = axis_angle([0, 0, 1], math.pi / 4.0)
    qa, qb = IDENT, O.quat_op(0, q2, q1)
    for r, e in [([10, 0, 0], False), ([1.4, 0, 0], False), ([-1.4, 0, 0], False), ([0, 1.4, 0], False),
                 ([0, -1.4, 0], False), ([1.4, 0.2, 0], False), ([-1.4, 0.2, 0], False),
                 ([-0.2, 1.4, 0], False), ([-0.2, -1.4, 0], False), ([0, 1.2, 0], True),
                 ([0, 1.2, 0.1], True), ([0.1, 1.2, 0.1], True), ([1.2, 0, 0], True), ([1.2, 0.1, 0], True),
                 ([1.2, 0.1, 0.1], True), ([-0.9, 0.9, 0], True), ([-0.9, 0.899, 0.001], True),
                 ([0.9, -0.9, 0], True), ([-0.9, 0.9, 0.1], True)]:
        symmetric(O.CUBE, r, qa, qb, e)


def _rotation_matrix(q):
    a, b, c, d = q / np.linalg.norm(q)
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def _sat_gap(ra, Ra, rb, Rb):
    """Largest separation over the 15 separating axes of two unit cubes
    (centres ra, rb; rotation matrices Ra, Rb): > 0 separated, < 0 overlapping."""
    axes = [Ra[:, i] for i in range(3)] + [Rb[:, i] for i in range(3)]
    for i in range(3):
        for j in range(3):
            c = np.cross(Ra[:, i], Rb[:, j])
            n = np.linalg.norm(c)
            if n > 1e-9:
                axes.append(c / n)
    d = rb - ra
    best = -np.inf
    for ax in axes:
        pa = 0.5 * np.sum(np.abs(Ra.T @ ax))
        pb = 0.5 * np.sum(np.abs(Rb.T @ ax))
        best = max(best, abs(float(d @ ax)) - pa - pb)
    return best


def test_cube_pairs_against_separating_axes():
    rng = np.random.default_rng(7)
    checked = 0
    for _ in range(3000):
        qa = rng.normal(size=4)
        qb = rng.normal(size=4)
        qa /= np.linalg.norm(qa)
        qb /= np.linalg.norm(qb)
        ra = rng.uniform(-5, 5, 3)
        rb = ra + rng.normal(size=3) * 0.8
        gap = _sat_gap(ra, _rotation_matrix(qa), rb, _rotation_matrix(qb))
        if abs(gap) < 1e-6:
            continue
        assert O.convex_intersects_global(O.CUBE, ra, qa, rb, qb) == (gap < 0), (ra, qa, rb, qb, gap)
        checked += 1
    assert checked > 2500


def test_quaternion_rotation_forms_agree():
    rng = np.random.default_rng(3)
    for _ in range(200):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        x = rng.normal(size=3)
        a = O.quat_op(1, q, np.r_[0.0, x])[1:]
        b = O.quat_op(2, q, np.r_[0.0, x])[1:]
        c = _rotation_matrix(q) @ x
        assert np.allclose(a, c, atol=1e-13) and np.allclose(b, c, atol=1e-13)
        # q * conj(q) = identity
        p = O.quat_op(0, q, conj(q))
        assert np.allclose(p, IDENT, atol=1e-15)


def test_quaternion_product_closed_form():
    # rotations about one axis add their angles
    for t1, t2 in [(0.3, 0.4), (1.0, -2.0), (math.pi / 2, math.pi / 2)]:
        p = O.quat_op(0, axis_angle([0, 0, 1], t1), axis_angle([0, 0, 1], t2))
        assert np.allclose(p, axis_angle([0, 0, 1], t1 + t2), atol=1e-15)


def test_versor_displacement_invariants():
    for sigma in (0.05, 0.1, 0.5, 3.0):
        q = O.versor_displacement(3, 1, 42, 17, sigma, 4000)
        assert np.allclose(np.linalg.norm(q, axis=1), 1.0, atol=4e-16)
        assert np.all(q[:, 0] >= -1e-15)  # theta <= pi: w = cos(theta/2) >= 0
        if sigma <= 0.1:
            # small rotations: |s| is chi-distributed with mean sigma sqrt(8/pi)
            theta = 2.0 * np.arctan2(np.linalg.norm(q[:, 1:], axis=1), q[:, 0])
            assert abs(theta.mean() / (sigma * math.sqrt(8.0 / math.pi)) - 1.0) < 0.03


def _sc_cubes(n, phi):
    L = (n**3 / phi) ** (1.0 / 3.0)
    a = L / n
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    pos = (np.stack([i.ravel(), j.ravel(), k.ravel()], 1) + 0.5) * a - L / 2
    q = np.tile(IDENT, (len(pos), 1))
    return np.array([L, L, L]), pos, q


def test_convex_sweep_stays_overlap_free_and_counts():
    L, pos, q = _sc_cubes(6, 0.55)
    s = O.ConvexOracle(L, pos, q, O.CUBE, seed=5)
    assert not s.any_overlap()
    s.trace(True)
    for step in range(4):
        a1, r1 = s.apply(False, 0.05)
        a2, r2 = s.apply(True, 0.1)
        assert a1 + r1 >= len(pos) and a2 + r2 >= len(pos)
        t = s.read_trace()
        assert (t["status"] != 2).sum() == a1 + r1 + a2 + r2
        s.increment_step()
        assert not s.any_overlap()
    bp, bq, sp, sq = s.state()
    assert np.allclose(np.linalg.norm(bq, axis=1), 1.0, atol=1e-14)
    assert np.array_equal(bq, sq)  # site orientation = body orientation * identity


def test_convex_sweep_is_thread_count_independent():
    L, pos, q = _sc_cubes(5, 0.5)
    states = []
    for threads in (1, 4):
        O.set_threads(threads)
        s = O.ConvexOracle(L, pos, q, O.CUBE, seed=11)
        for _ in range(3):
            s.apply(False, 0.08)
            s.apply(True, 0.2)
            s.increment_step()
        states.append(s.state())
    O.set_threads(O.lib().oracle_max_threads())
    for a, b in zip(*states):
        assert np.array_equal(a, b)
