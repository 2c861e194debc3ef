"""Pin the oracle (the CPU restatement) against the reference's own known-answer
tests, independent implementations (numpy SFC64) and closed forms.

Each test names the reference test it transcribes.  Values are data copied from
the reference's tests; the code under test is the restatement in oracle/.
"""
import json
import math
import pathlib

import numpy as np
import pytest

import _oracle as O

GOLDEN = pathlib.Path(__file__).with_name("golden")


# ---------------------------------------------------------------- hoomd-rand --
# hoomd-rand/src/sfc.rs:158-168 (values produced by numpy's SFC64)
SFC64_REFERENCE_OUTPUT = [
    4_237_781_876_154_851_393, 17_705_428_440_413_258_140, 1_322_197_197_711_907_681,
    822_724_228_132_957_142, 2_474_202_602_039_083_746, 5_912_426_283_212_852_001,
    15_821_317_571_115_833_757, 10_375_476_962_501_160_791, 16_772_721_532_102_950_986,
    10_344_472_337_605_944_266, 17_980_158_625_826_311_727, 5_068_054_200_821_024_954,
    7_387_740_087_731_720_141, 3_233_965_506_304_800_125, 805_889_469_043_559_033,
    5_406_730_423_683_535_818, 15_071_469_935_640_508_508,
]
# sfc.rs:128-155: the 12 warm-up outputs of state [0,0,0,1]
SFC64_WARMUP = [1, 2, 12, 150994975, 2533275243454595, 7601064794802825, 81067317779357880,
                939904496648135862, 8358283255545778679, 3766502869832372394,
                3749490695777847966, 10585829324311968283]


def test_sfc64_against_numpy_kat():
    # sfc.rs:190-200: seed_from_u64(0) == initialize(0, 0, 0, 1)
    assert [int(x) for x in O.sfc64_stream(0, 0, 0, 1, 17)] == SFC64_REFERENCE_OUTPUT


def test_sfc64_warmup_outputs():
    assert [int(x) for x in O.sfc64_stream(0, 0, 0, 1, 12, warm=False)] == SFC64_WARMUP


def test_sfc64_deep():
    # sfc.rs:207-214: value number 2^17 after seed_from_u64(0)
    assert int(O.sfc64_stream(0, 0, 0, 1, 2**17)[-1]) == 4_977_758_738_274_538_201


def test_counter_streams_match_numpy_vectors():
    data = json.loads((GOLDEN / "sfc64_counter_vectors.json").read_text())
    assert data["deep_seed0_2pow17"] == "4977758738274538201"
    for v in data["vectors"]:
        got = O.counter_stream(v["step"], v["substep"], v["seed"], v["index"], 32)
        assert [str(int(x)) for x in got] == v["draws"], v


def test_counter_independent():
    # hoomd-rand/src/counter.rs:179-210: each element of the counter changes the stream
    streams = [O.sample(0, 0, 0, 0, "f64", 256), O.sample(1, 0, 0, 0, "f64", 256),
               O.sample(0, 1, 0, 0, "f64", 256), O.sample(0, 0, 1, 0, "f64", 256),
               O.sample(0, 0, 0, 1 << 32, "f64", 256), O.sample(0, 0, 0, 1, "f64", 256),
               O.sample(0, 0, 0, 2, "f64", 256)]
    for i, a in enumerate(streams):
        for j, b in enumerate(streams):
            assert np.array_equal(a, b) == (i == j)


# --------------------------------------------------- rand adapters (unpinned) --
def test_uniform_f64_uses_high_53_bits():
    raw = O.counter_stream(7, 3, 42, 99, 64)
    got = O.sample(7, 3, 42, 99, "f64", 64)
    exact = np.array([float(int(x) >> 11) for x in raw]) / 2.0**53
    assert np.array_equal(got, exact)
    assert got.min() >= 0.0 and got.max() < 1.0


def test_open01_strictly_inside():
    x = O.sample(1, 2, 3, 4, "open01", 100000)
    assert x.min() > 0.0 and x.max() < 1.0
    raw = O.counter_stream(1, 2, 3, 4, 100000)
    # closed form: (m + 0.5) * 2^-52 with m the 52 high bits
    exact = (np.array([float(int(v) >> 12) for v in raw[:1000]]) + 0.5) / 2.0**52
    assert np.array_equal(x[:1000], exact)


@pytest.mark.parametrize("n", [1, 2, 3, 7, 1000, 2**31 + 11])
def test_index_in_range_and_canon_closed_form(n):
    idx = O.sample(5, 0, 42, 17, "index", 2000, param=n)
    assert idx.min() >= 0 and idx.max() < n
    # Re-derive the first draw from the raw stream: floor(u32 * n / 2^32), plus
    # one only when the low product word exceeds 2^32 - n and a second draw
    # overflows it.
    raw = O.counter_stream(5, 0, 42, 17, 2)
    x = int(raw[0]) & 0xFFFFFFFF
    hi, lo = (x * n) >> 32, (x * n) & 0xFFFFFFFF
    if lo > (2**32 - n) % 2**32:
        y = int(raw[1]) & 0xFFFFFFFF
        hi += 1 if lo + ((y * n) >> 32) > 0xFFFFFFFF else 0
    assert int(idx[0]) == hi


def test_index_uniform():
    idx = O.sample(11, 0, 42, 3, "index", 60000, param=3)
    counts = np.bincount(idx.astype(int), minlength=3)
    assert np.all(np.abs(counts - 20000) < 5 * math.sqrt(20000 * 2 / 3))


def test_ziggurat_tables_closed_forms():
    x, f = O.zig_tables()
    R, V = 3.6541528853610088, 0.00492867323399
    # values remembered from rand_distr's published table (first and last layers)
    assert x[0] == 3.910757959537090045 and x[1] == 3.654152885361008796
    assert x[2] == 3.449278298560964462 and x[255] == 0.215241895913273806 and x[256] == 0.0
    assert f[0] == 0.000477467764586655 and f[256] == 1.0
    assert x[1] == R
    assert np.all(np.diff(x[:257]) < 0)
    assert np.allclose(f, np.exp(-x * x / 2), rtol=1e-15, atol=1e-18)
    # every layer holds area V: x[i] * (f[i+1] - f[i]) == V  (ZIGNOR construction)
    areas = x[1:255] * (f[2:256] - f[1:255])
    assert np.allclose(areas, V, rtol=1e-9)
    assert math.isclose(x[0] * f[1], V, rel_tol=1e-12)


def test_standard_normal_moments_and_ks():
    from scipy import stats
    z = np.concatenate([O.sample(3, 0, 42, k, "normal", 20000) for k in range(10)])
    assert abs(z.mean()) < 5 / math.sqrt(len(z))
    assert abs(z.var() - 1) < 5 * math.sqrt(2 / len(z))
    assert stats.kstest(z, "norm").pvalue > 1e-4
    # the tail beyond R is sampled too
    assert np.sum(np.abs(z) > 3.6541528853610088) > 0


# ---------------------------------------------------- hoomd-mc checkerboard --
def test_multi_index_to_index_2d():  # hypercuboid.rs:430-442
    s = [12, 4]
    for m, want in [([0, 0], 0), ([0, 1], 1), ([0, 2], 2), ([0, 3], 3), ([1, 0], 4), ([1, 1], 5),
                    ([1, 2], 6), ([1, 3], 7), ([11, 3], 47)]:
        assert O.multi_to_index(m, s) == want


def test_multi_index_to_index_3d():  # hypercuboid.rs:444-461
    s = [12, 4, 6]
    for m, want in [([0, 0, 0], 0), ([0, 0, 5], 5), ([0, 1, 0], 6), ([0, 2, 0], 12), ([0, 3, 0], 18),
                    ([1, 0, 0], 24), ([2, 0, 0], 48), ([3, 3, 2], 92)]:
        assert O.multi_to_index(m, s) == want


def test_multi_index_to_index_4d():  # hypercuboid.rs:463-494
    s = [12, 4, 6, 8]
    assert O.multi_to_index([11, 3, 5, 7], s) == 2303
    assert O.multi_to_index([9, 0, 0, 0], s) == 9 * 4 * 6 * 8
    assert O.multi_to_index([3, 3, 2, 0], s) == 92 * 8


def test_compute_dimensions_exact():  # hypercuboid.rs:496-523
    w, s = O.compute_dimensions([16.0, 24.0], 2.0, True)
    assert list(w) == [2.0, 2.0] and list(s) == [8, 12]
    w, s = O.compute_dimensions([14.0, 22.0], 2.0, False)
    assert list(w) == [2.0, 2.0] and list(s) == [8, 12]


def test_compute_dimensions_expand():  # hypercuboid.rs:525-544
    w, s = O.compute_dimensions([15.0, 23.0], 1.0, True)
    assert list(w) == [15.0 / 14.0, 23.0 / 22.0] and list(s) == [14, 22]
    w, s = O.compute_dimensions([14.0, 22.0], 1.0, False)
    assert list(w) == [14.0 / 13.0, 22.0 / 21.0] and list(s) == [14, 22]


def test_compute_dimensions_too_large_range():  # hypercuboid.rs:196-199 assert
    assert O.compute_dimensions([3.0, 10.0], 2.0, True) is None


def test_space_indices_by_color_2d():  # hypercuboid.rs:546-583
    c = O.by_color([6, 4])
    assert c.tolist() == [[0, 2, 8, 10, 16, 18], [1, 3, 9, 11, 17, 19], [4, 6, 12, 14, 20, 22],
                          [5, 7, 13, 15, 21, 23]]


def test_space_indices_by_color_3d():  # hypercuboid.rs:585-684
    assert O.by_color([2, 2, 2]).tolist() == [[k] for k in range(8)]
    c = O.by_color([2, 6, 4])
    base = [[0, 2, 8, 10, 16, 18], [1, 3, 9, 11, 17, 19], [4, 6, 12, 14, 20, 22], [5, 7, 13, 15, 21, 23]]
    assert c.tolist() == base + [[24 + v for v in row] for row in base]


def test_space_indices_by_color_4d():  # hypercuboid.rs:686-790
    c = O.by_color([2, 4, 2, 4])
    first = [[0, 2, 16, 18], [1, 3, 17, 19], [4, 6, 20, 22], [5, 7, 21, 23], [8, 10, 24, 26],
             [9, 11, 25, 27], [12, 14, 28, 30], [13, 15, 29, 31]]
    assert c.tolist() == first + [[32 + v for v in row] for row in first]


def test_point_to_space_index_periodic():  # hypercuboid.rs:792-821
    pts = [[-8.0, -12.1], [-8.0, -12.0], [-8.0, -10.0], [-8.0, -8.0], [-8.0, -6.0], [-8.0, 11.9],
           [-8.0, 12.0], [-8.1, 0.0], [-8.0, 0.0], [-6.0, 0.0], [-4.0, 0.0], [-2.0, 0.0], [7.9, 0.0],
           [8.0, 0.0], [7.9, 11.9]]
    want = [-1, 0, 1, 2, 3, 11, 0, -1, 6, 18, 30, 42, 90, 6, 95]
    assert O.point_to_space_fixed(2.0, [16.0, 24.0], True, pts).tolist() == want


def test_point_to_space_index_nonperiodic():  # hypercuboid.rs:823-854
    pts = [[-7.0, -11.1], [-7.0, -11.0], [-7.0, -9.0], [-7.0, -7.0], [-7.0, -5.0], [-7.0, 10.0],
           [-7.0, 11.0], [-7.0, 12.9], [-7.0, 13.0], [-7.1, 1.0], [-7.0, 1.0], [-5.0, 1.0],
           [-3.0, 1.0], [-1.0, 1.0], [7.0, 1.0], [9.0, 1.0], [8.9, 12.9]]
    want = [-1, 0, 1, 2, 3, 10, 11, 11, -1, -1, 6, 18, 30, 42, 90, -1, 95]
    assert O.point_to_space_fixed(2.0, [14.0, 22.0], False, pts).tolist() == want


def test_random_origin_covers_box():
    # hypercuboid.rs:856-892 (periodic case): every point of the box maps to a space.
    rng = np.random.default_rng(1)
    for k in range(64):
        o, w, s = O.checkerboard(2, k, 0, 42, 1.5, [10.0, 7.0])
        assert np.all(o <= -np.array([5.0, 3.5])) and np.all(o > -np.array([5.0, 3.5]) - w)
        pts = rng.uniform([-5.0, -3.5], [5.0, 3.5], size=(256, 2))
        idx = ((pts - o) // w).astype(np.int64)
        idx[idx == s] = 0
        assert np.all((idx >= 0) & (idx < s))


def test_checkerboard_origin_draws():
    # ParallelSweep::update_checkerboard (parallel_sweep.rs:182): Counter index u64::MAX,
    # offsets drawn in axis order (hypercuboid.rs:325-330).
    o, w, s = O.checkerboard(3, 9, 2, 42, 1.0, [101.5, 101.5, 101.5])
    u = O.sample(9, 2, 42, 2**64 - 1, "f64", 3)
    assert list(s) == [100, 100, 100]
    assert np.array_equal(o, -101.5 / 2.0 - u * w)


# ------------------------------------------------- periodic cuboid boundary --
def test_wrap_1d():  # cuboid.rs:388-403
    assert O.wrap([20.0], [[5.0], [10.0], [25.0]]).ravel().tolist() == [5.0, -10.0, 5.0]


def test_wrap_2d():  # cuboid.rs:485-518
    L = [20.0, 20.0]
    nd = np.nextafter(10.0, 0.0)
    cases = [([5.0, 3.0], [5.0, 3.0]), ([-10.0, -10.0], [-10.0, -10.0]), ([nd, nd], [nd, nd]),
             ([10.0, 10.0], [-10.0, -10.0]), ([20.0, 20.0], [0.0, 0.0]), ([30.0, 30.0], [-10.0, -10.0]),
             ([25.0, -35.0], [5.0, 5.0])]
    for p, want in cases:
        assert O.wrap(L, [p]).ravel().tolist() == want


def test_wrap_doc_example():  # cuboid.rs:79-88
    assert O.wrap([10.0, 10.0], [[6.0, -15.0]]).ravel().tolist() == [-4.0, -5.0]


def test_wrap_moves_points_already_inside():
    # wrap computes (x/L)*L, which is not the identity in floating point even for
    # points inside the box; the GPU path must repeat every wrap the reference does.
    L = [101.55552806605625] * 3
    rng = np.random.default_rng(0)
    p = rng.uniform(-50.7, 50.7, size=(100000, 3))
    w1 = O.wrap(L, p)
    assert 0.05 < np.mean(w1 != p) < 0.2
    assert np.max(np.abs(w1 - p)) < 1e-13


def _as_set(a):
    return sorted(tuple(x) for x in np.asarray(a).tolist())


def test_ghosts_1d():  # cuboid.rs:405-443
    assert len(O.ghosts([20.0], 1.0, [0.0])) == 0
    assert len(O.ghosts([20.0], 1.0, [10.5])) == 0
    assert _as_set(O.ghosts([20.0], 1.0, [9.5])) == [(-10.5,)]
    assert _as_set(O.ghosts([20.0], 1.0, [-9.5])) == [(10.5,)]


def test_ghosts_2d():  # cuboid.rs:574-617
    L = [20.0, 10.0]
    assert len(O.ghosts(L, 1.0, [10.5, 0.0])) == 0
    assert len(O.ghosts(L, 1.0, [0.0, 5.5])) == 0
    assert _as_set(O.ghosts(L, 1.0, [9.5, 0.0])) == [(-10.5, 0.0)]
    assert _as_set(O.ghosts(L, 1.0, [-9.5, 0.0])) == [(10.5, 0.0)]
    assert _as_set(O.ghosts(L, 1.0, [0.0, 4.5])) == [(0.0, -5.5)]
    assert _as_set(O.ghosts(L, 1.0, [0.0, -4.5])) == [(0.0, 5.5)]
    assert _as_set(O.ghosts(L, 1.0, [9.5, 4.5])) == _as_set([[-10.5, 4.5], [9.5, -5.5], [-10.5, -5.5]])
    assert _as_set(O.ghosts(L, 1.0, [9.5, -4.5])) == _as_set([[-10.5, -4.5], [9.5, 5.5], [-10.5, 5.5]])
    assert _as_set(O.ghosts(L, 1.0, [-9.5, 4.5])) == _as_set([[10.5, 4.5], [-9.5, -5.5], [10.5, -5.5]])
    assert _as_set(O.ghosts(L, 1.0, [-9.5, -4.5])) == _as_set([[10.5, -4.5], [-9.5, 5.5], [10.5, 5.5]])


def test_ghosts_2d_bulk_has_none():  # cuboid.rs:520-548
    rng = np.random.default_rng(1)
    for p in rng.uniform([-9.0, -4.0], [9.0, 4.0], size=(1024, 2)):
        assert len(O.ghosts([20.0, 10.0], 1.0, p)) == 0


def test_ghosts_3d_corner_has_seven():
    g = O.ghosts([10.0, 10.0, 10.0], 1.0, [4.5, -4.5, 4.9])
    assert len(g) == 7 and len(set(map(tuple, g.tolist()))) == 7


# ---------------------------------------------------------------- VecCell ----
def test_stencil_order_3d():
    # vec_cell.rs:167-177: row-major cube stably sorted by the lexicographic
    # order of the squared offsets.
    st = O.stencil(3, 1)
    assert st.shape == (27, 3)
    assert st[0].tolist() == [0, 0, 0]
    assert st[1].tolist() == [0, 0, -1] and st[2].tolist() == [0, 0, 1]
    assert st[3].tolist() == [0, -1, 0] and st[4].tolist() == [0, 1, 0]
    key = [tuple(v * v for v in row) for row in st.tolist()]
    assert key == sorted(key)
    assert sorted(map(tuple, st.tolist())) == sorted((a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1)
                                                     for c in (-1, 0, 1))


def test_cell_index():  # vec_cell.rs:848-861
    assert O.cell_index(2.0, [0.0, 0.0, 0.0]).tolist() == [0, 0, 0]
    assert O.cell_index(2.0, [2.0, 0.0, 0.0]).tolist() == [1, 0, 0]
    assert O.cell_index(2.0, [0.0, 2.0, 0.0]).tolist() == [0, 1, 0]
    assert O.cell_index(2.0, [0.0, 0.0, 2.0]).tolist() == [0, 0, 1]
    assert O.cell_index(2.0, [-41.5, 18.5, -0.125]).tolist() == [-21, 9, -1]


# --------------------------------------------------------- translate / Ball --
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("r", [0.08, 0.5, 12.0])
def test_ball_inside_and_uniform(dim, r):
    # distribution.rs:82-104 and translate/cartesian.rs:65-100: |delta| < r, mean 0,
    # and uniform in the ball: E|v|^2 = D r^2 / (D + 2).
    import ctypes as C
    n = 20000
    v = np.zeros((n, dim))
    pos = np.zeros(dim)
    out = np.zeros(dim)
    ch = np.zeros(1, np.uint32)
    for k in range(n):
        O.lib().oracle_translate(dim, 1, 0, 42, k, 1, O._p(pos, O.D_P), r, O._p(out, O.D_P),
                                 O._p(ch, O.U32_P))
        v[k] = out
    rr = (v * v).sum(1)
    assert rr.max() < r * r
    assert np.all(np.abs(v.mean(0)) < 5 * r / math.sqrt(n))
    assert abs(rr.mean() / (r * r) - dim / (dim + 2)) < 0.02
