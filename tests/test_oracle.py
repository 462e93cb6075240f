"""The CPU oracle pinned against the reference's logged answers, hand-built
known answers and independent algorithms (brute force, sweep-and-prune)."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def box(x0, x1, y0, y1):
    return np.array([x0, x1, y0, y1], np.float32)


def test_predicate_known_answers(orc):
    # closed intervals (src/pbig/pbig.cu:996-1001): shared edge or corner overlaps
    assert orc.overlap(box(0, 1, 0, 1), box(1, 2, 0, 1))
    assert orc.overlap(box(0, 1, 0, 1), box(1, 2, 1, 2))
    assert not orc.overlap(box(0, 1, 0, 1), box(np.nextafter(np.float32(1), np.float32(2)), 2, 0, 1))
    assert orc.overlap(box(0, 10, 0, 10), box(2, 3, 2, 3))      # nesting
    assert orc.overlap(box(0, 10, 4, 5), box(4, 5, 0, 10))      # cross
    assert not orc.overlap(box(0, 10, 0, 1), box(0, 10, 2, 3))  # apart in y only
    assert orc.overlap(box(-0.0, 0, 0, 0), box(0, 0, -0.0, 0))  # -0.0 == 0.0
    # symmetric
    rng = np.random.default_rng(0)
    for _ in range(200):
        a = np.sort(rng.integers(0, 6, 4).astype(np.float32).reshape(2, 2), 1).reshape(4)
        b = np.sort(rng.integers(0, 6, 4).astype(np.float32).reshape(2, 2), 1).reshape(4)
        assert orc.overlap(a, b) == orc.overlap(b, a)
        want = a[0] <= b[1] and b[0] <= a[1] and a[2] <= b[3] and b[2] <= a[3]
        assert orc.overlap(a, b) == want


def test_brute_force_small_closed_form(orc):
    # k identical boxes: every pair overlaps
    b = np.tile(box(1, 2, 1, 2), (37, 1))
    assert len(orc.brute_pairs(b)) == 37 * 36 // 2
    # a chain of unit boxes touching at edges: k-1 pairs
    b = np.stack([box(i, i + 1, 0, 1) for i in range(20)])
    got = orc.sort_pairs(orc.brute_pairs(b))
    assert got.tolist() == [[i, i + 1] for i in range(19)]
    # a lattice of disjoint boxes: no pair
    g = [box(3 * i, 3 * i + 1, 3 * j, 3 * j + 1) for i in range(10) for j in range(10)]
    assert len(orc.brute_pairs(np.stack(g))) == 0


def test_cellmap_is_monotone_and_clamped(orc):
    inv, nrc = np.float32(1 / 64), 16
    assert orc.cellmap(-5.0, inv, nrc) == 0
    assert orc.cellmap(0.0, inv, nrc) == 0
    assert orc.cellmap(63.999996, inv, nrc) == 0
    assert orc.cellmap(64.0, inv, nrc) == 1
    assert orc.cellmap(1023.9999, inv, nrc) == 15
    assert orc.cellmap(5000.0, inv, nrc) == 15
    assert orc.cellmap(float("inf"), inv, nrc) == 15
    assert orc.cellmap(float("-inf"), inv, nrc) == 0
    vals = np.sort(np.random.default_rng(1).uniform(-100, 1200, 2000).astype(np.float32))
    cells = [orc.cellmap(float(v), inv, nrc) for v in vals]
    assert all(a <= b for a, b in zip(cells, cells[1:]))


def _random_boxes(rng, n, L, smin, smax):
    c = rng.uniform(0, L, size=(n, 2))
    w = rng.uniform(smin, smax, size=(n, 2))
    return np.stack([c[:, 0] - w[:, 0] / 2, c[:, 0] + w[:, 0] / 2,
                     c[:, 1] - w[:, 1] / 2, c[:, 1] + w[:, 1] / 2], 1).astype(np.float32)


@pytest.mark.parametrize("seed", range(6))
def test_three_algorithms_agree(orc, seed):
    rng = np.random.default_rng(seed)
    n = [500, 2000, 3000, 1500, 800, 2500][seed]
    L = [64.0, 256.0, 300.0, 128.0, 1000.0, 50.0][seed]
    b = _random_boxes(rng, n, L, 0.5, [4, 8, 40, 2, 200, 1][seed])
    brute = orc.sort_pairs(orc.brute_pairs(b))
    assert np.array_equal(brute, orc.sort_pairs(orc.sweep_pairs(b)))
    for cs in (1, 4, 16, 64, 13.0):
        grid = orc.sort_pairs(orc.grid_pairs(b, L, cs))
        assert np.array_equal(brute, grid), (seed, cs)


def test_grid_reports_each_pair_once(orc):
    # boxes spanning many cells: no duplicates (reference-cell rule)
    rng = np.random.default_rng(3)
    b = _random_boxes(rng, 400, 100.0, 5, 60)
    pairs = orc.grid_pairs(b, 100.0, 4)
    s = orc.sort_pairs(pairs)
    assert len(np.unique(s, axis=0)) == len(s) == len(pairs)
    assert np.all(s[:, 0] < s[:, 1])


def test_decide_cell_size_closed_form(orc):
    # 4M boxes, range 65536, mean extent 16 -> stops at cs=128 (20 boxes/cell) -> returns 256
    from pbig_amd import workloads

    b, L = workloads.c2_large_uniform(n=4194304)
    assert orc.decide_cell_size(b, L) == 256
    b1, L1 = workloads.c1_small_uniform()
    assert orc.decide_cell_size(b1, L1) == 64
    # a single box: the loop stops at once (fewer than 32 per cell)
    assert orc.decide_cell_size(np.array([[0, 1, 0, 1]], np.float32), 1024.0) == 1024


def test_reference_3d_generator_and_cell_rule_small(orc):
    b = orc.ref_generate(1000, 10000, 100, "UU", 3)
    # xmax = Inc(pos, size, range) <= range - 1; sizes < 100
    assert b[:, 1::2].max() <= 9999
    assert np.all(b[:, 1::2] - b[:, 0::2] < 100)
    # the reference's own trace at N=1e7: avgX = 49 and cs = 128
    # (benchmarks/test3/filter_file:5-13); its first 1000 boxes drive the rule
    big = orc.ref_generate(10**7, 10000, 100, "UN", 3)
    assert orc.decide_cell_size_u32(big, 3, 10000) == 128


def test_brute_u32_equals_grid_u32(orc):
    """The unsigned brute force reports what the KAT-pinned grid restatement
    counts, in 3-D and 2-D, at several cell sizes."""
    for dims in (3, 2):
        for sim in ("UU", "NU", "UN"):
            b = orc.ref_generate(3000, 2000, 120, sim, dims)
            bf = orc.brute_u32(b, dims)
            assert len(np.unique(bf.view(np.uint64))) == len(bf)
            for cs in (32, 64, 256, 2048):
                assert orc.grid_u32(b, dims, 2000, cs)["pairs"] == len(bf), (dims, sim, cs)


def test_reference_logged_answer_test1(orc):
    """benchmarks/test1/filter_file:7 -- 10^7 boxes, UU, S=100, range 1e4,
    3-D unsigned, srand(100): 49,514,252 pairs."""
    b = orc.ref_generate(10**7, 10000, 100, "UU", 3)
    cs = orc.decide_cell_size_u32(b, 3, 10000)
    assert cs == 128
    st = orc.grid_u32(b, 3, 10000, cs)
    assert st["pairs"] == 49514252


@pytest.mark.slow
def test_reference_logged_answer_test3(orc):
    """benchmarks/test3/filter_file:17-49 -- 10^7 boxes, UN (normal centres):
    cs 128, 493,039 cells, 26,668,901 memberships, max 3,543 in cell
    246,442, 200,096 jobs, 16,683,783,392 compared pairs, 1,121,815,231 pairs."""
    b = orc.ref_generate(10**7, 10000, 100, "UN", 3)
    st = orc.grid_u32(b, 3, 10000, 128)
    assert st["n_cells"] == 493039
    assert st["memberships"] == 26668901
    assert st["max_per_cell"] == 3543 and st["max_cell"] == 246442
    assert st["jobs256"] == 200096
    assert st["compared_pairs"] == 16683783392
    assert st["pairs"] == 1121815231


def test_golden_c1_fixture(orc):
    from pbig_amd import workloads

    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        man = json.load(f)["c1_uniform_16k"]
    b, L = workloads.c1_small_uniform()
    assert [float(v) for v in b[0]] == man["first_box"]
    golden = np.load(os.path.join(GOLDEN, "c1_pairs.npy"))
    assert len(golden) == man["pairs"]
    assert orc.pair_hash(golden) == man["sha256"]
    assert np.array_equal(orc.sort_pairs(orc.brute_pairs(b)), golden)
    assert np.array_equal(orc.sort_pairs(orc.grid_pairs(b, L, man["reference_cell_size"])), golden)


def test_planted_mistakes_are_caught(orc):
    """Plausible mistakes in a restatement change the C1 answer: an open
    interval, a dropped y test, or a reference cell taken from the max corner."""
    from pbig_amd import workloads

    b, L = workloads.c1_small_uniform()
    golden = np.load(os.path.join(GOLDEN, "c1_pairs.npy"))
    x0, x1, y0, y1 = (b[:, k] for k in range(4))
    i, j = golden[:, 0], golden[:, 1]
    strict = (x0[i] < x1[j]) & (x0[j] < x1[i]) & (y0[i] < y1[j]) & (y0[j] < y1[i])
    assert strict.all()  # these pairs all overlap, so an open-interval
    # predicate can only differ through exact contacts; build one and check
    touch = np.array([[0, 1, 0, 1], [1, 2, 0, 1]], np.float32)
    assert len(orc.brute_pairs(touch)) == 1
    # dropping the y tests would add pairs
    xo = (x0[:, None] <= x1[None, :]) & (x0[None, :] <= x1[:, None])
    assert int(np.triu(xo[:2000, :2000], 1).sum()) > int(((i < 2000) & (j < 2000)).sum())


def test_brute_count_pinned(orc):
    """orc_brute2f_count (the timed brute-force checker) against closed forms,
    the golden 16k count and the pair list it counts."""
    from pbig_amd import workloads

    assert orc.brute_count(np.tile(box(1, 2, 1, 2), (37, 1))) == 37 * 36 // 2
    assert orc.brute_count(np.stack([box(i, i + 1, 0, 1) for i in range(20)])) == 19
    g = [box(3 * i, 3 * i + 1, 3 * j, 3 * j + 1) for i in range(10) for j in range(10)]
    assert orc.brute_count(np.stack(g)) == 0
    assert orc.brute_count(np.zeros((0, 4), np.float32)) == 0
    rng = np.random.default_rng(11)
    for n, L, smax in ((700, 64.0, 6.0), (1500, 300.0, 40.0)):
        b = _random_boxes(rng, n, L, 0.5, smax)
        want = len(orc.brute_pairs(b))
        assert orc.brute_count(b, threads=1) == want
        assert orc.brute_count(b, threads=4) == want
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        man = json.load(f)["c1_uniform_16k"]
    b1, _ = workloads.c1_small_uniform()
    assert orc.brute_count(b1) == man["pairs"] == 10409


def _touch_counts(b, L, cs):
    """Cell membership counts by geometry, independent of the floor map: with
    nrc = ceil(L / cs) columns, column c covers [c cs, (c+1) cs), the first one
    extended to -inf and the last to +inf (clamping); a closed box [x0, x1]
    touches column c iff x0 < hi(c) and x1 >= lo(c).  Row-major, nrc*cy + cx."""
    nrc = max(1, int(np.ceil(L / cs)))
    lo = np.arange(nrc, dtype=np.float64) * cs
    hi = lo + cs
    lo[0], hi[-1] = -np.inf, np.inf
    x = (b[:, 0:1] < hi[None, :]) & (b[:, 1:2] >= lo[None, :])
    y = (b[:, 2:3] < hi[None, :]) & (b[:, 3:4] >= lo[None, :])
    return (y.T.astype(np.int64) @ x.astype(np.int64)).ravel(), nrc


@pytest.mark.parametrize("case", range(4))
def test_grid_stats_and_counts_pinned(orc, case):
    """orc_grid2f_counts and orc_grid2f_stats against the geometric count:
    integer-valued boxes land on cell edges (a box ending exactly on an edge
    touches the next cell), some boxes stick out of the domain on every side
    (clamped into the border cells), and L is not a multiple of the cell."""
    rng = np.random.default_rng(100 + case)
    L, cs, n = [(100.0, 4, 900), (1000.0, 64, 3000), (257.0, 16, 2000), (50.0, 8, 400)][case]
    c = rng.integers(-8, int(L) + 8, size=(n, 2)).astype(np.float32)
    w = rng.integers(0, [3, 90, 20, 30][case], size=(n, 2)).astype(np.float32)
    b = np.stack([c[:, 0], c[:, 0] + w[:, 0], c[:, 1], c[:, 1] + w[:, 1]], 1).astype(np.float32)
    b[: n // 10] = np.tile(box(5, 6, 5, 6), (n // 10, 1))  # one crowded cell
    want, nrc = _touch_counts(b, L, cs)
    got = orc.grid_counts(b, L, cs)
    assert got.shape == (nrc * nrc,)
    assert np.array_equal(got.astype(np.int64), want)
    st = orc.grid_stats(b, L, cs)
    assert st["n_cells"] == nrc * nrc
    assert st["memberships"] == int(want.sum())
    assert st["max_per_cell"] == int(want.max()) and st["max_cell"] == int(np.argmax(want))
    assert st["nonempty_cells"] == int((want > 0).sum())
    k = want[want > 1]
    assert st["compared_pairs"] == int((k * k // 2).sum())
    assert st["jobs256"] == int(((k + 255) // 256).sum())
    # the grid pipeline reports the same binning statistics with its pairs
    pairs, st2 = orc.grid_pairs(b, L, cs, with_stats=True)
    assert st2["pairs"] == len(pairs)
    assert {k2: v for k2, v in st2.items() if k2 != "pairs"} == {k2: v for k2, v in st.items() if k2 != "pairs"}


def _extent_boxes(ex, ey, n):
    """n boxes whose first len(ex) x/y extents are given, spread over [0, 1024)."""
    b = np.zeros((n, 4), np.float32)
    xs = (np.arange(n) * 37 % 1000).astype(np.float32)
    ys = (np.arange(n) * 91 % 1000).astype(np.float32)
    ex = np.resize(np.asarray(ex, np.float32), n)
    ey = np.resize(np.asarray(ey, np.float32), n)
    b[:, 0], b[:, 1], b[:, 2], b[:, 3] = xs, xs + ex, ys, ys + ey
    return b


def _rule_with(b, n, range_, accumulate):
    """The cell-size loop of pbig.cu:1316-1331 with the mean extent given by
    `accumulate` (for planting mistakes in the mean)."""
    x, y = accumulate(b[:1000, 1] - b[:1000, 0]), accumulate(b[:1000, 3] - b[:1000, 2])
    cs = 1
    while cs * 2 < range_:
        cs *= 2
    while True:
        nrc = int(np.ceil(range_ / cs))
        nmr = float(np.float32((1 + np.float32(x) / np.float32(cs)) * (1 + np.float32(y) / np.float32(cs))))
        per_cell = int(nmr * n / (nrc * nrc))
        if cs < x or per_cell < 32 or nmr > 4 or cs == 1:
            break
        cs //= 2
    return cs * 2


def test_decide_cell_size_int_truncated_mean(orc):
    """pbig.cu:1303-1308 accumulates the extents with an int init value
    (each step truncates) and divides by the sample count in integers.  On
    these inputs a float mean, or an int sum with a float division, crosses
    the cell side 64 and gives 128; the reference's arithmetic gives 64."""
    import pbig_amd

    n = 4096
    # x: 999 extents of 65.5 (each step truncates: +65) and one of 64: int sum
    # 64,999 -> 64 by integer division, float mean 65.4985, float division
    # 64.999; the loop stops where cs < x (pbig.cu:1322).
    # y (x extents 0): 999 of 193.5 and one of 192 -> 192 against 193.4985 or
    # 192.999; at cs 64, nmr = 1 + y/64 is exactly 4 (go on) or above (stop).
    ex = [65.5] * 999 + [64.0]
    ey = [193.5] * 999 + [192.0]
    for bx, by in ((ex, [8.0]), ([0.0], ey)):
        b = _extent_boxes(bx, by, n)
        assert orc.decide_cell_size(b, 1024.0) == 64
        assert pbig_amd.reference_cell_size(b, 1024.0) == 64

        def int_trunc(e):
            acc = 0
            for v in e:
                acc = int(np.float32(acc) + np.float32(v))
            return float(acc // len(e))

        assert _rule_with(b, n, 1024.0, int_trunc) == 64
        assert _rule_with(b, n, 1024.0, lambda e: float(np.mean(e.astype(np.float64)))) == 128
        assert _rule_with(b, n, 1024.0, lambda e: sum(int(np.float32(v)) for v in e) / len(e)) == 128
