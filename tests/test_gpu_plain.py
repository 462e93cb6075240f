"""GPU parity of the plain path (HIP kernels through the C-ABI) against the
CPU oracle.  Pair sets are compared sorted and must be bit-identical."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


PATHS = {"general": 1, "binned": 2, "auto": 0}


@pytest.fixture(scope="module", params=["general", "auto"])
def eng(torch_cuda, request):
    """An engine per implementation: the single-level general path, and the
    default (binned, falling back to general when a bin overflows LDS)."""
    import pbig_amd

    e = pbig_amd.Engine(0)
    e.configure(PATHS[request.param])
    e.path_name = request.param
    yield e
    e.close()


@pytest.fixture(scope="module")
def binned(torch_cuda):
    """An engine forced onto the binned path (errors instead of falling back)."""
    import pbig_amd

    e = pbig_amd.Engine(0)
    e.configure(PATHS["binned"])
    yield e
    e.close()


@pytest.fixture(scope="module")
def manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)


def run_device(torch, eng, boxes, L, cs, cap=None):
    import oracle

    d = torch.from_numpy(np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 4)).to("cuda:0")
    cap = cap if cap is not None else max(1024, 4 * boxes.shape[0])
    out = torch.empty((cap, 2), dtype=torch.int32, device="cuda:0")
    n = eng.check_device(d, L, cs, out, allow_overflow=True)
    if n > cap:  # rerun with room for every pair
        out = torch.empty((n, 2), dtype=torch.int32, device="cuda:0")
        assert eng.check_device(d, L, cs, out) == n
    return oracle.sort_pairs(out[:n].cpu().numpy().view(np.uint32))


def general_grid_stats(orc, st, boxes, cs, m):
    """The oracle's cell statistics of the grid the general path ran on: the
    whole grid's cells (the manifest's figures), or, when its cells were laid
    out from the bins, the binned grid, whose cells per row are rounded up to
    whole bins (boxes past the domain edge fold into its last row and
    column, not the domain's)."""
    if st["bins_per_row"] == 0:
        return m["grid_stats_at_reference_cell_size"]
    return orc.grid_stats(boxes, st["cells_per_row"] * cs, cs)


def test_c1_matches_golden_and_oracle(torch_cuda, eng, orc, manifest):
    from pbig_amd import workloads

    boxes, L = workloads.c1_small_uniform()
    m = manifest["c1_uniform_16k"]
    assert [float(v) for v in boxes[0]] == m["first_box"]
    cs = m["reference_cell_size"]
    got = run_device(torch_cuda, eng, boxes, L, cs)
    golden = np.load(os.path.join(GOLDEN, "c1_pairs.npy"))
    assert np.array_equal(got, golden)
    assert orc.pair_hash(got) == m["sha256"]
    st = eng.stats()
    if st["path"] == 1:
        ref = general_grid_stats(orc, st, boxes, cs, m)
        assert st["memberships"] == ref["memberships"]
        assert st["max_per_cell"] == ref["max_per_cell"]
    # cell_size = 0 selects the reference rule on the device boxes
    got0 = run_device(torch_cuda, eng, boxes, L, 0)
    assert np.array_equal(got0, golden)
    assert eng.stats()["cell_size"] == cs


@pytest.mark.parametrize("cs", [8, 16, 32, 64, 128, 256, 1024, 12.5])
def test_c1_any_cell_size(torch_cuda, eng, orc, cs):
    from pbig_amd import workloads

    boxes, L = workloads.c1_small_uniform()
    got = run_device(torch_cuda, eng, boxes, L, cs)
    assert np.array_equal(got, np.load(os.path.join(GOLDEN, "c1_pairs.npy")))


def test_c2_full_size(torch_cuda, eng, orc, manifest):
    from pbig_amd import workloads

    boxes, L = workloads.c2_large_uniform()
    m = manifest["c2_uniform_4m"]
    assert [float(v) for v in boxes[0]] == m["first_box"]
    got = run_device(torch_cuda, eng, boxes, L, m["reference_cell_size"], cap=4 * m["pairs"])
    assert len(got) == m["pairs"]
    assert orc.pair_hash(got) == m["sha256"]
    st = eng.stats()
    if st["path"] == 1:
        ref = general_grid_stats(orc, st, boxes, m["reference_cell_size"], m)
        assert st["memberships"] == ref["memberships"]


def test_c3_full_size(torch_cuda, eng, orc, manifest):
    from pbig_amd import workloads

    boxes, L = workloads.c3_clustered()
    m = manifest["c3_clustered_2m"]
    assert [float(v) for v in boxes[0]] == m["first_box"]
    got = run_device(torch_cuda, eng, boxes, L, m["reference_cell_size"], cap=4 * m["pairs"])
    assert len(got) == m["pairs"]
    assert orc.pair_hash(got) == m["sha256"]
    st = eng.stats()
    if st["path"] == 1:
        assert st["max_per_cell"] == m["grid_stats_at_reference_cell_size"]["max_per_cell"]


def edge_cases():
    rng = np.random.default_rng(7)
    cases = {}
    cases["empty"] = (np.zeros((0, 4), np.float32), 64.0, 16)
    cases["single"] = (np.array([[1, 2, 1, 2]], np.float32), 64.0, 16)
    # closed intervals: edge and corner contact count as overlap
    cases["touching"] = (np.array([[0, 1, 0, 1], [1, 2, 0, 1], [1, 2, 1, 2], [2, 3, 2, 3],
                                   [3.0000002, 4, 3, 4]], np.float32), 8.0, 1)
    # identical boxes: one dense cell, n(n-1)/2 pairs, several 64-box tiles
    cases["identical_300"] = (np.tile(np.array([[5, 6, 5, 6]], np.float32), (300, 1)), 64.0, 16)
    # tile boundaries inside one cell
    for k in (63, 64, 65, 127, 128, 129, 191, 257):
        c = rng.uniform(1, 15, size=(k, 2)).astype(np.float32)
        b = np.stack([c[:, 0] - 0.6, c[:, 0] + 0.6, c[:, 1] - 0.6, c[:, 1] + 0.6], 1).astype(np.float32)
        cases[f"one_cell_{k}"] = (b, 64.0, 16)
    # coordinates outside [0, range) and negative: fold into border cells
    c = rng.uniform(-40, 104, size=(3000, 2)).astype(np.float32)
    w = rng.uniform(0.5, 6, size=(3000, 2)).astype(np.float32)
    cases["outside_domain"] = (np.stack([c[:, 0], c[:, 0] + w[:, 0], c[:, 1], c[:, 1] + w[:, 1]], 1), 64.0, 8)
    # large boxes spanning many cells among small ones
    c = rng.uniform(0, 512, size=(4000, 2)).astype(np.float32)
    w = np.where(rng.uniform(size=(4000, 2)) < 0.02, rng.uniform(50, 300, size=(4000, 2)),
                 rng.uniform(0.5, 4, size=(4000, 2))).astype(np.float32)
    cases["mixed_sizes"] = (np.stack([c[:, 0], c[:, 0] + w[:, 0], c[:, 1], c[:, 1] + w[:, 1]], 1), 512.0, 16)
    # degenerate (zero-area) boxes and -0.0
    cases["degenerate"] = (np.array([[0, 0, 0, 0], [-0.0, 0, -0.0, 0], [0, 1, 0, 0], [0, 0, 0, 1],
                                     [1, 1, 1, 1]], np.float32), 4.0, 1)
    # subnormal coordinates
    tiny = np.float32(1e-40)
    cases["subnormal"] = (np.array([[tiny, 2 * tiny, 0, 1], [2 * tiny, 3 * tiny, 0, 1],
                                    [3 * tiny, 4 * tiny, 0, 1], [-tiny, 0, 0, 1]], np.float32), 4.0, 1)
    # infinite extents
    cases["infinite"] = (np.array([[-np.inf, np.inf, 0, 1], [3, 4, 0.5, 0.7], [10, 11, 5, 6]], np.float32),
                         16.0, 4)
    return cases


EDGE = edge_cases()


@pytest.mark.parametrize("name", sorted(EDGE))
def test_edge_cases_vs_brute_force(torch_cuda, eng, orc, name):
    boxes, L, cs = EDGE[name]
    want = orc.sort_pairs(orc.brute_pairs(boxes)) if len(boxes) else np.zeros((0, 2), np.uint32)
    got = run_device(torch_cuda, eng, boxes, L, cs) if len(boxes) else np.zeros((0, 2), np.uint32)
    if len(boxes):
        assert np.array_equal(got, want), (name, len(got), len(want))


def test_empty_input(torch_cuda, eng):
    d = torch_cuda.empty((0, 4), dtype=torch_cuda.float32, device="cuda:0")
    out = torch_cuda.empty((16, 2), dtype=torch_cuda.int32, device="cuda:0")
    assert eng.check_device(d, 64.0, 16, out) == 0


def test_invalid_box_is_rejected(torch_cuda, eng):
    import pbig_amd

    boxes = np.array([[0, 1, 0, 1], [2, 1, 0, 1]], np.float32)
    with pytest.raises(pbig_amd.PbigError) as e:
        run_device(torch_cuda, eng, boxes, 8.0, 1)
    assert e.value.status == pbig_amd.PBIG_E_INVALID_BOX
    boxes = np.array([[0, 1, 0, 1], [0, 1, np.nan, 1]], np.float32)
    with pytest.raises(pbig_amd.PbigError) as e:
        run_device(torch_cuda, eng, boxes, 8.0, 1)
    assert e.value.status == pbig_amd.PBIG_E_INVALID_BOX
    # the engine stays usable after an error
    assert len(run_device(torch_cuda, eng, np.array([[0, 1, 0, 1], [1, 2, 1, 2]], np.float32), 8.0, 1)) == 1


def test_overflow_reports_true_count(torch_cuda, eng):
    import pbig_amd

    boxes = np.tile(np.array([[5, 6, 5, 6]], np.float32), (100, 1))
    d = torch_cuda.from_numpy(boxes).to("cuda:0")
    out = torch_cuda.empty((10, 2), dtype=torch_cuda.int32, device="cuda:0")
    with pytest.raises(pbig_amd.PbigError) as e:
        eng.check_device(d, 64.0, 16, out)
    assert e.value.status == pbig_amd.PBIG_E_OVERFLOW
    assert eng.check_device(d, 64.0, 16, out, allow_overflow=True) == 4950


def test_async_matches_sync(torch_cuda, eng, orc):
    from pbig_amd import workloads

    boxes, L = workloads.c1_small_uniform()
    d = torch_cuda.from_numpy(boxes).to("cuda:0")
    out = torch_cuda.empty((65536, 2), dtype=torch_cuda.int32, device="cuda:0")
    res = torch_cuda.zeros(2, dtype=torch_cuda.int64, device="cuda:0")
    eng.check_device(d, L, 64, out)  # sizes the workspace
    for _ in range(3):
        eng.check_device_async(d, L, 64, out, res)
    torch_cuda.cuda.synchronize()
    n, err = [int(v) for v in res.cpu()]
    assert err == 0
    got = orc.sort_pairs(out[:n].cpu().numpy().view(np.uint32))
    assert np.array_equal(got, np.load(os.path.join(GOLDEN, "c1_pairs.npy")))


def test_host_entry_point_rectinfo(torch_cuda, eng, orc):
    import pbig_amd
    from pbig_amd import workloads

    boxes, L = workloads.c1_small_uniform()
    got = orc.sort_pairs(eng.check_host(pbig_amd.as_rects(boxes), L, 0))
    assert np.array_equal(got, np.load(os.path.join(GOLDEN, "c1_pairs.npy")))
    bad = pbig_amd.as_rects(boxes[:10])
    bad["rid"][3] = 7
    with pytest.raises(pbig_amd.PbigError) as e:
        eng.check_host(bad, L, 0)
    assert e.value.status == pbig_amd.PBIG_E_ARG


def test_host_entry_points_in_chunks(torch_cuda, eng, orc):
    """The host entry points copy the pairs back in chunks of 8 Mi pairs
    through two pinned buffers: 4,600 identical boxes (10,577,700 pairs, two
    chunks) beside 20,000 uniform ones.  The pairs spread into the caller's
    buffer, and the pairs handed chunk by chunk to a callback, both equal the
    oracle's set."""
    import pbig_amd
    from pbig_amd import workloads

    stack = np.tile(np.array([[3000.0, 3001.0, 3000.0, 3001.0]], np.float32), (4600, 1))
    rest = workloads.uniform(20000, 2048.0, 1.0, 8.0, 55)
    boxes = np.concatenate([rest[:9000], stack, rest[9000:]]).astype(np.float32)
    want = orc.sort_pairs(orc.sweep_pairs(boxes))
    assert want.shape[0] > (8 << 20)
    got = orc.sort_pairs(eng.check_host(pbig_amd.as_rects(boxes), 4096.0, 16.0, capacity=want.shape[0]))
    assert np.array_equal(got, want)
    chunks = []
    n = eng.check_host_each(pbig_amd.as_rects(boxes), 4096.0, 16.0, lambda p: chunks.append(p.copy()))
    assert n == want.shape[0] and len(chunks) == 2
    assert np.array_equal(orc.sort_pairs(np.concatenate(chunks)), want)


def test_host_entry_point_reports_the_first_bad_record(torch_cuda, eng):
    """Records are packed over several host threads; the error names the
    first bad record of the array, as a loop in order would."""
    import pbig_amd
    from pbig_amd import workloads

    boxes = workloads.uniform(300000, 8192.0, 1.0, 8.0, 56)
    r = pbig_amd.as_rects(boxes)
    r["xmin"][250000], r["xmax"][250000] = 5.0, 4.0  # inverted, later
    r["rid"][100000] = 7
    with pytest.raises(pbig_amd.PbigError) as e:
        eng.check_host(r, 8192.0, 16.0)
    assert e.value.status == pbig_amd.PBIG_E_ARG and "rect 100000 " in str(e.value)
    r["rid"][100000] = 100000
    with pytest.raises(pbig_amd.PbigError) as e:
        eng.check_host(r, 8192.0, 16.0)
    assert e.value.status == pbig_amd.PBIG_E_INVALID_BOX and "rect 250000 " in str(e.value)


def test_repeated_calls_shrinking_and_growing(torch_cuda, eng, orc):
    from pbig_amd import workloads

    for n, L, cs in ((50000, 2048.0, 32), (1000, 256.0, 16), (200000, 4096.0, 64), (10, 16.0, 4)):
        boxes = workloads.uniform(n, L, 1.0, 8.0, seed=n)
        got = run_device(torch_cuda, eng, boxes, L, cs)
        want = orc.sort_pairs(orc.grid_pairs(boxes, L, cs))
        assert np.array_equal(got, want), n


# ---------------------------------------------------------------------------
# Binned path forced on: it must either produce the exact set or refuse.

@pytest.mark.parametrize("cs", [16, 32, 64, 128, 256])
@pytest.mark.parametrize("bin_target", [300, 1500, 3000])
def test_binned_c1_cell_sizes_and_bin_sizes(torch_cuda, binned, orc, cs, bin_target):
    import pbig_amd
    from pbig_amd import workloads

    binned.configure(PATHS["binned"], bin_target)
    boxes, L = workloads.c1_small_uniform()
    try:
        got = run_device(torch_cuda, binned, boxes, L, cs)
    except pbig_amd.PbigError as e:
        # too dense for LDS, or more coarse bins than the partition supports
        assert e.status in (pbig_amd.PBIG_E_TOO_DENSE, pbig_amd.PBIG_E_ARG)
        return
    assert binned.stats()["path"] == PATHS["binned"]
    assert np.array_equal(got, np.load(os.path.join(GOLDEN, "c1_pairs.npy")))


@pytest.mark.parametrize("cs,bin_target", [(32, 2000), (64, 2000), (128, 2000), (32, 4096)])
def test_binned_c2_full_size(torch_cuda, binned, orc, manifest, cs, bin_target):
    from pbig_amd import workloads

    binned.configure(PATHS["binned"], bin_target)
    boxes, L = workloads.c2_large_uniform()
    m = manifest["c2_uniform_4m"]
    got = run_device(torch_cuda, binned, boxes, L, cs, cap=4 * m["pairs"])
    assert binned.stats()["path"] == PATHS["binned"]
    assert len(got) == m["pairs"]
    assert orc.pair_hash(got) == m["sha256"]


@pytest.mark.parametrize("name", sorted(EDGE))
def test_binned_edge_cases(torch_cuda, binned, orc, name):
    import pbig_amd

    binned.configure(PATHS["binned"], 2000)
    boxes, L, cs = EDGE[name]
    if not len(boxes):
        return
    want = orc.sort_pairs(orc.brute_pairs(boxes))
    try:
        got = run_device(torch_cuda, binned, boxes, L, cs)
    except pbig_amd.PbigError as e:
        assert e.status == pbig_amd.PBIG_E_TOO_DENSE, name
        return
    assert np.array_equal(got, want), (name, len(got), len(want))


def test_dense_spot_resolved_on_binned_path(torch_cuda, orc):
    """6000 boxes in a 2.5 x 2.5 spot of a 16-unit cell: the coarse bin and
    the fine cell overflow LDS; re-binning with finer cells resolves them."""
    import pbig_amd

    rng = np.random.default_rng(11)
    c = rng.uniform(10, 12, size=(6000, 2)).astype(np.float32)
    boxes = np.stack([c[:, 0], c[:, 0] + 0.5, c[:, 1], c[:, 1] + 0.5], 1).astype(np.float32)
    with pbig_amd.Engine(0) as e:
        e.configure(PATHS["binned"])  # the automatic choice sends a spot this dense to the general path
        got = run_device(torch_cuda, e, boxes, 64.0, 16)
        assert e.stats()["path"] == PATHS["binned"]
    assert np.array_equal(got, orc.sort_pairs(orc.brute_pairs(boxes)))


def test_auto_falls_back_on_unsplittable_stack(torch_cuda, orc):
    """6000 identical boxes share every cell at any refinement, more than a
    bin holds under either test schedule (4,096 general, 5,120 narrow): AUTO
    must return the exact set through the general path."""
    import pbig_amd

    boxes = np.tile(np.array([[20.0, 20.5, 30.0, 30.25]], np.float32), (6000, 1))
    boxes = np.concatenate([boxes, np.array([[1, 2, 1, 2], [20.5, 21, 30, 31]], np.float32)])
    with pbig_amd.Engine(0) as e:
        got = run_device(torch_cuda, e, boxes, 64.0, 16)
        assert e.stats()["path"] == PATHS["general"]
    assert np.array_equal(got, orc.sort_pairs(orc.brute_pairs(boxes)))


@pytest.mark.parametrize("cs", [16, 45, 64, 128])
def test_binned_c3_rebins_heavy_bins(torch_cuda, binned, orc, manifest, cs):
    """The clustered instance: coarse bins in the clusters exceed LDS and are
    re-binned one level down (recursively); the binned path must still give
    the exact set, without the general path."""
    from pbig_amd import workloads

    binned.configure(PATHS["binned"], 2000)
    boxes, L = workloads.c3_clustered()
    m = manifest["c3_clustered_2m"]
    got = run_device(torch_cuda, binned, boxes, L, cs, cap=4 * m["pairs"])
    assert binned.stats()["path"] == PATHS["binned"]
    assert len(got) == m["pairs"]
    assert orc.pair_hash(got) == m["sha256"]


def _dense_instance_settles(torch_cuda, orc, manifest, path, first_errors, levels_min=0):
    import pbig_amd
    from pbig_amd import workloads

    boxes, L = workloads.c3_clustered()
    m = manifest["c3_clustered_2m"]
    d = torch_cuda.from_numpy(boxes).to("cuda:0")
    out = torch_cuda.empty((m["pairs"] + 1024, 2), dtype=torch_cuda.int32, device="cuda:0")
    res = torch_cuda.zeros(2, dtype=torch_cuda.int64, device="cuda:0")
    with pbig_amd.Engine(0) as e:
        e.configure(PATHS[path], 2000, levels_min)
        e.check_device_async(d, L, 64.0, out, res)
        torch_cuda.cuda.synchronize()
        n_async, err = (int(v) for v in res.cpu().tolist())
        assert err in first_errors and n_async == 0, (n_async, err)
        n = e.check_device(d, L, 64.0, out)
        settled = e.stats()["path"]
        assert n == m["pairs"] and orc.pair_hash(orc.sort_pairs(out[:n].cpu().numpy().view(np.uint32))) == m["sha256"]
        for rep in range(2):
            res.zero_()
            out.fill_(-1)
            e.check_device_async(d, L, 64.0, out, res)
            torch_cuda.cuda.synchronize()
            assert res.cpu().tolist() == [m["pairs"], 0], rep
            assert orc.pair_hash(orc.sort_pairs(out[:n].cpu().numpy().view(np.uint32))) == m["sha256"], rep
    return settled


@pytest.mark.parametrize("levels_min", [0, 1 << 15])
def test_auto_dense_instance_settles(torch_cuda, orc, manifest, levels_min):
    """The clustered instance under the automatic choice: the first
    asynchronous call tries the single-pass launch, whose dense bins outgrow
    their fixed share of the bucket; it stops with error bit 5 and no pairs,
    and with bit 4 too, since the bins' totals (counted to the end) put over
    a quarter of the memberships in bins beyond LDS (mostly dense clusters).
    The synchronous call then sends the instance (2^21 boxes) to the general
    path at the default levels_min_boxes (2^20: mostly dense instances take
    the device levels from 64 times that), or to the device levels when 64
    times levels_min_boxes is within its size, and the next asynchronous
    calls follow it, with the exact set."""
    settled = _dense_instance_settles(torch_cuda, orc, manifest, "auto", (16, 32, 48), levels_min)
    assert settled == (PATHS["general"] if levels_min == 0 else PATHS["binned"])


def test_binned_dense_instance_on_device_levels(torch_cuda, orc, manifest):
    """The same instance on the binned path forced: the first asynchronous
    call reports its dense bins (error bit 5); the synchronous call runs the
    two-pass launch followed by re-binning levels planned on the device (the
    bins beyond LDS split into finer ones with no host work between levels),
    and the next asynchronous calls run that whole sequence on the device,
    with the exact set."""
    assert _dense_instance_settles(torch_cuda, orc, manifest, "binned", (8, 32)) == PATHS["binned"]


def test_binned_deep_rebinning(torch_cuda, binned, orc):
    """A dense core inside a sparse domain needs several re-binning levels."""
    from pbig_amd import workloads

    binned.configure(PATHS["binned"], 2000)
    core = workloads.uniform(60000, 400.0, 0.5, 2.0, 21) + np.float32(3000.0)
    rest = workloads.uniform(20000, 8192.0, 1.0, 8.0, 22)
    boxes = np.concatenate([core, rest]).astype(np.float32)
    got = run_device(torch_cuda, binned, boxes, 8192.0, 4.0)
    assert binned.stats()["path"] == PATHS["binned"]
    assert np.array_equal(got, orc.sort_pairs(orc.sweep_pairs(boxes)))


def _wide_share(boxes: np.ndarray, cs: float) -> float:
    """Share of boxes spanning more than two cells of side cs on an axis
    (the narrow test schedule keeps a launch up to one in 1,024)."""
    inv = np.float32(1.0) / np.float32(cs)
    f = np.floor(boxes * inv)
    return float(np.mean((f[:, 1] - f[:, 0] > 1) | (f[:, 3] - f[:, 2] > 1)))


def _lattice_wide_boxes(n: int, L: float, cs: float, seed: int) -> np.ndarray:
    """n boxes 2 to 40 cells wide on a quarter-cell lattice (edges on cell
    lines, touching the lattice boxes' edges), some of them flat."""
    rng = np.random.default_rng(seed)
    q = cs / 4
    x0 = rng.integers(-8, int(L / q), n) * q
    y0 = rng.integers(-8, int(L / q), n) * q
    w = rng.integers(8, 160, n) * q
    h = np.where(rng.random(n) < 0.2, 0.0, rng.integers(1, 160, n) * q)
    sw = rng.random(n) < 0.5  # which side is the wide one
    ww, hh = np.where(sw, w, h), np.where(sw, h, w)
    return np.stack([x0, x0 + ww, y0, y0 + hh], 1).astype(np.float32)


@pytest.mark.parametrize("bin_target", [200, 3500])
def test_narrow_kernel_on_a_touching_lattice(torch_cuda, binned, orc, bin_target):
    """Boxes exactly one cell wide on a half-cell lattice: every box spans two
    cells per axis, neighbours share edges and corners (closed intervals), and
    many pairs have their reference cell on a bin edge or in the halo.  The
    narrow-box test schedule runs; then one box two cells wide is added; then,
    on a lattice four times larger, 50 wide boxes (under one in 1,024: the
    narrow schedule's wide-box pass, with wide boxes crossing many bins,
    starting west of and south of the domain, and meeting each other with
    their reference cell in another bin); then, on the first lattice, enough
    wide boxes to send the launch to the general schedule.  Every launch must
    give the oracle's set."""
    cs = 8.0

    def lattice(side):
        g = np.arange(0, side, dtype=np.float32) * (cs / 2)
        x, y = np.meshgrid(g, g, indexing="ij")
        x, y = x.ravel(), y.ravel()
        return np.stack([x, x + cs, y, y + cs], 1).astype(np.float32), float(g[-1] + cs)

    boxes, L = lattice(120)
    big, L_big = lattice(240)
    binned.configure(PATHS["binned"], bin_target)
    got = run_device(torch_cuda, binned, boxes, L, cs)
    want = orc.sort_pairs(orc.brute_pairs(boxes))
    assert np.array_equal(got, want), (len(got), len(want))
    one = np.concatenate([boxes, np.array([[100.0, 100.0 + 2 * cs, 100.0, 101.0]], np.float32)])
    few = np.concatenate([big, _lattice_wide_boxes(50, L_big, cs, 5)])
    many = np.concatenate([boxes, _lattice_wide_boxes(3000, L, cs, 6)])
    assert _wide_share(one, cs) < 1 / 1024 and _wide_share(few, cs) < 1 / 1024 < _wide_share(many, cs)
    for case, side in ((one, L), (few, L_big), (many, L)):
        got = run_device(torch_cuda, binned, case, side, cs)
        want = orc.sort_pairs(orc.sweep_pairs(case))
        assert np.array_equal(got, want), (len(case), len(got), len(want))


@pytest.mark.parametrize("bin_target", [200, 3500])
def test_general_schedule_large_boxes_and_strips(torch_cuda, binned, orc, bin_target):
    """Under the general test schedule a box covering more than 16 fine cells
    of a bin is listed instead of replicated and tested against every
    membership of the bin (up to 256 listed per bin; the rest replicated),
    and bins beyond one strip lay out each strip by cell size.  30,000 small
    boxes in a 1,000-unit domain at cell 4 (dense cells, several strips per
    bin), 600 boxes 20 to 400 units wide (whole bins covered, listed boxes
    meeting each other with their reference cell in another bin, more than
    256 in some bins) and boxes west of and south of the domain: the pair
    set must be the oracle's."""
    from pbig_amd import workloads

    rng = np.random.default_rng(41)
    small = workloads.uniform(30000, 1000.0, 1.0, 6.0, 43)
    n = 1500
    x0 = rng.uniform(-50.0, 1000.0, n)
    y0 = rng.uniform(-50.0, 1000.0, n)
    w = rng.uniform(20.0, 400.0, n)
    h = rng.uniform(20.0, 400.0, n)
    large = np.stack([x0, x0 + w, y0, y0 + h], 1).astype(np.float32)
    boxes = np.concatenate([small, large]).astype(np.float32)
    binned.configure(PATHS["binned"], bin_target)
    got = run_device(torch_cuda, binned, boxes, 1000.0, 4.0)
    assert binned.stats()["path"] == PATHS["binned"]
    assert np.array_equal(got, orc.sort_pairs(orc.sweep_pairs(boxes)))


@pytest.mark.parametrize("bin_target", [16, 200, 3500, 5000])
def test_narrow_kernel_on_crowded_cells(torch_cuda, binned, orc, bin_target):
    """20,000 boxes of side up to one cell in a 256 x 256 domain at cell 16:
    about 78 boxes per cell, so every box has hundreds of partners in its own
    cell and the row below, and the pair stage fills and flushes many times
    inside one partner loop.  A bin target of 16 plans bins of one cell per
    side (the division by the bin side is then the identity); 5,000 plans
    bins near the narrow schedule's capacity, some over it (re-binned)."""
    rng = np.random.default_rng(11)
    n = 20000
    c = rng.uniform(0, 256, size=(n, 2)).astype(np.float32)
    w = rng.uniform(1, 16, size=(n, 2)).astype(np.float32)
    boxes = np.stack([c[:, 0] - w[:, 0] / 2, c[:, 0] + w[:, 0] / 2, c[:, 1] - w[:, 1] / 2, c[:, 1] + w[:, 1] / 2],
                     1).astype(np.float32)
    binned.configure(PATHS["binned"], bin_target)
    got = run_device(torch_cuda, binned, boxes, 256.0, 16.0)
    want = orc.sort_pairs(orc.brute_pairs(boxes))
    assert np.array_equal(got, want), (len(got), len(want))


def test_auto_keeps_large_narrow_bins_binned(torch_cuda, orc, manifest):
    """At cell 32 and bin target 4,096 the 4M instance's bins hold about 4,200
    memberships: over what the general test schedule takes (4,096), within
    the narrow-box schedule's 5,120.  The automatic choice must not count
    them as heavy; the pairs are the manifest's."""
    import pbig_amd
    from pbig_amd import workloads

    eng = pbig_amd.Engine(0)
    try:
        eng.configure(PATHS["auto"], 4096)
        boxes, L = workloads.c2_large_uniform()
        m = manifest["c2_uniform_4m"]
        got = run_device(torch_cuda, eng, boxes, L, 32.0, cap=4 * m["pairs"])
        assert eng.stats()["path"] == PATHS["binned"]
        assert len(got) == m["pairs"]
        assert orc.pair_hash(got) == m["sha256"]
    finally:
        eng.close()


@pytest.mark.parametrize("levels_min", [0, 1])
def test_unsplittable_bin(torch_cuda, orc, levels_min):
    """6,000 identical boxes among 20,000 uniform ones share one fine cell at
    every refinement; their bin is beyond LDS but holds under a quarter of
    the memberships.  At the default levels_min_boxes (2^20) the automatic
    choice sends this 26,000-box instance to the general path; at 1 the
    device levels cannot split the bin, so it is tested whole on the device
    (every pair of its memberships under the reference-cell rule) and the
    call stays on the binned path.  Either way the set is exact, and the
    asynchronous call after it follows the path that worked."""
    import pbig_amd
    from pbig_amd import workloads

    same = np.tile(np.array([[100.0, 101.0, 100.0, 101.0]], np.float32), (6000, 1))
    rest = workloads.uniform(20000, 4096.0, 1.0, 8.0, 77)
    boxes = np.concatenate([rest[:7000], same, rest[7000:]]).astype(np.float32)
    want = orc.sort_pairs(orc.sweep_pairs(boxes))
    d = torch_cuda.from_numpy(boxes).to("cuda:0")
    out = torch_cuda.empty((len(want) + 4096, 2), dtype=torch_cuda.int32, device="cuda:0")
    res = torch_cuda.zeros(2, dtype=torch_cuda.int64, device="cuda:0")
    with pbig_amd.Engine(0) as e:
        e.configure(PATHS["auto"], 2000, levels_min)
        n = e.check_device(d, 4096.0, 8.0, out)
        assert n == len(want)
        assert e.stats()["path"] == (PATHS["binned"] if levels_min == 1 else PATHS["general"])
        assert np.array_equal(orc.sort_pairs(out[:n].cpu().numpy().view(np.uint32)), want)
        e.check_device_async(d, 4096.0, 8.0, out, res)
        torch_cuda.cuda.synchronize()
        assert res.cpu().tolist() == [len(want), 0]
        assert np.array_equal(orc.sort_pairs(out[:n].cpu().numpy().view(np.uint32)), want)


@pytest.mark.parametrize("levels_min", [0, 1])
def test_auto_spot_in_uniform_instance(torch_cuda, orc, levels_min):
    """100,000 uniform boxes and a spot of 15,000 small ones: the spot's bin
    is beyond LDS but holds under a quarter of the memberships (not mostly
    dense clusters).  With levels_min_boxes above the instance's size the
    automatic choice sends it to the general path; at 1 the device levels
    split the spot's bin, the call stays on the binned path, and the
    asynchronous calls after it run the levels too.  The set is exact
    every time."""
    import pbig_amd
    from pbig_amd import workloads

    rest = workloads.uniform(100000, 8192.0, 1.0, 8.0, 91)
    spot = workloads.uniform(15000, 60.0, 0.5, 2.0, 92) + np.float32(4000.0)
    boxes = np.concatenate([rest[:50000], spot, rest[50000:]]).astype(np.float32)
    want = orc.sort_pairs(orc.sweep_pairs(boxes))
    d = torch_cuda.from_numpy(boxes).to("cuda:0")
    out = torch_cuda.empty((len(want) + 4096, 2), dtype=torch_cuda.int32, device="cuda:0")
    res = torch_cuda.zeros(2, dtype=torch_cuda.int64, device="cuda:0")
    with pbig_amd.Engine(0) as e:
        e.configure(PATHS["auto"], 2000, levels_min)
        n = e.check_device(d, 8192.0, 8.0, out)
        assert n == len(want)
        assert e.stats()["path"] == (PATHS["binned"] if levels_min == 1 else PATHS["general"])
        assert np.array_equal(orc.sort_pairs(out[:n].cpu().numpy().view(np.uint32)), want)
        for rep in range(2):
            res.zero_()
            out.fill_(-1)
            e.check_device_async(d, 8192.0, 8.0, out, res)
            torch_cuda.cuda.synchronize()
            assert res.cpu().tolist() == [len(want), 0], rep
            assert np.array_equal(orc.sort_pairs(out[:n].cpu().numpy().view(np.uint32)), want), rep


def test_general_cells_from_bins_dense_and_large(torch_cuda, orc):
    """The general path's cells laid out from the bins (cells per bin counted
    and placed in LDS): a bin holding one cell of 3,000 memberships (placed
    by the count kernel's own stage), the cluster's bins of more than eight
    placement windows (counted and placed in slices), boxes spanning hundreds
    of cells (the membership array grows and the call reruns), asynchronous
    calls after the synchronous one, all against sweep-and-prune; then a
    cell past the 262,144-box tile limit, which the call refuses."""
    import pbig_amd
    from pbig_amd import workloads

    stack = np.tile(np.array([[3000.0, 3001.0, 3000.0, 3001.0]], np.float32), (3000, 1))  # (its own bin)
    cluster = workloads.uniform(30000, 200.0, 1.0, 4.0, 93) + np.float32(200.0)
    background = workloads.uniform(20000, 4096.0, 1.0, 8.0, 94)
    rng = np.random.default_rng(95)
    c = rng.uniform(0, 4096, size=(50, 2)).astype(np.float32)
    w = rng.uniform(100, 400, size=(50, 2)).astype(np.float32)
    wide = np.stack([c[:, 0], c[:, 0] + w[:, 0], c[:, 1], c[:, 1] + w[:, 1]], 1).astype(np.float32)
    boxes = np.concatenate([background[:10000], stack, wide, cluster, background[10000:]]).astype(np.float32)
    want = orc.sort_pairs(orc.sweep_pairs(boxes))
    d = torch_cuda.from_numpy(boxes).to("cuda:0")
    out = torch_cuda.empty((len(want) + 4096, 2), dtype=torch_cuda.int32, device="cuda:0")
    res = torch_cuda.zeros(2, dtype=torch_cuda.int64, device="cuda:0")
    with pbig_amd.Engine(0) as e:
        e.configure(PATHS["general"], 4096)
        n = e.check_device(d, 4096.0, 16.0, out)
        st = e.stats()
        assert st["path"] == PATHS["general"] and st["bins_per_row"] > 0  # the cells came from the bins
        assert st["max_per_cell"] >= 3000
        assert n == len(want)
        assert np.array_equal(orc.sort_pairs(out[:n].cpu().numpy().view(np.uint32)), want)
        for rep in range(2):
            res.zero_()
            out.fill_(-1)
            e.check_device_async(d, 4096.0, 16.0, out, res)
            torch_cuda.cuda.synchronize()
            assert res.cpu().tolist() == [len(want), 0], rep
            assert np.array_equal(orc.sort_pairs(out[:n].cpu().numpy().view(np.uint32)), want), rep
        over = torch_cuda.from_numpy(np.tile(np.array([[5.0, 6.0, 5.0, 6.0]], np.float32), (262145, 1))).to("cuda:0")
        with pytest.raises(pbig_amd.PbigError) as err:
            e.check_device(over, 64.0, 16.0, out)
        assert err.value.status == pbig_amd.PBIG_E_TOO_DENSE


@pytest.mark.parametrize("cell", [45.254833, 22.627417])
def test_general_clustered_sliced_layout(torch_cuda, orc, manifest, cell):
    """The clustered 2M instance on the general path at bin target 4,096: its
    cluster cores' bins (up to ~58,000 coarse memberships) are counted in
    slices by many workgroups and placed in slices; at cell 22.6 the grid has
    2,116 bins, so the count kernel's tickets run over three chunks of bins.
    The sorted set's hash matches the manifest, synchronously and in two
    asynchronous calls after it."""
    import pbig_amd
    from pbig_amd import workloads

    boxes, L = workloads.c3_clustered()
    m = manifest["c3_clustered_2m"]
    d = torch_cuda.from_numpy(boxes).to("cuda:0")
    out = torch_cuda.empty((m["pairs"] + 1024, 2), dtype=torch_cuda.int32, device="cuda:0")
    res = torch_cuda.zeros(2, dtype=torch_cuda.int64, device="cuda:0")
    with pbig_amd.Engine(0) as e:
        e.configure(PATHS["general"], 4096)
        n = e.check_device(d, L, cell, out)
        st = e.stats()
        assert st["path"] == PATHS["general"] and st["bins_per_row"] > 0
        if cell < 30.0:
            assert st["bins_per_row"] ** 2 > 2048  # more than two chunks of 1,024 bins
        assert n == m["pairs"]
        assert orc.pair_hash(orc.sort_pairs(out[:n].cpu().numpy().view(np.uint32))) == m["sha256"]
        for rep in range(2):
            res.zero_()
            out.fill_(-1)
            e.check_device_async(d, L, cell, out, res)
            torch_cuda.cuda.synchronize()
            assert res.cpu().tolist() == [m["pairs"], 0], rep
            assert orc.pair_hash(orc.sort_pairs(out[:n].cpu().numpy().view(np.uint32))) == m["sha256"], rep
