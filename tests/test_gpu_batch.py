"""GPU parity of batches: many instances in one launch sequence, each
instance's pairs compared (sorted, bit-identical) with the CPU oracle run on
that instance alone.  Covers the synchronous entry point, the asynchronous
launch, HIP-graph capture and replay, the general-path fallback for a crowded
instance inside a batch, empty and one-box instances and pair overflow."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def eng(torch_cuda):
    import pbig_amd

    e = pbig_amd.Engine(0)
    yield e
    e.close()


def concat(instances):
    """[(boxes, L)] -> (boxes, offsets, ranges)."""
    n = [b.shape[0] for b, _ in instances]
    offsets = np.zeros(len(n) + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(n)
    boxes = np.concatenate([b for b, _ in instances]).astype(np.float32) if offsets[-1] else \
        np.zeros((0, 4), np.float32)
    return boxes, offsets, np.array([L for _, L in instances], dtype=np.float64)


def expected(boxes, offsets):
    import oracle

    return [oracle.sort_pairs(oracle.sweep_pairs(boxes[offsets[i]:offsets[i + 1]]))
            for i in range(len(offsets) - 1)]


def split(pairs, io):
    import oracle

    return [oracle.sort_pairs(pairs[int(o):int(o) + int(c)]) for o, c in io]


def to_dev(torch, boxes):
    return torch.from_numpy(np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 4)).to("cuda:0")


def check_all(got, want):
    assert len(got) == len(want)
    for i, (g, w) in enumerate(zip(got, want)):
        assert np.array_equal(g, w), f"instance {i}: {g.shape[0]} pairs vs {w.shape[0]} expected"


def small_batch(seed=41, k=48, nmax=20000):
    from pbig_amd import workloads

    boxes, offsets, ranges = workloads.batched(k, seed, nmin=64, nmax=nmax)
    return boxes, offsets, ranges.astype(np.float64)


@pytest.mark.parametrize("cell", [8.0, 32.0, 100.0])
def test_batch_check_matches_oracle(torch_cuda, eng, cell):
    import pbig_amd

    boxes, offsets, ranges = small_batch()
    want = expected(boxes, offsets)
    total = sum(w.shape[0] for w in want)
    d = to_dev(torch_cuda, boxes)
    pairs = torch_cuda.empty((total + 1024, 2), dtype=torch_cuda.int32, device="cuda:0")
    with pbig_amd.Batch(eng, offsets, ranges, cell) as b:
        n, io = b.check(d, pairs)
        assert n == total
        assert int(io[-1, 0] + io[-1, 1]) == total
        assert np.array_equal(io[1:, 0], io[:-1, 0] + io[:-1, 1])  # instances in order, back to back
        check_all(split(pairs[:n].cpu().numpy().view(np.uint32), io), want)


def test_batch_alone_equals_in_mix(torch_cuda, eng):
    """An instance gives the same sorted pairs alone (plain path) and inside a batch."""
    import oracle
    import pbig_amd

    boxes, offsets, ranges = small_batch(seed=43, k=16)
    d = to_dev(torch_cuda, boxes)
    pairs = torch_cuda.empty((4 * boxes.shape[0], 2), dtype=torch_cuda.int32, device="cuda:0")
    with pbig_amd.Batch(eng, offsets, ranges, 32.0) as b:
        n, io = b.check(d, pairs)
        in_mix = split(pairs[:n].cpu().numpy().view(np.uint32), io)
    eng.configure(pbig_amd.PATH_BINNED)
    try:
        for i in range(16):
            sub = to_dev(torch_cuda, boxes[offsets[i]:offsets[i + 1]])
            out = torch_cuda.empty((4 * sub.shape[0] + 1024, 2), dtype=torch_cuda.int32, device="cuda:0")
            m = eng.check_device(sub, float(ranges[i]), 32.0, out)
            alone = oracle.sort_pairs(out[:m].cpu().numpy().view(np.uint32))
            assert np.array_equal(alone, in_mix[i]), i
    finally:
        eng.configure(pbig_amd.PATH_AUTO)


def test_batch_launch_and_graph_replay(torch_cuda, eng):
    import pbig_amd

    torch = torch_cuda
    boxes, offsets, ranges = small_batch(seed=47, k=64)
    want = expected(boxes, offsets)
    total = sum(w.shape[0] for w in want)
    d = to_dev(torch, boxes)
    cap = total + 4096
    with pbig_amd.Batch(eng, offsets, ranges, 32.0) as b:
        pairs = torch.empty((cap, 2), dtype=torch.int32, device="cuda:0")
        n, io = b.check(d, pairs)  # sizes the workspace
        assert n == total
        io_dev = torch.zeros((len(want), 2), dtype=torch.int64, device="cuda:0")
        res = torch.zeros(2, dtype=torch.int64, device="cuda:0")
        pairs.fill_(-1)
        b.launch(d, pairs, io_dev, res)
        torch.cuda.synchronize()
        assert res.cpu().tolist() == [total, 0]
        assert np.array_equal(io_dev.cpu().numpy().astype(np.uint64), io)
        check_all(split(pairs[:n].cpu().numpy().view(np.uint32), io), want)
        # graph: replayed several times, eager work and checks between replays
        b.capture(d, pairs, io_dev, res)
        for r in range(4):
            pairs.fill_(-1)
            io_dev.zero_()
            res.fill_(-7)
            torch.cuda.synchronize()
            b.replay()
            torch.cuda.synchronize()
            assert res.cpu().tolist() == [total, 0], r
            io_r = io_dev.cpu().numpy().astype(np.uint64)
            assert np.array_equal(io_r, io), r
            check_all(split(pairs[:n].cpu().numpy().view(np.uint32), io_r), want)


def test_batch_edge_instances(torch_cuda, eng):
    """Empty, one-box, two identical, touching-edge and all-overlapping instances."""
    import pbig_amd
    from pbig_amd import workloads

    inst = [
        (np.zeros((0, 4), np.float32), 64.0),
        (np.array([[1, 2, 1, 2]], np.float32), 64.0),
        (np.array([[5, 6, 5, 6], [5, 6, 5, 6]], np.float32), 64.0),
        (np.array([[0, 1, 0, 1], [1, 2, 1, 2], [2, 3, 0, 1]], np.float32), 4.0),  # shared edges/corners
        (workloads.uniform(2000, 256.0, 1.0, 8.0, 9), 256.0),
        (np.zeros((0, 4), np.float32), 10.0),
        (np.tile(np.array([[10, 300, 10, 300]], np.float32), (300, 1)), 512.0),  # 44,850 pairs
        (workloads.uniform(5000, 1024.0, 1.0, 64.0, 10), 1024.0),
    ]
    boxes, offsets, ranges = concat(inst)
    want = expected(boxes, offsets)
    total = sum(w.shape[0] for w in want)
    d = to_dev(torch_cuda, boxes)
    pairs = torch_cuda.empty((total + 16, 2), dtype=torch_cuda.int32, device="cuda:0")
    with pbig_amd.Batch(eng, offsets, ranges, 16.0) as b:
        n, io = b.check(d, pairs)
        assert n == total
        check_all(split(pairs[:n].cpu().numpy().view(np.uint32), io), want)


@pytest.mark.parametrize("host", [True, False])
def test_batch_crowded_instance(torch_cuda, eng, host):
    """A 6000-box stack in one fine cell overflows LDS at every refinement.
    With host re-binning that instance runs on the general path; with device
    levels its bin is tested whole on the device, and the batch's later
    launches complete there.  Its pairs land in its slot between the
    others."""
    import pbig_amd
    from pbig_amd import workloads

    stack = np.tile(np.array([[3, 4, 3, 4]], np.float32), (6000, 1))  # 17,997,000 pairs
    inst = [(workloads.uniform(3000, 512.0, 1.0, 8.0, 11), 512.0), (stack, 64.0),
            (workloads.uniform(4000, 512.0, 1.0, 8.0, 12), 512.0)]
    boxes, offsets, ranges = concat(inst)
    want = expected(boxes, offsets)
    total = sum(w.shape[0] for w in want)
    d = to_dev(torch_cuda, boxes)
    pairs = torch_cuda.empty((total, 2), dtype=torch_cuda.int32, device="cuda:0")
    eng.configure(pbig_amd.PATH_AUTO, host_levels=host)
    try:
        with pbig_amd.Batch(eng, offsets, ranges, 16.0) as b:
            n, io = b.check(d, pairs)
            assert n == total
            assert [int(c) for _, c in io] == [w.shape[0] for w in want]
            check_all(split(pairs[:n].cpu().numpy().view(np.uint32), io), want)
            res = torch_cuda.zeros(2, dtype=torch_cuda.int64, device="cuda:0")
            dio = torch_cuda.zeros((3, 2), dtype=torch_cuda.int64, device="cuda:0")
            pairs.fill_(-1)
            b.launch(d, pairs, dio, res)
            torch_cuda.cuda.synchronize()
            if host:  # the asynchronous launch reports the overflow instead
                assert res[1].item() & 8
            else:
                assert res.tolist() == [total, 0]
                check_all(split(pairs[:n].cpu().numpy().view(np.uint32), dio.cpu().numpy()), want)
    finally:
        eng.configure(pbig_amd.PATH_AUTO)


def test_batch_overflow_reports_counts(torch_cuda, eng):
    import pbig_amd

    boxes, offsets, ranges = small_batch(seed=53, k=8)
    want = expected(boxes, offsets)
    total = sum(w.shape[0] for w in want)
    d = to_dev(torch_cuda, boxes)
    pairs = torch_cuda.empty((total // 2, 2), dtype=torch_cuda.int32, device="cuda:0")
    with pbig_amd.Batch(eng, offsets, ranges, 32.0) as b:
        n, io = b.check(d, pairs, allow_overflow=True)
        assert n == total
        assert [int(c) for _, c in io] == [w.shape[0] for w in want]
        with pytest.raises(pbig_amd.PbigError) as e:
            b.check(d, pairs)
        assert e.value.status == pbig_amd.PBIG_E_OVERFLOW


def test_batch_rejects_bad_arguments(torch_cuda, eng):
    import pbig_amd

    boxes, offsets, ranges = small_batch(seed=59, k=4)
    d = to_dev(torch_cuda, boxes)
    pairs = torch_cuda.empty((16, 2), dtype=torch_cuda.int32, device="cuda:0")
    with pytest.raises(pbig_amd.PbigError):
        pbig_amd.Batch(eng, offsets, ranges, 0.0)  # cell size must be > 0
    with pbig_amd.Batch(eng, offsets, ranges, 32.0) as b:
        with pytest.raises(pbig_amd.PbigError):
            b.check(d[:-1], pairs)  # fewer boxes than the instances cover
    bad = boxes.copy()
    bad[offsets[2] + 3] = [5, 4, 0, 1]
    with pbig_amd.Batch(eng, offsets, ranges, 32.0) as b:
        with pytest.raises(pbig_amd.PbigError) as e:
            b.check(to_dev(torch_cuda, bad), pairs)
        assert e.value.status == pbig_amd.PBIG_E_INVALID_BOX


@pytest.mark.slow
def test_batch_c4_full(torch_cuda, eng):
    """The 4,096-instance batch (120M boxes) as the bench runs it: cell sizes
    from the autotuner's table, the table's bin target, one HIP-graph capture
    replayed twice (with other work between the replays); and once more
    through the synchronous entry point at cell 32.  Every instance's sorted
    pair set against the oracle's grid restatement, bit for bit."""
    import oracle
    import pbig_amd
    from pbig_amd import autotune, workloads

    boxes, offsets, ranges = workloads.batched(4096, workloads.SEEDS["c4"])
    table = autotune.load_table()
    cells = autotune.cell_sizes(offsets, ranges, table, boxes)
    eng.configure(pbig_amd.PATH_AUTO, int(table["bin_target"]))
    d = to_dev(torch_cuda, boxes)
    cap = int(0.2 * boxes.shape[0])
    pairs = torch_cuda.empty((cap, 2), dtype=torch_cuda.int32, device="cuda:0")
    io_dev = torch_cuda.zeros((4096, 2), dtype=torch_cuda.int64, device="cuda:0")
    res = torch_cuda.zeros(2, dtype=torch_cuda.int64, device="cuda:0")
    runs = []
    with pbig_amd.Batch(eng, offsets, ranges.astype(np.float64), cells) as b:
        n0, io0 = b.check(d, pairs)  # sizes the workspace
        b.capture(d, pairs, io_dev, res)
        for rep in range(2):
            pairs.fill_(-1)
            io_dev.zero_()
            torch_cuda.cuda.synchronize()
            b.replay()
            torch_cuda.cuda.synchronize()
            assert res.tolist() == [n0, 0], rep
            io = io_dev.cpu().numpy()
            assert np.array_equal(io.astype(np.uint64), io0)
            runs.append((pairs[:n0].cpu().numpy().view(np.uint32).copy(), io))
    eng.configure(pbig_amd.PATH_AUTO)
    with pbig_amd.Batch(eng, offsets, ranges.astype(np.float64), 32.0) as b:
        n, io = b.check(d, pairs)
        runs.append((pairs[:n].cpu().numpy().view(np.uint32).copy(), io))
    for i in range(4096):
        sub = boxes[offsets[i]:offsets[i + 1]]
        w = oracle.sort_pairs(oracle.grid_pairs(sub, float(ranges[i]), 32, threads=16))
        for k, (got, io) in enumerate(runs):
            g = oracle.sort_pairs(got[int(io[i, 0]):int(io[i, 0] + io[i, 1])])
            assert np.array_equal(g, w), (i, k)


@pytest.mark.parametrize("path", ["binned", "auto"])
def test_batch_clustered_instances_rebinned(torch_cuda, eng, path):
    """Clustered instances whose coarse bins exceed LDS, mixed with uniform
    ones: heavy bins are re-binned (by levels on the device) and their pairs
    placed in their instance's run.  Exact, and again on a second check."""
    import pbig_amd
    from pbig_amd import workloads

    eng.configure(pbig_amd.PATH_BINNED if path == "binned" else pbig_amd.PATH_AUTO)

    inst = []
    for t in range(10):
        n = [3000, 40000, 120000, 7000, 250000][t % 5]
        L = 65536.0 * float(np.sqrt(n / 2097152.0))
        if t % 2:
            inst.append((workloads.clustered(n, L, 100 + t), L))
        else:
            inst.append((workloads.uniform(n, L, 1.0, 32.0, 100 + t), L))
    boxes, offsets, ranges = concat(inst)
    want = expected(boxes, offsets)
    total = sum(w.shape[0] for w in want)
    d = to_dev(torch_cuda, boxes)
    pairs = torch_cuda.empty((total + 1024, 2), dtype=torch_cuda.int32, device="cuda:0")
    cells = [float(np.float32(L / np.sqrt(offsets[i + 1] - offsets[i]))) for i, (_, L) in enumerate(inst)]
    with pbig_amd.Batch(eng, offsets, ranges, cells) as b:
        n, io = b.check(d, pairs)
        assert n == total
        check_all(split(pairs[:n].cpu().numpy().view(np.uint32), io), want)
        # and again: the plan is restored after the splice
        pairs.fill_(-1)
        n2, io2 = b.check(d, pairs)
        assert n2 == total and np.array_equal(io2, io)
        check_all(split(pairs[:n].cpu().numpy().view(np.uint32), io2), want)


class host_levels:
    """The engine configured to re-bin bins beyond LDS on the host (no device
    levels) inside the block; the automatic defaults after it."""

    def __init__(self, eng):
        self.eng = eng

    def __enter__(self):
        import pbig_amd

        self.eng.configure(pbig_amd.PATH_AUTO, host_levels=True)

    def __exit__(self, *exc):
        import pbig_amd

        self.eng.configure(pbig_amd.PATH_AUTO)


def test_batch_finish_after_launch(torch_cuda, eng):
    """Host re-binning: an asynchronous launch that reports bins beyond LDS is
    completed by finish() from its own results (no second launch), exactly."""
    import pbig_amd
    from pbig_amd import workloads

    inst = [(workloads.clustered(150000, 17000.0, 300), 17000.0),
            (workloads.uniform(50000, 3000.0, 1.0, 8.0, 301), 3000.0)]
    boxes, offsets, ranges = concat(inst)
    want = expected(boxes, offsets)
    total = sum(w.shape[0] for w in want)
    d = to_dev(torch_cuda, boxes)
    pairs = torch_cuda.empty((total + 1024, 2), dtype=torch_cuda.int32, device="cuda:0")
    cells = [40.0, 13.0]
    with host_levels(eng), pbig_amd.Batch(eng, offsets, ranges, cells) as b:
        res = torch_cuda.zeros(2, dtype=torch_cuda.int64, device="cuda:0")
        b.check(d, pairs)  # sizes the workspace
        pairs.fill_(-1)
        b.launch(d, pairs, None, res)
        torch_cuda.cuda.synchronize()
        assert res[1].item() == 8  # bins beyond LDS, nothing else
        n, io = b.check(d, pairs, finish=True)
        assert n == total
        check_all(split(pairs[:n].cpu().numpy().view(np.uint32), io), want)


def mixed_batch(seed, k=10):
    """Clustered instances whose coarse bins exceed LDS among uniform ones."""
    from pbig_amd import workloads

    inst = []
    for t in range(k):
        n = [3000, 40000, 120000, 7000, 250000][t % 5]
        L = 65536.0 * float(np.sqrt(n / 2097152.0))
        if t % 2:
            inst.append((workloads.clustered(n, L, seed + t), L))
        else:
            inst.append((workloads.uniform(n, L, 1.0, 32.0, seed + t), L))
    boxes, offsets, ranges = concat(inst)
    cells = [float(np.float32(L / np.sqrt(offsets[i + 1] - offsets[i]))) for i, (_, L) in enumerate(inst)]
    return boxes, offsets, ranges, cells


def test_batch_levels_on_device(torch_cuda, eng):
    """Device levels for a batch: the first asynchronous launch reports bins
    beyond LDS; finish() runs the launch again with levels planned on the
    device over every such bin of the batch, and each instance's run holds
    its level-0 pairs followed by its levels' pairs.  The batch keeps the
    levels: its next launches complete on the device (error word 0), also
    captured in a graph and replayed (with eager work between replays), and
    unresolved() flags none.  Re-planned for other instances, its launch
    completes, or regrows through finish().  Exact every time."""
    import pbig_amd

    eng.configure(pbig_amd.PATH_AUTO)
    boxes, offsets, ranges, cells = mixed_batch(400)
    want = expected(boxes, offsets)
    total = sum(w.shape[0] for w in want)
    d = to_dev(torch_cuda, boxes)
    pairs = torch_cuda.empty((total + 1024, 2), dtype=torch_cuda.int32, device="cuda:0")
    k = len(want)
    io = torch_cuda.zeros((k, 2), dtype=torch_cuda.int64, device="cuda:0")
    res = torch_cuda.zeros(2, dtype=torch_cuda.int64, device="cuda:0")

    def check_launch(tag):
        torch_cuda.cuda.synchronize()
        assert res.tolist() == [total, 0], (tag, res.tolist())
        hio = io.cpu().numpy()
        check_all(split(pairs[:total].cpu().numpy().view(np.uint32), hio), want)

    with pbig_amd.Batch(eng, offsets, ranges, cells) as b:
        b.launch(d, pairs, io, res)
        torch_cuda.cuda.synchronize()
        assert res[1].item() & 8, res.tolist()  # bins beyond LDS (and perhaps a workspace to grow)
        n, hio = b.check(d, pairs, finish=True)
        assert n == total
        check_all(split(pairs[:n].cpu().numpy().view(np.uint32), hio), want)
        for rep in range(2):
            pairs.fill_(-1)
            io.zero_()
            b.launch(d, pairs, io, res)
            check_launch(f"launch {rep}")
        assert not b.unresolved().any()
        b.capture(d, pairs, io, res)
        for rep in range(3):
            pairs.fill_(-1)
            io.zero_()
            res.zero_()
            torch_cuda.cuda.synchronize()
            b.replay()
            check_launch(f"replay {rep}")
            torch_cuda.ones(1 << 20, device="cuda:0").sum().item()  # eager work between replays
        # other instances on the same batch object
        boxes2, offsets2, ranges2, cells2 = mixed_batch(500, k=7)
        want2 = expected(boxes2, offsets2)
        total2 = sum(w.shape[0] for w in want2)
        d2 = to_dev(torch_cuda, boxes2)
        pairs2 = torch_cuda.empty((total2 + 1024, 2), dtype=torch_cuda.int32, device="cuda:0")
        io2 = torch_cuda.zeros((len(want2), 2), dtype=torch_cuda.int64, device="cuda:0")
        b.update(offsets2[:-1], np.diff(offsets2), ranges2, cells2)
        b.launch(d2, pairs2, io2, res)
        torch_cuda.cuda.synchronize()
        n2, err2 = res.tolist()
        if err2:
            n2, hio2 = b.check(d2, pairs2, finish=True)
        else:
            hio2 = io2.cpu().numpy()
        assert n2 == total2
        check_all(split(pairs2[:n2].cpu().numpy().view(np.uint32), hio2), want2)


def test_batch_unresolved_refuses_after_overflow(torch_cuda, eng):
    """Host re-binning: a launch whose only error is bins beyond LDS but whose
    pairs exceed the launch's capacity: the pairs past the capacity were
    dropped, so unresolved() must not vouch for the other instances' runs."""
    import oracle
    import pbig_amd
    from pbig_amd import workloads

    inst = [(workloads.uniform(50000, 3000.0, 1.0, 8.0, 311), 3000.0),
            (workloads.clustered(150000, 17000.0, 310), 17000.0)]
    boxes, offsets, ranges = concat(inst)
    want = expected(boxes, offsets)
    total = sum(w.shape[0] for w in want)
    d = to_dev(torch_cuda, boxes)
    big = torch_cuda.empty((total + 1024, 2), dtype=torch_cuda.int32, device="cuda:0")
    small = torch_cuda.empty((want[0].shape[0] // 2, 2), dtype=torch_cuda.int32, device="cuda:0")
    with host_levels(eng), pbig_amd.Batch(eng, offsets, ranges, [13.0, 40.0]) as b:
        b.check(d, big)  # sizes the workspace
        res = torch_cuda.zeros(2, dtype=torch_cuda.int64, device="cuda:0")
        b.launch(d, small, None, res)
        torch_cuda.cuda.synchronize()
        n_launch, err = res.tolist()
        assert err == 8 and n_launch > small.shape[0]
        with pytest.raises(pbig_amd.PbigError) as e:
            b.unresolved()
        assert e.value.status == pbig_amd.PBIG_E_OVERFLOW
        # with room for the launch's pairs, the first instance is complete
        io = torch_cuda.zeros((2, 2), dtype=torch_cuda.int64, device="cuda:0")
        b.launch(d, big, io, res)
        torch_cuda.cuda.synchronize()
        assert res[1].item() == 8
        flags = b.unresolved()
        assert flags.tolist() == [False, True]
        o, c = (int(v) for v in io[0].tolist())
        got = big[o:o + c].cpu().numpy().view(np.uint32)
        check_all([oracle.sort_pairs(got)], want[:1])
