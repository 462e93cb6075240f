"""GPU tests of the serving loop: every request's sorted pair set equals the
oracle's and the plain path's on the request alone, while requests share
batches, wait for pool space and overlap on the device."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def make_requests(count, seed, nmax, rate):
    from pbig_amd import workloads
    from pbig_amd.serve import Request

    reqs = []
    for q in workloads.serving_requests(count, seed, rate, nmin=1000, nmax=nmax):
        reqs.append(Request(rid=q["id"], boxes=workloads.request_boxes(q), L=q["L"], arrival=q["t"]))
    return reqs


def check_requests(torch, reqs):
    import oracle
    import pbig_amd

    eng = pbig_amd.Engine(0)
    try:
        for r in reqs:
            want = oracle.sort_pairs(oracle.sweep_pairs(r.boxes))
            p = r.pairs.cpu().numpy().view(np.uint32) if hasattr(r.pairs, "cpu") else r.pairs
            got = oracle.sort_pairs(p)
            assert np.array_equal(got, want), f"request {r.rid}: {got.shape[0]} vs {want.shape[0]} pairs"
            d = torch.from_numpy(r.boxes).to("cuda:0")
            out = torch.empty((max(4 * r.n, 1024), 2), dtype=torch.int32, device="cuda:0")
            n = eng.check_device(d, r.L, r.cell_size, out)
            alone = oracle.sort_pairs(out[:n].cpu().numpy().view(np.uint32))
            assert np.array_equal(alone, got), f"request {r.rid}: alone differs from in the mix"
    finally:
        eng.close()


def test_serve_all_at_once_small_pool(torch_cuda):
    """A pool of 150,000 boxes for 24 requests of up to 60,000 boxes: requests
    wait for space, batches hold several requests, results match alone."""
    from pbig_amd.serve import Server

    reqs = make_requests(24, 71, 60000, 1000.0)
    srv = Server(hbm_budget=2 * 150_000 * 16, slots=2, pool_frac=0.5)
    try:
        stats = srv.run(reqs, realtime=False)
    finally:
        srv.close()
    assert stats["requests"] == 24
    assert max(srv.batch_sizes) > 1
    assert stats["batches"] > 2  # the pool forced several rounds
    check_requests(torch_cuda, reqs)


def test_serve_poisson_realtime(torch_cuda):
    from pbig_amd.serve import Server

    reqs = make_requests(40, 73, 40000, 2000.0)
    srv = Server(hbm_budget=1 << 30, slots=3)
    try:
        stats = srv.run(reqs, realtime=True)
    finally:
        srv.close()
    assert stats["requests"] == 40
    assert all(r.t_done >= r.arrival for r in reqs)
    assert stats["latency_ms_p99"] >= stats["latency_ms_p50"] > 0
    check_requests(torch_cuda, reqs)


def test_serve_per_instance_baseline_matches(torch_cuda):
    import oracle
    import pbig_amd
    from pbig_amd import autotune
    from pbig_amd.serve import serve_per_instance

    reqs = make_requests(8, 79, 30000, 1000.0)
    for r in reqs:
        r.cell_size = float(autotune.cell_sizes([0, r.n], [r.L])[0])
    eng = pbig_amd.Engine(0)
    try:
        stats = serve_per_instance(eng, reqs, realtime=False)
    finally:
        eng.close()
    assert stats["requests"] == 8
    for r in reqs:
        assert np.array_equal(oracle.sort_pairs(r.pairs), oracle.sort_pairs(oracle.sweep_pairs(r.boxes)))


def test_serve_device_resident(torch_cuda):
    """Requests hand over device box arrays and get device pair lists."""
    from pbig_amd.serve import Server

    reqs = make_requests(30, 83, 50000, 5000.0)
    for r in reqs:
        r.device = torch_cuda.from_numpy(r.boxes).to("cuda:0")
    srv = Server(hbm_budget=2 * 200_000 * 16, slots=3, results="device")
    try:
        stats = srv.run(reqs, realtime=True)
    finally:
        srv.close()
    assert stats["requests"] == 30
    assert all(r.pairs.shape[0] == r.n_pairs for r in reqs)
    check_requests(torch_cuda, reqs)


def overfull_requests():
    from pbig_amd import workloads
    from pbig_amd.serve import Request

    reqs = []
    for i in range(10):
        if i % 3 == 0:
            n = 150_000
            b = workloads.clustered(n, 65536.0 * (n / 2097152) ** 0.5, 11 + i)
        else:
            n = 5_000 + 3_000 * i
            b = workloads.uniform(n, 65536.0 * (n / 4194304) ** 0.5, 1.0, 32.0, 11 + i)
        reqs.append(Request(rid=i, boxes=b, L=65536.0 * (n / 2097152) ** 0.5 if i % 3 == 0
                            else 65536.0 * (n / 4194304) ** 0.5))
    # a fixed table (the scenario needs the batch's bins to outgrow LDS)
    table = {"version": 1, "bin_target": 3000, "source": "test",
             "classes": [{"max_n": None, "boxes_per_cell": 1.0}]}
    return reqs, table


@pytest.mark.parametrize("host_levels", [True, False])
def test_serve_batch_with_overfull_bins(torch_cuda, host_levels):
    """Light load, batches whose clustered request has bins beyond LDS.  With
    host re-binning, such a batch hands over its other requests at once and
    reruns that one alone; with levels on the device, a slot's first finish
    settles its batch on them and its later batches complete on the device.
    Every request's pairs equal the oracle's and its own run."""
    from pbig_amd.serve import Server

    reqs, table = overfull_requests()
    # (mixed batches: grouped by kind, the clustered request would not share one)
    srv = Server(hbm_budget=1 << 30, slots=3, table=table, host_levels=host_levels, group_by_schedule=False)
    try:
        srv.run(reqs, realtime=False)  # sizes the slots' workspaces
        for r in reqs:
            r.pairs = None
        # arrivals 2 ms apart, in threes: light load, batches of up to three
        for r in reqs:
            r.arrival = 0.002 * (r.rid // 3)
        stats = srv.run(reqs, realtime=True)
    finally:
        srv.close()
    assert stats["requests"] == 10
    if host_levels:
        assert stats["requeued"] >= 1, "no batch was left incomplete: the case is not exercised"
    check_requests(torch_cuda, reqs)


def test_serve_c5_stream_full_sizes(torch_cuda):
    """The serving configuration at its stated request sizes: the first 16
    requests of the Poisson stream the bench serves (seed 5, n log-uniform
    in [1,000, 1,000,000], 50/50 uniform and clustered), arriving at 2,000
    requests/s, HBM budget 8 GB.  Three requests hold more than 500,000
    boxes.  Every request's pairs equal the oracle's sweep-and-prune set and
    its own plain-path run."""
    from pbig_amd import autotune, workloads
    from pbig_amd.serve import Request, Server

    table = autotune.load_table()
    spec = workloads.serving_requests(400, workloads.SEEDS["c5"], 2000.0)[:16]
    reqs = []
    for q in spec:
        b = workloads.request_boxes(q)
        reqs.append(Request(rid=q["id"], boxes=b, L=q["L"], arrival=q["t"],
                            cell_size=float(autotune.cell_sizes([0, b.shape[0]], [q["L"]], table, b)[0])))
    assert sum(r.n > 500_000 for r in reqs) >= 2
    assert {q["kind"] for q in spec} == {"uniform", "clustered"}
    assert max(r.n for r in reqs) > 900_000
    srv = Server(hbm_budget=8 << 30, slots=3, table=table)
    try:
        stats = srv.run(reqs, realtime=True)
    finally:
        srv.close()
    assert stats["requests"] == 16
    check_requests(torch_cuda, reqs)
