"""Benchmark of the plain path on the 4,194,304-box uniform instance.

  python bench.py [--gpus 1] [--steps K] [--warmup W]

One step = one pass of the whole pipeline (binning, scan, scatter, per-cell
test, compaction) over 4,194,304 float32 boxes (centres U[0,65536)^2, sides
U[1,32], seed 2) already resident in HBM, with the pair list written to HBM.
The reference's cell-size rule gives 256 for this instance.  Prints one JSON
line: overlaps/s (whole job), the roofline of the dominant kernel from HIP
events recorded around every kernel of the timed steps, and the CPU
baseline (the oracle's grid restatement on all host cores, plus the
reference checker's brute-force predicate on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak, GB/s


def alg_bytes(N, C, M, P):
    """Algorithmic bytes per kernel: 32N + 12C + 24M + 8P in total.
    count 16N + 4C; scan 8C; scatter 16N + 4M; test 20M + 8P."""
    return {
        "count": 16 * N + 4 * C,
        "scan": 8 * C,
        "scatter": 16 * N + 4 * M,
        "test": 20 * M + 8 * P,
        "total": 32 * N + 12 * C + 24 * M + 8 * P,
    }


def reference_grid(boxes, L, cs):
    """Cells C and memberships M of the reference's single-level grid at cell
    size cs (cells per row ceil(L / cs); a coordinate's cell is
    floor(v * (1/cs)) in float32, clamped into the grid), the terms that
    price the algorithmic bytes: counted here with numpy from the boxes, the
    same on every path."""
    import numpy as np

    nrc = max(1, int(np.ceil(float(L) / float(cs))))
    inv = np.float32(1.0) / np.float32(cs)
    b = np.asarray(boxes, dtype=np.float32)
    M = 0
    for i0 in range(0, b.shape[0], 1 << 20):
        q = b[i0:i0 + (1 << 20)]
        cell = lambda v: np.clip(np.floor(v * inv), 0, nrc - 1).astype(np.int64)
        M += int(((cell(q[:, 1]) - cell(q[:, 0]) + 1) * (cell(q[:, 3]) - cell(q[:, 2]) + 1)).sum())
    return nrc * nrc, M


def cpu_baseline(boxes, L, cs, n_pairs):
    """CPU baselines on the host cores (about 20 s in all):
    * value: the oracle (CPU restatement of the reference's grid algorithm)
      over the full 4M instance on all host threads, median of 5 after one
      discarded warm-up; also once on one thread, and on the clustered 2M
      instance (median of 3);
    * the reference checker's brute-force predicate (src/pbig/pbig.cu:996-1001)
      on the 16,384-box instance in full (one thread and all threads, median of
      5) and on the first 2^18 boxes of the 4M instance (all threads, one run).
    """
    import oracle
    import pbig_amd
    from pbig_amd import workloads

    oracle.build()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = min(threads, 16)

    def timed(fn, reps, warm=True):
        if warm:
            fn()
        ts, out = [], None
        for _ in range(reps):
            t0 = time.perf_counter()
            out = fn()
            ts.append(time.perf_counter() - t0)
        return statistics.median(ts), out

    t, got = timed(lambda: oracle.grid_pairs(boxes, L, cs, threads=threads), 5)
    assert len(got) == n_pairs
    t1, _ = timed(lambda: oracle.grid_pairs(boxes, L, cs, threads=1), 1, warm=False)
    b3, L3 = workloads.c3_clustered()
    cs3 = pbig_amd.reference_cell_size(b3, L3)
    t3, got3 = timed(lambda: oracle.grid_pairs(b3, L3, cs3, threads=threads), 3)
    b1, _ = workloads.c1_small_uniform()
    tests1 = b1.shape[0] * (b1.shape[0] - 1) / 2
    tb1, n1 = timed(lambda: oracle.brute_count(b1, threads=1), 5)
    tbn, _ = timed(lambda: oracle.brute_count(b1, threads=threads), 5)
    sample = 1 << 18
    tests2 = sample * (sample - 1) / 2
    tb2, n2 = timed(lambda: oracle.brute_count(boxes[:sample], threads=threads), 1, warm=False)
    return {
        "value": n_pairs / t,
        "unit": "overlaps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"full 4,194,304-box instance through the oracle's grid restatement at cell size {cs} "
                  f"(median of 5 after 1 warm-up, {t:.3f} s per run, {threads} threads)",
        "seconds_per_instance": t,
        "seconds_per_instance_one_thread": t1,
        "clustered_2m": {"seconds": t3, "pairs": len(got3), "cell_size": cs3, "overlaps_per_s": len(got3) / t3,
                         "threads": threads},
        "bruteforce_checker": {
            "predicate": "src/pbig/pbig.cu:996-1001, all pairs",
            "instance_16k": {"pairs": n1, "tests": int(tests1), "seconds_one_thread": tb1,
                             "tests_per_s_one_thread": tests1 / tb1, "seconds_all_threads": tbn,
                             "tests_per_s_all_threads": tests1 / tbn},
            "first_2pow18_of_4m": {"pairs": n2, "tests": int(tests2), "seconds_all_threads": tb2,
                                   "tests_per_s_all_threads": tests2 / tb2,
                                   "extrapolated_full_instance_seconds": tb2 * (boxes.shape[0] / sample) ** 2},
        },
        "host_cpu": _cpu_model(),
    }


def read_traffic(path_name, cs, bin_target):
    """HBM bytes per launch by kernel, measured with rocprofv3 FETCH_SIZE and
    WRITE_SIZE (each scaled by the same counter's reading for a copy of known
    size) and recorded in profiles/ for this exact configuration."""
    fn = os.path.join(ROOT, "profiles", f"traffic_plain_uniform_4m_{path_name}_cs{int(cs)}_bt{bin_target}.txt")
    if not os.path.exists(fn):
        return None, None
    out = {}
    with open(fn) as f:
        for line in f:
            if line.startswith("kernel="):
                kv = dict(tok.split("=", 1) for tok in line.split())
                out[kv["kernel"]] = float(kv["traffic_bytes"])
    return out, os.path.relpath(fn, ROOT)


def clustered_leg(torch, eng_cls, steps=50, reps=5):
    """The clustered 2,097,152-box instance (seed 3), boxes resident, at the
    autotuner's cell size and bin target.
    * One synchronous call per step, per path: the binned path forced (its
      heavy coarse bins are re-binned, which needs the host between levels),
      the general path forced, and the automatic choice.
    * The automatic choice as a stream of asynchronous calls (after the
      synchronous call that settles its path): `steps` calls timed with two
      events, then again with HIP events around every kernel; the roofline
      of the dominant kernel and of the whole call, with the algorithmic
      bytes priced at the reference rule's cell size like the 4M leg's."""
    import pbig_amd
    from pbig_amd import autotune, workloads

    boxes, L = workloads.c3_clustered()
    n = boxes.shape[0]
    dev = torch.device("cuda:0")
    d = torch.from_numpy(boxes).to(dev)
    pairs = torch.empty((40 * n, 2), dtype=torch.int32, device=dev)
    cell = float(autotune.cell_sizes([0, n], [L], boxes=boxes)[0])
    ref_cs = pbig_amd.reference_cell_size(boxes, L)
    out = {"workload": "2,097,152 float32 rects, 90% in 64 Gaussian clusters (sigma 327.68) + 10% uniform "
                       "in [0,65536)^2, log-normal sides, seed 3, boxes resident in HBM",
           "cell_size": cell, "reference_cell_size": ref_cs}
    # (the algorithmic bytes at the reference rule's cell: its grid counted
    # in numpy; the pair count is the same at every cell size)
    C_ref, M_ref = reference_grid(boxes, L, ref_cs)
    k_ref, B = None, None
    bin_target = int(autotune.load_table()["bin_target"])
    for name, path in (("binned", pbig_amd.PATH_BINNED), ("general", pbig_amd.PATH_GENERAL),
                       ("auto", pbig_amd.PATH_AUTO)):
        e = eng_cls(0)
        e.configure(path, bin_target)
        k = e.check_device(d, L, cell, pairs)
        if k_ref is None:
            k_ref, B = k, alg_bytes(n, C_ref, M_ref, k)
        assert k == k_ref, (name, k, k_ref)
        torch.cuda.synchronize()
        t = []
        for _ in range(reps):
            t0 = time.perf_counter()
            k2 = e.check_device(d, L, cell, pairs)
            t.append(time.perf_counter() - t0)
            assert k2 == k
        st = e.stats()
        ms = statistics.median(t) * 1e3
        out[name] = {"ms_per_call": ms, "pairs": k, "overlaps_per_s": k / (ms / 1e3), "path": st["path"]}
        if name == "auto":
            stream = torch.cuda.Stream(device=dev)
            res = torch.zeros(2, dtype=torch.int64, device=dev)
            with torch.cuda.stream(stream):
                for _ in range(3):
                    e.check_device_async(d, L, cell, pairs, res, stream=stream)
            torch.cuda.synchronize()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            with torch.cuda.stream(stream):
                a.record(stream)
                for _ in range(steps):
                    e.check_device_async(d, L, cell, pairs, res, stream=stream)
                b.record(stream)
            torch.cuda.synchronize()
            assert res.cpu().tolist() == [k, 0], res.cpu().tolist()
            ms_async = a.elapsed_time(b) / steps
            e.timing_start(steps)
            with torch.cuda.stream(stream):
                for _ in range(steps):
                    e.check_device_async(d, L, cell, pairs, res, stream=stream)
            torch.cuda.synchronize()
            per_kernel, calls, path_name = e.timing_read()
            assert res.cpu().tolist() == [k, 0]
            avg = {kk: v / calls for kk, v in per_kernel.items() if v > 0}
            kernel_alg = {"count": B["count"], "scatter": B["scatter"], "test": B["test"],
                          "scan_reduce": B["scan"] / 2, "scan_apply": B["scan"] / 2, "clear": 4 * C_ref,
                          "bin_hist": B["count"], "bin_scan": B["scan"], "bin_starts": 0.0,
                          "bin_scatter": B["scatter"], "bin_test": B["test"],
                          "bin_place": B["count"] + B["scan"] + B["scatter"],
                          "fine_count": B["count"], "fine_place": B["scatter"]}
            dom = max(avg, key=lambda kk: avg[kk])
            achieved = kernel_alg.get(dom, 0.0) / (avg[dom] / 1e3) / 1e9
            dom_bytes, dom_ms = kernel_alg.get(dom, 0.0), avg[dom]
            if path_name == "deep":
                # per-level times only: the whole call (level 0, the device
                # levels and the dense-bin kernel) against its algorithmic bytes
                dom, dom_bytes, dom_ms = "whole call (level 0 and device levels)", B["total"], ms_async
                achieved = dom_bytes / (dom_ms / 1e3) / 1e9
            out["auto_async"] = {
                "ms_per_call": ms_async, "overlaps_per_s": k / (ms_async / 1e3), "path": path_name,
                "kernel_avg_ms": avg,
                "alg_bytes_per_call": B["total"],
                "alg_GBps": B["total"] / (ms_async / 1e3) / 1e9,
                "frac_of_peak": B["total"] / (ms_async / 1e3) / 1e9 / HBM_PEAK_GBS,
                "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                             "alg_bytes_per_launch": dom_bytes, "avg_launch_ms": dom_ms},
            }
        e.close()
    del d, pairs
    torch.cuda.empty_cache()
    return out


def compat_leg(torch, reps=3):
    """The reference's own benchmark inputs (its shipped build: unsigned 3-D
    boxes, src/pbig/main.cc:204-335 generator, srand(100), 10^7 boxes, range
    10^4, S = 100): test1 (uniform centres) and test3 (normal centres), with
    the reference's cell-size rule.  Timed as the reference times pbig_check
    (copy in, check, pairs handed over: into a host buffer, and chunk by
    chunk to a callback as pbig_check's report gets them), and with boxes
    resident and pairs only counted.  The reference's published times are on a different,
    much older GPU (context only)."""
    import numpy as np

    import oracle
    import pbig_amd

    oracle.build()
    out = {}
    published = {"test1_UU": (533211.0, 49514252, "benchmarks/test1/test1.set.csv:8"),
                 "test3_UN": (1742764.0, 1121815231, "benchmarks/test3/test3.set.csv:8")}
    eng = pbig_amd.Engine(0)
    for name, sim in (("test1_UU", "UU"), ("test3_UN", "UN")):
        b = oracle.ref_generate(10**7, 10000, 100, sim, 3)
        d = torch.from_numpy(b.view(np.int32)).to("cuda:0")
        k = eng.check_u32(d, 3, 10000, 0)
        assert k == published[name][1], (name, k)
        torch.cuda.synchronize()
        t = []
        for _ in range(reps):
            t0 = time.perf_counter()
            eng.check_u32(d, 3, 10000, 0)
            t.append(time.perf_counter() - t0)
        resident_ms = statistics.median(t) * 1e3
        # end to end: host records in, every pair back on the host
        t = []
        for _ in range(2):
            t0 = time.perf_counter()
            pairs = eng.check_host_u32(b, 3, 10000, 0, capacity=k)
            t.append(time.perf_counter() - t0)
            assert pairs.shape[0] == k
            del pairs
        e2e_ms = statistics.median(t) * 1e3
        # the pairs handed to a callback chunk by chunk (pbig_check's report
        # runs this way), here one that counts them
        t = []
        for _ in range(2):
            seen = [0]

            def count(p):
                seen[0] += p.shape[0]

            t0 = time.perf_counter()
            assert eng.check_host_u32_each(b, 3, 10000, 0, count) == k and seen[0] == k
            t.append(time.perf_counter() - t0)
        each_ms = statistics.median(t) * 1e3
        out[name] = {"pairs": k, "resident_count_ms": resident_ms, "host_records_to_host_pairs_ms": e2e_ms,
                     "host_records_to_callbacks_ms": each_ms,
                     "reference_published_us": published[name][0], "reference_source": published[name][2]}
        del d
    eng.close()
    torch.cuda.empty_cache()
    return out


def batched_leg(torch, steps, warmup, baseline=True):
    """The 4,096-instance workload (n log-uniform in [64, 262144], L = 16
    sqrt(n), sides U[1,8], seed 4): one HIP-graph replay of the batch per
    step, against per-instance calls of the plain path on the same instances
    (synchronous calls, as pbig_check is, and asynchronous calls queued back
    to back)."""
    import numpy as np

    import pbig_amd
    from pbig_amd import autotune, workloads

    boxes, offsets, ranges = workloads.batched(4096, workloads.SEEDS["c4"])
    k = offsets.shape[0] - 1
    N = int(boxes.shape[0])
    table = autotune.load_table()
    cells = autotune.cell_sizes(offsets, ranges, table, boxes)
    dev = torch.device("cuda:0")
    eng = pbig_amd.Engine(0)
    eng.configure(pbig_amd.PATH_AUTO, int(table["bin_target"]))
    d = torch.from_numpy(boxes).to(dev)
    del boxes
    cap = N // 4
    pairs = torch.empty((cap, 2), dtype=torch.int32, device=dev)
    b = pbig_amd.Batch(eng, offsets, ranges.astype(np.float64), cells)
    total, io = b.check(d, pairs)
    info = b.info()
    res = torch.zeros(2, dtype=torch.int64, device=dev)
    io_dev = torch.zeros((k, 2), dtype=torch.int64, device=dev)
    b.capture(d, pairs, io_dev, res)
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        b.replay(stream)
    torch.cuda.synchronize()
    s_ev, e_ev = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    s_ev.record(stream)
    for _ in range(steps):
        b.replay(stream)
    e_ev.record(stream)
    torch.cuda.synchronize()
    wall_batch = (time.perf_counter() - t0) / steps
    ms_batch = s_ev.elapsed_time(e_ev) / steps
    r = res.cpu().tolist()
    assert r == [total, 0], (r, total)
    assert np.array_equal(io_dev.cpu().numpy().astype(np.uint64), io)

    batch_s = ms_batch / 1e3
    out = {
        "workload": "4,096 instances, n log-uniform [64, 262144], L = 16 sqrt(n), sides U[1,8], seed 4, "
                    "boxes resident in HBM",
        "instances": k, "boxes": N, "pairs": total, "chunks": info["n_chunks"], "bins": info["n_bins"],
        "cell_sizes": "autotuner table (" + str(table.get("source", "default")) + ")",
        "bin_target": int(table["bin_target"]),
        "graph_replay_ms": ms_batch,
        "graph_replay_host_wall_ms": wall_batch * 1e3,
        "boxes_per_s": N / batch_s,
        "pairs_per_s": total / batch_s,
    }
    if not baseline:
        b.close()
        eng.close()
        return out
    # Per-instance plain path on the same device array and cell sizes.
    views = [d[int(offsets[i]):int(offsets[i + 1])] for i in range(k)]
    L = [float(x) for x in ranges]
    cs = [float(x) for x in cells]

    def sync_loop():
        tot = 0
        for i in range(k):
            tot += eng.check_device(views[i], L[i], cs[i], pairs)
        return tot

    sync_loop()  # warm-up: plans and workspace
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tot_sync = sync_loop()
    wall_sync = time.perf_counter() - t0
    assert tot_sync == total, (tot_sync, total)

    res1 = torch.zeros(2, dtype=torch.int64, device=dev)

    def async_loop():
        for i in range(k):
            eng.check_device_async(views[i], L[i], cs[i], pairs, res1, stream=stream)

    async_loop()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s_ev.record(stream)
    async_loop()
    e_ev.record(stream)
    torch.cuda.synchronize()
    wall_async = time.perf_counter() - t0
    ms_async = s_ev.elapsed_time(e_ev)
    assert res1.cpu().tolist() == [int(io[-1, 1]), 0]
    b.close()
    eng.close()
    del d, pairs, views
    torch.cuda.empty_cache()
    out.update({
        "per_instance_sync_ms": wall_sync * 1e3,
        "per_instance_async_ms": ms_async,
        "per_instance_async_host_wall_ms": wall_async * 1e3,
        "speedup_vs_per_instance_sync": wall_sync / batch_s,
        "speedup_vs_per_instance_async": (max(ms_async, wall_async * 1e3) / 1e3) / batch_s,
    })
    return out


def serving_leg(torch, count, rate, budget_gb):
    """Poisson stream of requests (n log-uniform in [1,000, 1,000,000], 50/50
    uniform and clustered, seed 5): the continuous-batching server against
    the same stream served one request at a time by the plain path.

    Two settings: "resident" (each request hands over a device box array and
    its pairs stay in HBM for a consumer on the GPU) and "host" (boxes
    uploaded from pinned host memory, pairs copied back to pinned host
    memory).  In both, the consumer reads the pairs in place and keeps their
    count.  One server per setting serves every run, as a long-running
    server would (its slots' workspaces and device levels settle in the
    warm-up run).  Capacity: every request arrives at once (median of three
    runs by makespan).  Latency: arrivals at the stated rate (after one
    warm-up run, the median of three runs by p99)."""
    import numpy as np

    import pbig_amd
    from pbig_amd import autotune, workloads
    from pbig_amd.serve import Request, Server, count_only, keep_copy, serve_per_instance

    table = autotune.load_table()
    spec = workloads.serving_requests(count, workloads.SEEDS["c5"], rate)
    dev = torch.device("cuda:0")
    reqs = []
    for q in spec:
        b = workloads.request_boxes(q)
        r = Request(rid=q["id"], boxes=b, L=q["L"], arrival=q["t"],
                    cell_size=float(autotune.cell_sizes([0, b.shape[0]], [q["L"]], table, b)[0]),
                    narrow=autotune.sides_bounded(b))
        r.host = torch.from_numpy(b).pin_memory()
        r.device = r.host.to(dev)
        reqs.append(r)
    torch.cuda.synchronize()
    total_boxes = sum(r.n for r in reqs)
    res = {"workload": f"{count} requests, n log-uniform [1000, 1000000], 50/50 uniform (L = 65536 sqrt(n/4194304), "
                       f"sides U[1,32]) and clustered (L = 65536 sqrt(n/2097152)), seed 5",
           "boxes": total_boxes, "hbm_budget_gb": budget_gb, "pool_share": 0.3, "arrival_rate_per_s": rate, "slots": 3}
    eng = pbig_amd.Engine(0)
    eng.configure(pbig_amd.PATH_AUTO, int(table["bin_target"]))

    def reset():
        for r, q in zip(reqs, spec):
            r.arrival = q["t"]
            r.pairs = None

    srv = None  # one server per setting serves every run, as a long-running one would

    def server(mode, realtime, consumer=count_only):
        reset()
        srv.consumer = consumer
        return srv.run(reqs, realtime=realtime)

    def one(mode, realtime, consumer=count_only):
        reset()
        return serve_per_instance(eng, reqs, realtime=realtime, results=mode, consumer=consumer)

    for mode in ("device", "host"):
        key = "resident" if mode == "device" else "host"
        if mode == "host":
            for r in reqs:
                r.device = None
        srv = Server(hbm_budget=int(budget_gb * (1 << 30)), slots=3, pool_frac=0.3, table=table, results=mode)
        server(mode, False)  # warm-up: plans, workspaces, pinned buffers, the slots' device levels
        one(mode, False)

        def median_run(fn):  # a makespan of tens of ms: the median of three runs
            runs = sorted((fn(mode, False) for _ in range(3)), key=lambda o: o["makespan_s"])
            return runs[1]

        cap_s, cap_o = median_run(server), median_run(one)
        # at the stated rate: one warm-up run (the slots' first plain and
        # general-path calls size their workspaces), then the median of three
        # runs by p99 latency
        server(mode, True)
        one(mode, True)

        def median_p99(fn):
            runs = sorted((fn(mode, True) for _ in range(3)), key=lambda o: o["latency_ms_p99"])
            return runs[1]

        lat_s, lat_o = median_p99(server), median_p99(one)
        res[key] = {"server_capacity": cap_s, "per_instance_capacity": cap_o,
                    "speedup_capacity": cap_s["boxes_per_s"] / cap_o["boxes_per_s"],
                    "server_at_rate": lat_s, "per_instance_at_rate": lat_o}
        if mode == "device":
            # every request's pairs alone (plain path) and in the mix (server),
            # compared by count and order-independent sums over the pairs
            server(mode, False, consumer=keep_copy)
            mix = {r.rid: r.pairs for r in reqs}
            one(mode, False, consumer=keep_copy)
            ok = True
            for r in reqs:
                a, b = mix[r.rid], r.pairs
                if a.shape[0] != b.shape[0]:
                    ok = False
                    break
                if a.shape[0]:
                    ka = (a[:, 1].to(torch.int64) << 32) | (a[:, 0].to(torch.int64) & 0xFFFFFFFF)
                    kb = (b[:, 1].to(torch.int64) << 32) | (b[:, 0].to(torch.int64) & 0xFFFFFFFF)
                    if not torch.equal(torch.sort(ka).values, torch.sort(kb).values):
                        ok = False
                        break
            res["results_equal_alone_and_in_mix"] = ok
            for r in reqs:
                r.pairs = None
            mix = None
        srv.close()
        srv = None
        torch.cuda.empty_cache()
    eng.close()
    return res


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def calibrate_copy(torch):
    """A 256 MiB device-to-device copy, three times: a kernel of known traffic
    (256 MiB read, 256 MiB written) for scaling rocprofv3's counters."""
    src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    # Cell 32 is the autotuner table's one box per cell at this density
    # (pbig_amd/tuning.json); every box then spans at most two cells per axis
    # and the narrow-box test schedule runs.  Its best bin target here is
    # 4,096: bins of 64 x 64 cells, 1,024 of them, four on each of the 256
    # test workgroups, about 4,200 memberships each (the schedule takes up to
    # 5,120).  The general test schedule's best is cell 64, bin target 2,000.
    ap.add_argument("--cell-size", type=float, default=32.0,
                    help="fine cell side; 0 = the reference rule (256 on this instance)")
    ap.add_argument("--path", choices=("auto", "general", "binned"), default="auto")
    ap.add_argument("--bin-target", type=int, default=4096)
    ap.add_argument("--no-batched", action="store_true", help="skip the 4,096-instance batch leg")
    ap.add_argument("--no-batched-baseline", action="store_true",
                    help="batch leg without the per-instance comparison loops (for profiling)")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-clustered", action="store_true", help="skip the clustered 2M-box leg")
    ap.add_argument("--no-compat", action="store_true", help="skip the reference's own 10^7-box unsigned inputs")
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--serve-budget-gb", type=float, default=8.0)
    ap.add_argument("--batched-steps", type=int, default=10)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 256 MiB device-to-device copy (torch) three times, a kernel of known "
                         "traffic for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE")
    ap.add_argument("--only", choices=("clustered",),
                    help="run only this leg (steps = its asynchronous calls) and print its figures, not the "
                         "bench line: a profiler then sees that leg's kernels alone")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.only:
        import torch

        import pbig_amd

        if args.calibrate_copy:
            calibrate_copy(torch)
        print(json.dumps({args.only: clustered_leg(torch, pbig_amd.Engine, steps=args.steps)}))
        return

    import numpy as np
    import torch

    import pbig_amd
    from pbig_amd import workloads

    boxes, L = workloads.c2_large_uniform()
    N = boxes.shape[0]
    ref_cs = pbig_amd.reference_cell_size(boxes, L)
    cs = args.cell_size or ref_cs
    dev = torch.device("cuda:0")
    d_boxes = torch.from_numpy(boxes).to(dev)
    eng = pbig_amd.Engine(0)

    cap = 3 * 1024 * 1024
    d_pairs = torch.empty((cap, 2), dtype=torch.int32, device=dev)
    # B_alg is priced at the reference rule's cell size (cells C and
    # memberships M of the single-level grid), whatever cell size runs.
    eng.configure(pbig_amd.PATH_GENERAL)
    n_ref = eng.check_device(d_boxes, L, ref_cs, d_pairs)
    C_ref, M_ref = reference_grid(boxes, L, ref_cs)
    # The measured configuration; this call also sizes the workspace.
    eng.configure(pbig_amd.PATH_AUTO if args.path == "auto" else
                  {"general": pbig_amd.PATH_GENERAL, "binned": pbig_amd.PATH_BINNED}[args.path], args.bin_target)
    n_pairs = eng.check_device(d_boxes, L, cs, d_pairs)
    assert n_pairs == n_ref, (n_pairs, n_ref)
    st = eng.stats()
    B = alg_bytes(N, C_ref, M_ref, n_pairs)

    if args.calibrate_copy:
        calibrate_copy(torch)

    stream = torch.cuda.Stream(device=dev)
    result = torch.zeros(2, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        for _ in range(args.warmup):
            eng.check_device_async(d_boxes, L, cs, d_pairs, result, stream=stream)
    torch.cuda.synchronize()

    # Timed region: K steps back to back on one stream, bracketed by two
    # events and a synchronisation on each side.
    start_ev = torch.cuda.Event(enable_timing=True)
    end_ev = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.cuda.stream(stream):
        start_ev.record(stream)
        for _ in range(args.steps):
            eng.check_device_async(d_boxes, L, cs, d_pairs, result, stream=stream)
        end_ev.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = start_ev.elapsed_time(end_ev)
    res = result.cpu().tolist()
    assert res[1] == 0 and res[0] == n_pairs, res
    # Kernel durations: the same K steps again with a HIP event recorded
    # around every kernel on the launch stream (the events add gaps between
    # kernels, so the step time above is taken without them).
    eng.timing_start(args.steps)
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        for _ in range(args.steps):
            eng.check_device_async(d_boxes, L, cs, d_pairs, result, stream=stream)
    torch.cuda.synchronize()
    per_kernel, calls, path_name = eng.timing_read()
    res = result.cpu().tolist()
    assert res[1] == 0 and res[0] == n_pairs, res

    ms_per_step = gpu_ms / args.steps
    value = n_pairs / (ms_per_step / 1e3)
    avg = {k: v / calls for k, v in per_kernel.items()}
    if path_name == "binned":
        # A single instance's test kernel writes the output itself and its
        # scatter scans the bucket offsets: the offsets, pair scan and
        # compaction kernels are not launched (their slots would only time
        # two back-to-back events).
        avg.pop("bin_starts", None)
        avg.pop("bin_pair_scan", None)
        avg.pop("bin_compact", None)
    # B_alg terms by kernel (the binned kernels do the same four jobs)
    # (bin_starts moves a few KB: counted in bin_scan; the single-instance test kernel writes the
    # output itself, so it carries the 8P term)
    # (the single-pass launch's place kernel does the count, scan and scatter
    # jobs in one pass over the boxes)
    kernel_alg = {"count": B["count"], "scatter": B["scatter"], "test": B["test"],
                  "scan_reduce": B["scan"] / 2, "scan_apply": B["scan"] / 2, "clear": 4 * C_ref,
                  "bin_hist": B["count"], "bin_scan": B["scan"], "bin_starts": 0.0, "bin_scatter": B["scatter"],
                  "bin_test": B["test"], "bin_place": B["count"] + B["scan"] + B["scatter"]}
    dominant = max(avg, key=lambda k: avg[k])
    traffic, traffic_src = read_traffic(path_name, cs, args.bin_target)
    achieved = kernel_alg[dominant] / (avg[dominant] / 1e3) / 1e9
    out = {
        "metric": "overlaps/s & HBM GB/s (4M boxes); batched vs per-instance; serve p50/p99",
        "value": value,
        "unit": "overlaps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "plain large instance: 4,194,304 float32 rects, centres U[0,65536)^2, "
                        "sides U[1,32], seed 2, boxes resident in HBM",
            "n_boxes": N, "range": L, "cell_size": cs, "reference_cell_size": ref_cs,
            "pairs": n_pairs, "path": path_name, "cells": st["n_cells"],
            "bins_per_row": st["bins_per_row"], "cells_per_bin_side": st["cells_per_bin_side"],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic.get(dominant) if traffic else None,
            "traffic_source": traffic_src if traffic else None,
            "alg_bytes_per_launch": kernel_alg[dominant],
            "avg_launch_ms": avg[dominant],
        },
        "pipeline": {
            "alg_bytes_per_step": B["total"],
            "alg_GBps": B["total"] / (ms_per_step / 1e3) / 1e9,
            "frac_of_peak": B["total"] / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS,
            "kernel_avg_ms": avg,
            "kernel_sum_ms": sum(avg.values()),
            "host_wall_ms_per_step": wall * 1e3 / args.steps,
        },
    }
    eng.close()
    del d_boxes, d_pairs
    torch.cuda.empty_cache()
    if not args.no_batched:
        out["batched"] = batched_leg(torch, args.batched_steps, 2, baseline=not args.no_batched_baseline)
    if not args.no_clustered:
        out["clustered"] = clustered_leg(torch, pbig_amd.Engine)
    if not args.no_compat:
        out["reference_compat"] = compat_leg(torch)
    if not args.no_serving:
        out["serving"] = serving_leg(torch, args.serve_requests, args.serve_rate, args.serve_budget_gb)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(boxes, L, ref_cs, n_pairs)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
