"""Serving loop: continuous batching of box-intersection requests on one GPU.

A request is one instance (boxes, domain side).  The reference serves one
instance per blocking pbig_check call (src/pbig/pbig.h:10, copy-in at
pbig.cu:1401, per-call grid allocation at 1385 and free at 1979-1980); here a
server keeps request box sets resident in an HBM pool and runs many requests
per launch:

* Admission.  A request is admitted when its boxes fit in the pool (first-fit
  over a free list, the pool being the box share of the HBM budget); its
  boxes are then copied into the pool on a copy stream, from host memory or
  from the device array the client handed over.  Requests are admitted in
  arrival order.
* Continuous batching.  Each of ``slots`` batch slots owns a pbig_hip_batch
  (workspace reused across batches and bounded by the slot's share of the
  budget), a stream and result buffers.  Whenever a slot is free, the
  admitted requests form its next batch: the slot's stream waits for their
  uploads, the batch is re-planned and launched, and the result words and
  the per-request (offset, count) table come back by asynchronous copies.
  Requests that arrive while batches run are admitted as pool space frees up
  and go into the next free slot, so uploads, the kernels of several batches
  and result copies overlap on the device.  A batch of one request (light
  load) runs on the plain path instead, on the slot's own context: one
  launch, no batch plan, reading the request's own device array when it
  handed one over; a request whose bins show mostly dense clusters is
  launched again, asynchronously, on the general path.
* Completion work off the loop.  A launch whose result word asks for host
  work (bins beyond LDS, regrowth) is finished by a synchronous call in a
  worker thread; every slot has its own context, so
  the loop keeps admitting, launching and delivering for the other slots
  meanwhile.  A batch's bins beyond LDS are re-binned by levels planned on
  the device; the first such finish settles the slot's batch on them, and
  its later launches complete on the device (regrowing when a batch needs
  more than the levels hold).  With ``host_levels`` (re-binning by the host
  between levels), under light load (no more requests waiting than slots),
  when a batch's only trouble is bins beyond LDS, its requests without such
  a bin are delivered at once and the others are queued again as a batch of
  their own (one request alone runs on the plain path, whose dense exit is
  a short general-path call), so one clustered request does not hold up
  its batch; under a backlog the batch is finished whole, which costs less
  device time.
* Completion.  With ``results="device"`` a request's pairs stay in HBM (a
  view of the slot's pair buffer, valid while the consumer runs); with
  ``results="host"`` they are copied to pinned host memory first.  The
  consumer callback receives them; its default keeps a copy in the request.
  The request's pool space is freed at once.  Latency = delivery - arrival.

Each request's pair set does not depend on what else is in its batch: the
instance is planned and tested alone inside the batch (its own grid, bins and
output run); the tests check this against the request run alone.
"""
from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

import pbig_amd
from pbig_amd import autotune

_ERR_OVERFULL = 8  # error word: a coarse bin did not fit in LDS (re-binned on the host)
_ERR_DENSE = 16    # error word: the instance is mostly dense clusters (the general path takes it)
_ERR_BINCAP = 32   # error word: a bin outgrew its share of the single-pass launch


@dataclass
class Request:
    """One instance: boxes (n, 4) float32 {xmin,xmax,ymin,ymax}, domain side L.
    The boxes come from ``device`` (a device tensor), else ``host`` (a pinned
    torch tensor), else ``boxes`` (numpy)."""
    rid: int
    boxes: np.ndarray | None
    L: float
    arrival: float = 0.0
    cell_size: float = 0.0           # 0 = the autotuner's table
    narrow: bool | None = None       # sides bounded (autotune.sides_bounded); None = from `boxes` at submission
    host: object = None
    device: object = None
    # filled by the server
    pairs: object = None             # numpy (host results) or device tensor
    n_pairs: int = -1
    t_admit: float = 0.0
    t_launch: float = 0.0
    t_ready: float = 0.0             # kernels done, result words on the host
    t_done: float = 0.0
    batch_id: int = -1
    _off: int = -1
    _alone: bool = False             # rerun alone: its batch left it incomplete

    @property
    def n(self) -> int:
        src = self.device if self.device is not None else (self.host if self.host is not None else self.boxes)
        return int(src.shape[0])

    @property
    def latency(self) -> float:
        return self.t_done - self.arrival


def keep_copy(req: Request, pairs, stream):
    """Default consumer: keep the pairs in the request (host: numpy copy;
    device: a device copy ordered on the slot's stream)."""
    if isinstance(pairs, np.ndarray):
        req.pairs = pairs.copy()
    else:
        import torch

        with torch.cuda.stream(stream):
            req.pairs = pairs.clone()


def count_only(req: Request, pairs, stream):
    """Consumer that keeps only the number of pairs (already in n_pairs)."""
    req.pairs = None


class Pool:
    """First-fit allocator over a device array of boxes (16 bytes each)."""

    def __init__(self, capacity_boxes: int):
        self.capacity = int(capacity_boxes)
        self.free = [(0, self.capacity)]  # sorted (start, length)
        self.used = 0

    def alloc(self, n: int) -> int:
        n = max(int(n), 1)
        for k, (s, ln) in enumerate(self.free):
            if ln >= n:
                if ln == n:
                    self.free.pop(k)
                else:
                    self.free[k] = (s + n, ln - n)
                self.used += n
                return s
        return -1

    def release(self, s: int, n: int):
        n = max(int(n), 1)
        self.used -= n
        self.free.append((s, n))
        self.free.sort()
        merged = []
        for a, ln in self.free:
            if merged and merged[-1][0] + merged[-1][1] == a:
                merged[-1] = (merged[-1][0], merged[-1][1] + ln)
            else:
                merged.append((a, ln))
        self.free = merged


@dataclass
class _Slot:
    stream: object
    batch: object = None
    pairs: object = None             # device (cap, 2) int32
    io: object = None                # device (k, 2) int64
    res: object = None               # device (2,) int64
    h_res: object = None             # pinned
    h_io: object = None              # pinned
    h_pairs: object = None           # pinned (host results)
    ev: object = None
    state: str = "free"              # free | running | finishing | copying
    mode: str = "batch"              # batch | plain (a single request on the plain path)
    eng: object = None               # the slot's own context (its batch and plain launches)
    future: object = None            # completion work running in a worker thread
    reqs: list = field(default_factory=list)
    bid: int = -1
    total: int = 0
    src: object = None               # a plain launch's boxes (the request's own device array, or its pool run)
    general: bool = False            # a plain launch relaunched on the general path
    bin_target: int = 0              # the slot's configured bin target (0: the table's)


class Server:
    """Continuous-batching server on one GPU (see the module docstring).

    hbm_budget: bytes of device memory the server may hold: the box pool
    (pool_frac of the budget) plus the slots' workspaces and pair buffers
    (each slot's batches are capped at slot budget / bytes_per_box boxes; a
    single larger request still runs, alone).
    unbounded_bin_target: the bin target of batches of requests whose box
    sides are unbounded (Request.narrow False; their dense bins are re-binned
    by levels, which do better on smaller bins); 0 = the table's.
    """

    def __init__(self, device: int = 0, hbm_budget: int = 8 << 30, slots: int = 2, pool_frac: float = 0.5,
                 table: dict | None = None, bytes_per_box: int = 96, results: str = "host", consumer=None,
                 max_batch_boxes: int | None = None, host_levels: bool = False, group_by_schedule: bool = True,
                 unbounded_bin_target: int = 3000):
        import torch

        if results not in ("host", "device"):
            raise ValueError("results must be 'host' or 'device'")
        self.torch = torch
        self.dev = torch.device(f"cuda:{device}")
        torch.zeros(1, device=self.dev)  # torch's HIP runtime before the engine's context
        self.table = table or autotune.load_table()
        self.host_levels = host_levels
        self.group_by_schedule = group_by_schedule
        self.unbounded_bin_target = int(unbounded_bin_target)
        self.results = results
        self.consumer = consumer or keep_copy
        self.eng = pbig_amd.Engine(device)
        self._configure(self.eng, pbig_amd.PATH_AUTO)
        pool_boxes = int(hbm_budget * pool_frac) // 16
        self.pool = Pool(pool_boxes)
        self.d_pool = torch.empty((pool_boxes, 4), dtype=torch.float32, device=self.dev)
        self.slot_budget = int(hbm_budget * (1 - pool_frac)) // max(slots, 1)
        self.max_batch_boxes = max_batch_boxes or max(self.slot_budget // bytes_per_box, 1)
        self.copy_stream = torch.cuda.Stream(device=self.dev)
        self.slots = [_Slot(stream=torch.cuda.Stream(device=self.dev)) for _ in range(slots)]
        for sl in self.slots:
            sl.eng = pbig_amd.Engine(device)
            self._configure(sl.eng, pbig_amd.PATH_AUTO)
        self.pool_exec = ThreadPoolExecutor(max_workers=max(slots, 1))
        self.queue: list[Request] = []       # arrived, not admitted
        self.admitted: list[Request] = []    # upload issued, not yet batched
        self.upload_ev = None
        self.done: list[Request] = []
        self.n_batches = 0
        self.batch_sizes: list[int] = []
        self.n_requeued = 0                  # requests a batch left incomplete, rerun alone

    def _configure(self, eng, path: int, bin_target: int = 0):
        eng.configure(path, int(bin_target or self.table["bin_target"]), host_levels=self.host_levels)

    # -- client side ---------------------------------------------------------
    def submit(self, req: Request):
        if req.n > self.pool.capacity:
            raise ValueError(f"request {req.rid}: {req.n} boxes exceed the pool of {self.pool.capacity}")
        if req.cell_size <= 0:
            req.cell_size = float(autotune.cell_sizes([0, req.n], [req.L], self.table,
                                                      req.boxes if req.boxes is not None else None)[0])
        if req.narrow is None and req.boxes is not None:
            req.narrow = autotune.sides_bounded(req.boxes)
        req._alone = False
        self.queue.append(req)

    def close(self):
        self.pool_exec.shutdown(wait=True)
        self.torch.cuda.synchronize()
        for s in self.slots:
            if s.batch is not None:
                s.batch.close()
                s.batch = None
            if s.eng is not None:
                s.eng.close()
                s.eng = None
        self.eng.close()

    # -- server loop -----------------------------------------------------------
    def _admit(self, now: float) -> bool:
        torch = self.torch
        progressed = False
        # device arrays handed over are copied into the pool together, a few
        # multi-tensor copy launches for all the requests admitted here
        # instead of one copy launch each
        dst, src_dev = [], []
        while self.queue:
            r = self.queue[0]
            off = self.pool.alloc(r.n)
            if off < 0:
                break
            self.queue.pop(0)
            r._off = off
            r.t_admit = now
            if r.n:
                if r.device is not None and r.device.device == self.dev:
                    dst.append(self.d_pool[off:off + r.n])
                    src_dev.append(r.device)
                else:
                    src = r.device if r.device is not None else (
                        r.host if r.host is not None else torch.from_numpy(np.ascontiguousarray(r.boxes)))
                    with torch.cuda.stream(self.copy_stream):
                        self.d_pool[off:off + r.n].copy_(src, non_blocking=True)
            self.admitted.append(r)
            progressed = True
        if dst:
            with torch.cuda.stream(self.copy_stream):
                torch._foreach_copy_(dst, src_dev, non_blocking=True)
        if progressed:
            self.upload_ev = torch.cuda.Event()
            self.upload_ev.record(self.copy_stream)
        return progressed

    def _ensure_buffers(self, s: _Slot, k: int, n_boxes: int):
        torch = self.torch
        cap = max(4096, 2 * n_boxes)  # two pairs per box before the first regrowth
        if s.pairs is None or s.pairs.shape[0] < cap:
            s.pairs = torch.empty((cap, 2), dtype=torch.int32, device=self.dev)
            if self.results == "host":
                s.h_pairs = torch.empty((cap, 2), dtype=torch.int32, pin_memory=True)
        if s.io is None or s.io.shape[0] < k:
            kk = max(k, 256)
            s.io = torch.empty((kk, 2), dtype=torch.int64, device=self.dev)
            s.h_io = torch.empty((kk, 2), dtype=torch.int64, pin_memory=True)
        if s.res is None:
            s.res = torch.zeros(2, dtype=torch.int64, device=self.dev)
            s.h_res = torch.zeros(2, dtype=torch.int64, pin_memory=True)

    def _launch(self, s: _Slot, now: float):
        torch = self.torch
        reqs, rest, total = [], [], 0
        # Requests a batch left incomplete form a batch of their own: one
        # alone runs on the plain path, several are finished as a whole, and
        # none is split off again.
        redo = bool(self.admitted) and self.admitted[0]._alone
        # Like requests together: the test schedule is chosen per launch, and
        # one instance with wide boxes sends a whole batch to the general
        # schedule, so a batch takes the oldest request's kind (sides
        # bounded or not) and leaves the others for the next free slot.
        kind = None
        for r in self.admitted:
            if self.group_by_schedule and r._alone == redo:
                if kind is None:
                    kind = bool(r.narrow)
                if bool(r.narrow) != kind:
                    rest.append(r)
                    continue
            if r._alone != redo:
                rest.append(r)
            elif not reqs or total + r.n <= self.max_batch_boxes:
                reqs.append(r)
                total += r.n
            else:
                rest.append(r)
        self.admitted = rest
        starts = [r._off for r in reqs]
        counts = [r.n for r in reqs]
        Ls = [r.L for r in reqs]
        cells = [r.cell_size for r in reqs]
        self._ensure_buffers(s, len(reqs), total)
        s.mode = "plain" if len(reqs) == 1 and reqs[0].n > 0 else "batch"
        if s.mode == "plain":
            self._configure(s.eng, pbig_amd.PATH_AUTO)  # no memory of earlier requests
            s.bin_target = 0
            r = reqs[0]
            s.general = False
            if r.device is not None:  # the request's own device array: no wait for its pool copy
                s.src = r.device
            else:
                if self.upload_ev is not None:
                    s.stream.wait_event(self.upload_ev)
                s.src = self.d_pool[r._off:r._off + r.n]
            s.eng.check_device_async(s.src, r.L, r.cell_size, s.pairs, s.res, stream=s.stream)
        else:
            # batches choose their paths automatically (also after a request
            # that ended on the general path), at the bin target of their kind
            bt = self.unbounded_bin_target if self.group_by_schedule and kind is False else 0
            if s.general or bt != s.bin_target:
                self._configure(s.eng, pbig_amd.PATH_AUTO, bt)
                s.general = False
                s.bin_target = bt
            if s.batch is None:
                s.batch = pbig_amd.Batch.from_instances(s.eng, starts, counts, Ls, cells)
            else:
                s.batch.update(starts, counts, Ls, cells, stream=s.stream)
            if self.upload_ev is not None:
                s.stream.wait_event(self.upload_ev)
            s.batch.launch(self.d_pool, s.pairs, s.io, s.res, stream=s.stream)
        k = len(reqs)
        with torch.cuda.stream(s.stream):
            s.h_res.copy_(s.res, non_blocking=True)
            if s.mode == "batch":
                s.h_io[:k].copy_(s.io[:k], non_blocking=True)
        s.ev = torch.cuda.Event()
        s.ev.record(s.stream)
        for r in reqs:
            r.t_launch = now
        s.reqs = reqs
        s.state = "running"
        s.bid = self.n_batches
        self.n_batches += 1
        self.batch_sizes.append(k)

    def _deliver(self, s: _Slot, pairs, now: float):
        io = s.h_io[:len(s.reqs)].numpy()
        for r, (o, c) in zip(s.reqs, io):
            r.n_pairs = int(c)
            self.consumer(r, pairs[int(o):int(o) + int(c)], s.stream)
            r.t_done = now
            r.batch_id = s.bid
            self.pool.release(r._off, r.n)
            self.done.append(r)
        s.reqs = []
        s.state = "free"

    def _grow_pairs(self, s: _Slot, total: int):
        torch = self.torch
        cap = total + total // 4
        s.pairs = torch.empty((cap, 2), dtype=torch.int32, device=self.dev)
        if self.results == "host":
            s.h_pairs = torch.empty((cap, 2), dtype=torch.int32, pin_memory=True)

    def _finish(self, s: _Slot, err: int) -> int:
        """Completion of a launch whose result word asks for host work, run in
        a worker thread on the slot's own context and stream: the plain path's
        synchronous call (a dense request straight on the general path), or
        the batch's finish from its launch (re-binning of dense bins,
        workspace regrowth), then a rerun if the pairs did not fit."""
        torch = self.torch
        if s.mode == "plain":
            r = s.reqs[0]
            view = s.src
            if err & 16:
                self._configure(s.eng, pbig_amd.PATH_GENERAL)
            try:
                total = s.eng.check_device(view, r.L, r.cell_size, s.pairs, stream=s.stream, allow_overflow=True)
                if total > s.pairs.shape[0]:
                    self._grow_pairs(s, total)
                    total = s.eng.check_device(view, r.L, r.cell_size, s.pairs, stream=s.stream)
            finally:
                # the slot's later batches choose their paths automatically again
                self._configure(s.eng, pbig_amd.PATH_AUTO)
                s.bin_target = 0
            s.h_io[0, 0] = 0
            s.h_io[0, 1] = total
            return total
        total, io = s.batch.check(self.d_pool, s.pairs, allow_overflow=True, stream=s.stream, finish=True)
        if total > s.pairs.shape[0]:
            self._grow_pairs(s, total)
            total, io = s.batch.check(self.d_pool, s.pairs, stream=s.stream)
        s.h_io[:len(s.reqs)].copy_(torch.from_numpy(io.astype(np.int64)))
        return total

    def _relaunch_general(self, s: _Slot):
        torch = self.torch
        r = s.reqs[0]
        s.general = True
        self._configure(s.eng, pbig_amd.PATH_GENERAL)
        s.bin_target = 0
        s.eng.check_device_async(s.src, r.L, r.cell_size, s.pairs, s.res, stream=s.stream)
        with torch.cuda.stream(s.stream):
            s.h_res.copy_(s.res, non_blocking=True)
        s.ev = torch.cuda.Event()
        s.ev.record(s.stream)

    def _light(self) -> bool:
        """Light load: no more requests waiting than there are slots."""
        return len(self.queue) + len(self.admitted) <= len(self.slots)

    def _split(self, s: _Slot) -> bool:
        """A batch whose only trouble was bins beyond LDS: its requests
        without such a bin are complete and stay in the slot for delivery;
        the others go back to the head of the queue, to run as a batch of
        their own, instead of the whole batch waiting for their host work.
        False when every request is affected."""
        if any(r._alone for r in s.reqs):
            return False  # a batch of requests split off before: finished whole
        flags = s.batch.unresolved(stream=s.stream)
        if flags.all() or not flags.any():
            return False
        io = s.h_io[:len(s.reqs)].clone()
        keep = [i for i, f in enumerate(flags) if not f]
        back = [r for r, f in zip(s.reqs, flags) if f]
        s.h_io[:len(keep)].copy_(io[keep])
        s.reqs = [s.reqs[i] for i in keep]
        for r in back:
            r._alone = True
        self.admitted[:0] = back
        self.n_requeued += len(back)
        return True

    def _results_ready(self, s: _Slot, total: int, now: float) -> bool:
        torch = self.torch
        s.total = total
        if self.results == "device":
            self._deliver(s, s.pairs, now)
            return True
        if total:
            with torch.cuda.stream(s.stream):
                s.h_pairs[:total].copy_(s.pairs[:total], non_blocking=True)
        s.ev = torch.cuda.Event()
        s.ev.record(s.stream)
        s.state = "copying"
        return True

    def _progress(self, s: _Slot, now: float) -> bool:
        if s.state == "running" and s.ev.query():
            total, err = (int(v) for v in s.h_res.tolist())
            for r in s.reqs:
                r.t_ready = now
            if (err == _ERR_OVERFULL and s.mode == "batch" and len(s.reqs) > 1 and
                    total <= s.pairs.shape[0] and self._light() and self._split(s)):
                return self._results_ready(s, total, now)
            if (s.mode == "plain" and not s.general and err & _ERR_DENSE and not err & ~(_ERR_DENSE | _ERR_BINCAP)):
                # mostly dense clusters: the general path, asynchronously again
                self._relaunch_general(s)
                return True
            if err != 0 or total > s.pairs.shape[0]:
                s.future = self.pool_exec.submit(self._finish, s, err)
                s.state = "finishing"
                return True
            if s.mode == "plain":
                s.h_io[0, 0] = 0
                s.h_io[0, 1] = total
            return self._results_ready(s, total, now)
        if s.state == "finishing" and s.future.done():
            total = s.future.result()  # re-raises the worker's error
            s.future = None
            for r in s.reqs:
                r.t_ready = now
            return self._results_ready(s, total, now)
        if s.state == "copying" and s.ev.query():
            self._deliver(s, s.h_pairs.numpy().view(np.uint32), now)
            return True
        return False

    def poll(self, now: float) -> bool:
        progressed = self._admit(now)
        for s in self.slots:
            progressed |= self._progress(s, now)
        if self.admitted:
            for s in self.slots:
                if s.state == "free" and self.admitted:
                    self._launch(s, now)
                    progressed = True
        if progressed:
            progressed |= self._admit(now)
        return progressed

    def busy(self) -> bool:
        return bool(self.queue or self.admitted or any(s.state != "free" for s in self.slots))

    def run(self, requests: list[Request], realtime: bool = True) -> dict:
        """Serve requests arriving at their ``arrival`` times (seconds from
        the start; with realtime=False all arrive at once); returns
        throughput and latency figures."""
        order = sorted(requests, key=lambda r: r.arrival)
        self.done, self.batch_sizes, self.n_batches, self.n_requeued = [], [], 0, 0  # (this run's figures)
        t0 = time.perf_counter()
        k = 0
        while k < len(order) or self.busy():
            now = time.perf_counter() - t0
            while k < len(order) and (order[k].arrival <= now or not realtime):
                if not realtime:
                    order[k].arrival = now
                self.submit(order[k])
                k += 1
            progressed = self.poll(time.perf_counter() - t0)
            if not progressed and k < len(order) and realtime and not self.busy():
                wait = order[k].arrival - (time.perf_counter() - t0)
                if wait > 2e-4:
                    time.sleep(min(wait - 1e-4, 0.002))
            elif not progressed and any(s.state == "finishing" for s in self.slots):
                time.sleep(5e-5)  # let a worker thread take the interpreter lock
        t_end = time.perf_counter() - t0
        out = summarize(order, t_end, self.batch_sizes)
        out["requeued"] = self.n_requeued
        return out


def summarize(reqs: list[Request], t_end: float, batch_sizes=None) -> dict:
    lat = np.array([r.latency for r in reqs], dtype=np.float64) * 1e3
    boxes = sum(r.n for r in reqs)
    pairs = sum(r.n_pairs for r in reqs)
    span = t_end - min(r.arrival for r in reqs)
    out = {
        "requests": len(reqs),
        "boxes": boxes,
        "pairs": pairs,
        "makespan_s": span,
        "boxes_per_s": boxes / span,
        "pairs_per_s": pairs / span,
        "requests_per_s": len(reqs) / span,
        "latency_ms_p50": float(np.percentile(lat, 50)),
        "latency_ms_p99": float(np.percentile(lat, 99)),
        "latency_ms_mean": float(lat.mean()),
    }
    if batch_sizes:
        out["batches"] = len(batch_sizes)
        out["mean_requests_per_batch"] = float(np.mean(batch_sizes))
    return out


def serve_per_instance(eng, requests: list[Request], realtime: bool = True, results: str = "host",
                       consumer=None) -> dict:
    """The same stream served one request at a time by the plain path, the
    reference's calling pattern: (upload,) synchronous check, pairs back
    (to pinned host memory, or left on the device)."""
    import torch

    consumer = consumer or keep_copy
    dev = torch.device("cuda:0")
    order = sorted(requests, key=lambda r: r.arrival)
    cap = 1 << 20
    d_pairs = torch.empty((cap, 2), dtype=torch.int32, device=dev)
    h_pairs = torch.empty((cap, 2), dtype=torch.int32, pin_memory=True) if results == "host" else None
    stream = torch.cuda.current_stream()
    t0 = time.perf_counter()
    for r in order:
        now = time.perf_counter() - t0
        if realtime and r.arrival > now:
            while time.perf_counter() - t0 < r.arrival:
                pass
        elif not realtime:
            r.arrival = now
        if r.device is not None:
            d = r.device
        else:
            src = r.host if r.host is not None else torch.from_numpy(np.ascontiguousarray(r.boxes))
            d = src.to(dev, non_blocking=True)
        n = eng.check_device(d, r.L, r.cell_size, d_pairs, stream=stream, allow_overflow=True) if r.n else 0
        if n > cap:
            cap = n + n // 4
            d_pairs = torch.empty((cap, 2), dtype=torch.int32, device=dev)
            if results == "host":
                h_pairs = torch.empty((cap, 2), dtype=torch.int32, pin_memory=True)
            n = eng.check_device(d, r.L, r.cell_size, d_pairs, stream=stream)
        r.n_pairs = n
        if results == "host":
            if n:
                h_pairs[:n].copy_(d_pairs[:n], non_blocking=True)
                stream.synchronize()
            consumer(r, h_pairs.numpy().view(np.uint32)[:n], stream)
        else:
            consumer(r, d_pairs[:n], stream)
        r.t_done = time.perf_counter() - t0
    torch.cuda.synchronize()
    t_end = time.perf_counter() - t0
    return summarize(order, t_end)
