"""Autotuner of the binned path's two knobs.

* Fine cell side.  Expressed as boxes per cell, q = rho * c^2 (rho = boxes per
  unit area of the instance's domain), so that one value serves instances of
  any size and density.  q is tuned per size class: small instances pay more
  per-bin fixed cost and prefer other cells than large ones.  When the boxes
  are given, the cell is never narrower than the widest box of a sample (the
  first 1,000 boxes, as the reference's cell-size rule samples them,
  src/pbig/pbig.cu:1296-1308) if their sides are bounded (sample maximum at
  most 2.5 times the sample median): then every box spans at most two cells
  per axis and the narrow-box test schedule runs, whatever q is.
* Work tile: the number of boxes a coarse bin (the unit one test workgroup
  holds in LDS) aims for, the context's ``bin_target``.

``tune`` times candidate settings on a workload with HIP-graph replays of a
batch and keeps the fastest; ``save_table``/``load_table`` persist the choice
as JSON beside this file, and ``cell_sizes`` applies a table to a batch.  The
reference picks a power-of-two cell side from the first 1000 boxes
(src/pbig/pbig.cu:1285-1357) and has no tile knob; its rule remains available
as ``pbig_amd.reference_cell_size``.
"""
from __future__ import annotations

import json
import math
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE_PATH = os.path.join(HERE, "tuning.json")

# Used when no tuned table exists.  q = 1 is the cell side that measured best
# on the 4,194,304-box uniform instance (cell 32 at rho = 2^-10: boxes of side
# up to 32 span at most two cells per axis, so the narrow-box test kernel runs).
DEFAULT_TABLE = {
    "version": 1,
    "bin_target": 3000,
    "classes": [
        {"max_n": 4096, "boxes_per_cell": 1.0},
        {"max_n": 65536, "boxes_per_cell": 1.0},
        {"max_n": None, "boxes_per_cell": 1.0},
    ],
    "source": "default",
}


def load_table(path: str | None = None) -> dict:
    p = path or TABLE_PATH
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return json.loads(json.dumps(DEFAULT_TABLE))


def save_table(table: dict, path: str | None = None) -> str:
    p = path or TABLE_PATH
    with open(p, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
        f.write("\n")
    return p


def class_of(n: int, table: dict) -> int:
    for k, c in enumerate(table["classes"]):
        if c["max_n"] is None or n <= c["max_n"]:
            return k
    return len(table["classes"]) - 1


SAMPLE = 1000        # boxes sampled for the side floor
BOUNDED_RATIO = 2.5  # sample max / sample median at most this: the sides are bounded
HEADROOM = 1.1       # the floor's margin over the sample's widest side


def side_floor(boxes) -> float:
    """HEADROOM times the widest side among the first SAMPLE boxes when the
    sides look bounded (max <= BOUNDED_RATIO * median), else 0 (no floor)."""
    b = np.asarray(boxes[:SAMPLE], dtype=np.float32).reshape(-1, 4)
    if b.shape[0] == 0:
        return 0.0
    side = np.maximum(b[:, 1] - b[:, 0], b[:, 3] - b[:, 2]).astype(np.float64)
    mx, med = float(side.max()), float(np.median(side))
    return HEADROOM * mx if mx <= BOUNDED_RATIO * med else 0.0


def sides_bounded(boxes) -> bool:
    """The first SAMPLE boxes' sides look bounded (see side_floor): at the
    floored cell every box spans at most two cells per axis, and a batch of
    such instances runs the narrow-box test schedule."""
    return side_floor(boxes) > 0.0


def cell_side(n: int, L: float, q: float, floor: float = 0.0) -> float:
    """Cell side holding q box origins on average, sqrt(q L^2 / n), and at
    least `floor`, rounded to a float32 value; at least L / 4096 and at most
    L."""
    if n <= 0:
        return float(np.float32(max(L, 1.0)))
    c = max(math.sqrt(q * L * L / n), floor)
    c = min(max(c, L / 4096.0, 1e-3), max(L, 1e-3))
    return float(np.float32(c))


def cell_sizes(offsets, ranges, table: dict | None = None, boxes=None) -> np.ndarray:
    """Per-instance fine cell sides for a batch (offsets: k+1 box offsets;
    boxes, optional: the concatenated boxes, for the side floor)."""
    t = table or load_table()
    offsets = np.asarray(offsets, dtype=np.int64)
    n = np.diff(offsets)
    ranges = np.asarray(ranges, dtype=np.float64)
    out = np.empty(n.shape[0], dtype=np.float32)
    for i in range(n.shape[0]):
        q = t["classes"][class_of(int(n[i]), t)]["boxes_per_cell"]
        fl = side_floor(boxes[offsets[i]:offsets[i + 1]]) if boxes is not None else 0.0
        out[i] = cell_side(int(n[i]), float(ranges[i]), q, fl)
    return out


def time_batch(engine, d_boxes, starts, counts, ranges, cells, reps: int = 5) -> float:
    """Milliseconds per HIP-graph replay of the batch whose instance i is
    d_boxes[starts[i] : starts[i] + counts[i]] (the workspace is sized by one
    synchronous check first)."""
    import torch

    import pbig_amd

    with pbig_amd.Batch.from_instances(engine, starts, counts, ranges, cells) as b:
        cap = max(1024, int(np.sum(counts)))
        pairs = torch.empty((cap, 2), dtype=torch.int32, device=d_boxes.device)
        n, _ = b.check(d_boxes, pairs, allow_overflow=True)
        if n > cap:
            pairs = torch.empty((n, 2), dtype=torch.int32, device=d_boxes.device)
            b.check(d_boxes, pairs)
        res = torch.zeros(2, dtype=torch.int64, device=d_boxes.device)
        b.capture(d_boxes, pairs, None, res)
        stream = torch.cuda.current_stream()
        b.replay(stream)
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record(stream)
        for _ in range(reps):
            b.replay(stream)
        e.record(stream)
        torch.cuda.synchronize()
        r = res.cpu().tolist()
        if r[1] != 0 or r[0] != n:
            raise RuntimeError(f"tuning replay reported {r} (expected [{n}, 0])")
        return s.elapsed_time(e) / reps


def tune(engine, boxes, offsets, ranges, qs=(0.25, 0.5, 1.0, 2.0, 4.0), bin_targets=(1000, 2000, 3000, 4096),
         reps: int = 5, base: dict | None = None, log=print) -> dict:
    """Pick boxes-per-cell per size class, then the bin target, by timing
    graph replays of the workload's instances on the GPU."""
    import torch

    t = json.loads(json.dumps(base or DEFAULT_TABLE))
    offsets = np.asarray(offsets, dtype=np.int64)
    ranges = np.asarray(ranges, dtype=np.float64)
    n = np.diff(offsets)
    d = torch.from_numpy(np.ascontiguousarray(boxes, dtype=np.float32)).to("cuda:0")
    measured = {"classes": [], "bin_target": {}}
    floors = np.array([side_floor(boxes[offsets[i]:offsets[i + 1]]) for i in range(n.shape[0])])
    engine.configure(0, t["bin_target"])
    for k, cls in enumerate(t["classes"]):
        idx = np.array([i for i in range(n.shape[0]) if class_of(int(n[i]), t) == k], dtype=np.int64)
        if idx.size == 0:
            measured["classes"].append({})
            continue
        times = {}
        for q in qs:
            cells = np.array([cell_side(int(n[i]), float(ranges[i]), q, floors[i]) for i in idx], dtype=np.float32)
            times[q] = time_batch(engine, d, offsets[idx], n[idx], ranges[idx], cells, reps)
            log(f"autotune class {k} ({idx.size} instances): q={q} -> {times[q]:.3f} ms")
        cls["boxes_per_cell"] = float(min(times, key=times.get))
        measured["classes"].append({str(q): v for q, v in times.items()})
    cells = cell_sizes(offsets, ranges, t, boxes)
    for bt in bin_targets:
        engine.configure(0, int(bt))
        measured["bin_target"][str(bt)] = time_batch(engine, d, offsets[:-1], n, ranges, cells, reps)
        log(f"autotune bin_target {bt} -> {measured['bin_target'][str(bt)]:.3f} ms")
    t["bin_target"] = int(min(bin_targets, key=lambda b: measured["bin_target"][str(b)]))
    engine.configure(0, t["bin_target"])
    t["measured_ms"] = measured
    t["source"] = f"tuned {time.strftime('%Y-%m-%d')} on {torch.cuda.get_device_name(0)}"
    return t


def main():
    """python -m pbig_amd.autotune: tune on the 4,096-instance workload and
    write the table."""
    import argparse

    import pbig_amd
    from pbig_amd import workloads

    ap = argparse.ArgumentParser()
    ap.add_argument("--instances", type=int, default=4096)
    ap.add_argument("--out", default=TABLE_PATH)
    args = ap.parse_args()
    import torch

    boxes, offsets, ranges = workloads.batched(args.instances, workloads.SEEDS["c4"])
    torch.zeros(1, device="cuda:0")  # torch's HIP runtime first, then the engine's context
    eng = pbig_amd.Engine(0)
    t = tune(eng, boxes, offsets, ranges)
    print("wrote", save_table(t, args.out))
    eng.close()


if __name__ == "__main__":
    main()
