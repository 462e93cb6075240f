"""CPU oracle for the pairwise box intersection path (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package, and only as the checker.  The product path
(``pbig_amd``) never imports it.

The C++ restatement lives in ``pbig_oracle.cpp`` (built by ``Makefile`` into
``liboracle.so``); this module is a ctypes wrapper over it plus numpy helpers
for sorting and hashing pair sets.

Pinning (each function below is checked in ``tests/test_oracle.py``):
  * ``brute_pairs``  -- reference checker predicate, src/pbig/pbig.cu:996-1001;
    pinned by hand-built known answers (touching edges, nesting, corners).
  * ``sweep_pairs``  -- an independent algorithm; must equal ``brute_pairs``.
  * ``grid_pairs``   -- the reference's grid pipeline (pbig.cu:174-221,
    321-427, 562-603, 642-932); must equal ``brute_pairs``/``sweep_pairs``.
  * ``decide_cell_size`` -- pbig.cu:1285-1357; pinned by the reference's own
    trace (benchmarks/test3/filter_file:5-13 -> cs = 128) and by the closed
    form of the loop.
  * ``ref_generate`` + ``grid_u32`` -- the reference's 3-D unsigned mode and
    generator (src/pbig/main.cc:204-335); pinned by the reference's logged
    answers (benchmarks/test1/filter_file:7 -> 49,514,252 pairs;
    benchmarks/test3/filter_file:17-49 -> cs, cells, memberships, max per
    cell, jobs, compared pairs and 1,121,815,231 pairs).
  * ``brute_u32`` -- the same predicate by brute force; must equal
    ``grid_u32``'s count on in-range inputs.
  * ``brute_count`` -- closed forms (k identical boxes, a chain, a lattice),
    the 16k instance's 10,409 pairs and ``len(brute_pairs)`` on random
    inputs, on one thread and several.
  * ``grid_counts`` / ``grid_stats`` -- a geometric count (a closed box
    touches a half-open cell, border cells extended for clamping), on
    integer boxes that end exactly on cell edges and stick out of the
    domain; ``grid_pairs``' own statistics must agree.
  * ``decide_cell_size`` also on constructed inputs where a float mean, or
    an int sum divided in float, changes the answer (64 against 128).
"""
from __future__ import annotations

import ctypes
import hashlib
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class GridStats(ctypes.Structure):
    _fields_ = [
        ("n_cells", ctypes.c_uint64),
        ("memberships", ctypes.c_uint64),
        ("max_per_cell", ctypes.c_uint64),
        ("max_cell", ctypes.c_uint64),
        ("compared_pairs", ctypes.c_uint64),
        ("jobs256", ctypes.c_uint64),
        ("nonempty_cells", ctypes.c_uint64),
        ("pairs", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    P = ctypes.c_void_p
    u32, u64, i32, f64 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_double
    sig = {
        "orc_overlap2f": (i32, [P, P]),
        "orc_brute2f": (u64, [P, u32, P, u64, i32]),
        "orc_brute2f_count": (u64, [P, u32, i32]),
        "orc_sweep2f": (u64, [P, u32, P, u64, i32]),
        "orc_decide_cell_size2f": (u32, [P, u32, f64]),
        "orc_grid2f": (u64, [P, u32, f64, u32, P, u64, i32, P]),
        "orc_grid2f_stats": (None, [P, u32, f64, u32, P]),
        "orc_grid2f_counts": (None, [P, u32, f64, u32, P]),
        "orc_cellmap": (u32, [ctypes.c_float, ctypes.c_float, u32]),
        "orc_ref_generate": (None, [u32, u32, u32, ctypes.c_char_p, i32, P]),
        "orc_decide_cell_size_u32": (u32, [P, u32, i32, u32]),
        "orc_grid_u32": (u64, [P, u32, i32, u32, u32, i32, P]),
        "orc_brute_u32": (u64, [P, u32, i32, P, u64, i32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _boxes(boxes) -> np.ndarray:
    b = np.ascontiguousarray(boxes, dtype=np.float32)
    if b.ndim != 2 or b.shape[1] != 4:
        raise ValueError("boxes must be an (n, 4) float32 array {xmin,xmax,ymin,ymax}")
    return b


def _run_pairs(fn, b, *extra):
    n = b.shape[0]
    cnt = fn(b.ctypes.data, n, None, 0, *extra)
    out = np.empty((max(int(cnt), 1), 2), dtype=np.uint32)
    cnt2 = fn(b.ctypes.data, n, out.ctypes.data, int(cnt), *extra)
    assert cnt2 == cnt
    return out[: int(cnt)]


def overlap(a, b) -> bool:
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    return bool(_load().orc_overlap2f(a.ctypes.data, b.ctypes.data))


def brute_pairs(boxes, threads: int = 0) -> np.ndarray:
    """Every overlapping pair (lo, hi) by O(n^2) brute force, unsorted."""
    b = _boxes(boxes)
    return _run_pairs(_load().orc_brute2f, b, threads)


def brute_count(boxes, threads: int = 0) -> int:
    b = _boxes(boxes)
    return int(_load().orc_brute2f_count(b.ctypes.data, b.shape[0], threads))


def sweep_pairs(boxes, threads: int = 0) -> np.ndarray:
    b = _boxes(boxes)
    return _run_pairs(_load().orc_sweep2f, b, threads)


def decide_cell_size(boxes, range_: float) -> int:
    b = _boxes(boxes)
    return int(_load().orc_decide_cell_size2f(b.ctypes.data, b.shape[0], float(range_)))


def grid_pairs(boxes, range_: float, cell_size: int, threads: int = 0, with_stats: bool = False):
    """The reference's grid algorithm restated on the CPU; pairs (lo, hi) unsorted."""
    b = _boxes(boxes)
    lib = _load()
    st = GridStats()
    n = b.shape[0]
    cnt = lib.orc_grid2f(b.ctypes.data, n, float(range_), int(cell_size), None, 0, threads, None)
    out = np.empty((max(int(cnt), 1), 2), dtype=np.uint32)
    lib.orc_grid2f(b.ctypes.data, n, float(range_), int(cell_size), out.ctypes.data, int(cnt),
                   threads, ctypes.byref(st))
    out = out[: int(cnt)]
    return (out, st.as_dict()) if with_stats else out


def grid_stats(boxes, range_: float, cell_size: int) -> dict:
    b = _boxes(boxes)
    st = GridStats()
    _load().orc_grid2f_stats(b.ctypes.data, b.shape[0], float(range_), int(cell_size), ctypes.byref(st))
    return st.as_dict()


def grid_counts(boxes, range_: float, cell_size: int) -> np.ndarray:
    """Per-cell membership counts (row-major cells, nrc*cy + cx)."""
    b = _boxes(boxes)
    nrc = max(1, int(np.ceil(float(range_) / float(cell_size))))
    out = np.empty(nrc * nrc, dtype=np.uint32)
    _load().orc_grid2f_counts(b.ctypes.data, b.shape[0], float(range_), int(cell_size), out.ctypes.data)
    return out


def cellmap(v: float, inv: float, nrc: int) -> int:
    return int(_load().orc_cellmap(float(v), float(inv), int(nrc)))


def ref_generate(n: int, range_: int, size: int, simcase: str, dims: int = 3) -> np.ndarray:
    """The reference driver's glibc rand() generator (src/pbig/main.cc:204-335)."""
    out = np.empty((n, 2 * dims), dtype=np.uint32)
    _load().orc_ref_generate(n, range_, size, simcase.encode(), dims, out.ctypes.data)
    return out


def brute_u32(boxes: np.ndarray, dims: int, threads: int = 0) -> np.ndarray:
    """Every overlapping pair (lo, hi) of unsigned boxes (n, 2*dims) by brute force."""
    b = np.ascontiguousarray(boxes, dtype=np.uint32).reshape(-1, 2 * dims)
    lib = _load()
    n = b.shape[0]
    cnt = lib.orc_brute_u32(b.ctypes.data, n, dims, None, 0, threads)
    out = np.empty((max(int(cnt), 1), 2), dtype=np.uint32)
    lib.orc_brute_u32(b.ctypes.data, n, dims, out.ctypes.data, int(cnt), threads)
    return out[: int(cnt)]


def decide_cell_size_u32(boxes: np.ndarray, dims: int, range_: int) -> int:
    b = np.ascontiguousarray(boxes, dtype=np.uint32)
    return int(_load().orc_decide_cell_size_u32(b.ctypes.data, b.shape[0], dims, range_))


def grid_u32(boxes: np.ndarray, dims: int, range_: int, cell_size: int, threads: int = 0) -> dict:
    b = np.ascontiguousarray(boxes, dtype=np.uint32)
    st = GridStats()
    _load().orc_grid_u32(b.ctypes.data, b.shape[0], dims, range_, cell_size, threads, ctypes.byref(st))
    return st.as_dict()


# ---------------------------------------------------------------------------
# Pair-set helpers
# ---------------------------------------------------------------------------

def sort_pairs(pairs) -> np.ndarray:
    """Normalise to (lo, hi) and sort lexicographically; returns (k, 2) uint32."""
    p = np.asarray(pairs, dtype=np.uint32).reshape(-1, 2)
    lo = np.minimum(p[:, 0], p[:, 1]).astype(np.uint64)
    hi = np.maximum(p[:, 0], p[:, 1]).astype(np.uint64)
    key = np.sort((lo << np.uint64(32)) | hi)
    out = np.empty((key.shape[0], 2), dtype=np.uint32)
    out[:, 0] = (key >> np.uint64(32)).astype(np.uint32)
    out[:, 1] = (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    return out


def pair_hash(sorted_pairs) -> str:
    """SHA-256 of the sorted (lo, hi) uint32 little-endian pair list."""
    p = np.ascontiguousarray(sorted_pairs, dtype="<u4")
    return hashlib.sha256(p.tobytes()).hexdigest()
