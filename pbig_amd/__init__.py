"""pbig_amd -- MI355X-native pairwise box intersection engine.

Python host binding over the C-ABI of ``include/pbig_hip.h`` (ctypes).  The
compute path is the HIP library ``libpbig_hip.so`` built by
``pbig_amd/build.py``; importing this module never falls back to a CPU
implementation: if the library is missing, every entry point raises.

Names follow the reference API (src/pbig/pbig.h:8-11): ``Engine`` is the
handle that ``pbig_init``/``pbig_finalize`` manage in the reference, and
``Engine.check*`` is ``pbig_check``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpbig_hip.so")

PBIG_OK = 0
PBIG_E_ARG = -1
PBIG_E_INVALID_BOX = -2
PBIG_E_OVERFLOW = -3
PBIG_E_HIP = -4
PBIG_E_NOMEM = -5
PBIG_E_TOO_DENSE = -6

# The C-ABI entry points declared in include/pbig_hip.h.
EXPORTS = (
    "pbig_hip_create",
    "pbig_hip_destroy",
    "pbig_hip_last_error",
    "pbig_hip_reference_cell_size",
    "pbig_hip_check_device",
    "pbig_hip_check_device_async",
    "pbig_hip_check_device_noalloc",
    "pbig_hip_reserve",
    "pbig_hip_check_host",
    "pbig_hip_check_host_each",
    "pbig_hip_last_stats",
    "pbig_hip_timing_start",
    "pbig_hip_timing_read",
    "pbig_hip_configure",
    "pbig_hip_batch_create",
    "pbig_hip_batch_destroy",
    "pbig_hip_batch_update",
    "pbig_hip_batch_check",
    "pbig_hip_batch_finish",
    "pbig_hip_batch_unresolved",
    "pbig_hip_batch_launch",
    "pbig_hip_batch_capture",
    "pbig_hip_batch_replay",
    "pbig_hip_batch_info",
    "pbig_hip_reference_cell_size_u32",
    "pbig_hip_check_u32",
    "pbig_hip_check_host_u32",
    "pbig_hip_check_host_u32_each",
    "pbig_hip_rice_encode",
    "pbig_hip_rice_decode",
)

# pbig_hip_pairs_fn: (user, pairs {lo, hi} uint32 x n, n), pairs valid during the call.
PAIRS_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint64)

PATH_AUTO, PATH_GENERAL, PATH_BINNED = 0, 1, 2
# pbig_hip_timing_read also reports the binned path's single-pass launch of
# one instance (no histogram or scan pass) as path 3.
TIMING_ONEPASS = 3
# ... and the two-pass launch followed by re-binning levels planned on the
# device as path 4 (one duration per level).
TIMING_DEEP = 4
# ... and the general path's tile test on a cell layout built from the bins
# (the binned count, scan and scatter, then a per-bin fine-cell sort) as
# path 5.
TIMING_SORTED = 5
PATH_NAMES = {PATH_GENERAL: "general", PATH_BINNED: "binned", TIMING_ONEPASS: "onepass", TIMING_DEEP: "deep",
              TIMING_SORTED: "sorted"}
# Kernels of each path in launch order (pbig_hip_timing_read).
KERNELS = {
    PATH_GENERAL: ("clear", "count", "scan_reduce", "scan_apply", "scatter", "test"),
    PATH_BINNED: ("bin_hist", "bin_scan", "bin_starts", "bin_scatter", "bin_test", "bin_pair_scan",
                  "bin_compact"),
    TIMING_ONEPASS: ("bin_place", "bin_test"),
    TIMING_DEEP: ("level0", "level1", "level2", "level3"),
    TIMING_SORTED: ("clear", "bin_hist", "bin_scan", "bin_starts", "bin_scatter", "fine_count", "fine_place",
                    "test"),
}

# numpy view of the reference's RectInfo (src/pbig/rect.h:15-18, CType=float)
RECT_DTYPE = np.dtype([("xmin", "<f4"), ("xmax", "<f4"), ("ymin", "<f4"), ("ymax", "<f4"),
                       ("rid", "<u4")])


class PbigError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"pbig status {status}: {message}")
        self.status = status


class Stats(ctypes.Structure):
    _fields_ = [
        ("cell_size_num", ctypes.c_uint32),
        ("cell_size_den", ctypes.c_uint32),
        ("cells_per_row", ctypes.c_uint32),
        ("reserved0", ctypes.c_uint32),
        ("n_boxes", ctypes.c_uint64),
        ("n_cells", ctypes.c_uint64),
        ("memberships", ctypes.c_uint64),
        ("tile_pairs", ctypes.c_uint64),
        ("n_pairs", ctypes.c_uint64),
        ("max_per_cell", ctypes.c_uint64),
        ("path", ctypes.c_int32),
        ("bins_per_row", ctypes.c_uint32),
        ("cells_per_bin_side", ctypes.c_uint32),
        ("reserved1", ctypes.c_uint32),
    ]

    def as_dict(self):
        d = {name: int(getattr(self, name)) for name, _ in self._fields_ if not name.startswith("reserved")}
        d["cell_size"] = d["cell_size_num"] / max(d["cell_size_den"], 1)
        return d


class Instance(ctypes.Structure):
    """pbig_hip_instance: one instance of a batch."""
    _fields_ = [("box_offset", ctypes.c_uint64), ("n_boxes", ctypes.c_uint32),
                ("cell_size", ctypes.c_float), ("range", ctypes.c_double)]


class Config(ctypes.Structure):
    _fields_ = [("path", ctypes.c_int32), ("bin_target", ctypes.c_uint32),
                ("levels_min_boxes", ctypes.c_uint32), ("host_levels", ctypes.c_uint32)]


_lib = None


def library_path() -> str:
    return LIB_PATH


def load_library(path: str | None = None):
    """Load libpbig_hip.so (no compute, no GPU needed) and declare signatures."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(p)
    P, u32, u64, i32, f32, f64 = (ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
                                  ctypes.c_float, ctypes.c_double)
    sig = {
        "pbig_hip_create": (i32, [ctypes.POINTER(P), i32]),
        "pbig_hip_destroy": (i32, [P]),
        "pbig_hip_last_error": (ctypes.c_char_p, [P]),
        "pbig_hip_reference_cell_size": (u32, [P, u32, f64]),
        "pbig_hip_check_device": (i32, [P, P, u32, f64, f32, P, u64, ctypes.POINTER(u64), P]),
        "pbig_hip_check_device_async": (i32, [P, P, u32, f64, f32, P, u64, P, P]),
        "pbig_hip_check_device_noalloc": (i32, [P, P, u32, f64, f32, P, u64, P, P]),
        "pbig_hip_reserve": (i32, [P, u64, u64, u64]),
        "pbig_hip_check_host": (i32, [P, P, u32, f64, f32, P, u64, ctypes.POINTER(u64)]),
        "pbig_hip_check_host_each": (i32, [P, P, u32, f64, f32, PAIRS_FN, P, ctypes.POINTER(u64)]),
        "pbig_hip_last_stats": (i32, [P, ctypes.POINTER(Stats)]),
        "pbig_hip_timing_start": (i32, [P, i32]),
        "pbig_hip_timing_read": (i32, [P, ctypes.POINTER(ctypes.c_double), i32, ctypes.POINTER(i32),
                                       ctypes.POINTER(i32)]),
        "pbig_hip_configure": (i32, [P, ctypes.POINTER(Config)]),
        "pbig_hip_batch_create": (i32, [P, ctypes.POINTER(Instance), u32, ctypes.POINTER(P)]),
        "pbig_hip_batch_destroy": (i32, [P]),
        "pbig_hip_batch_update": (i32, [P, ctypes.POINTER(Instance), u32, P]),
        "pbig_hip_batch_check": (i32, [P, P, u64, P, u64, P, ctypes.POINTER(u64), P]),
        "pbig_hip_batch_finish": (i32, [P, P, u64, P, u64, P, ctypes.POINTER(u64), P]),
        "pbig_hip_batch_unresolved": (i32, [P, P, ctypes.POINTER(u32), P]),
        "pbig_hip_batch_launch": (i32, [P, P, u64, P, u64, P, P, P]),
        "pbig_hip_batch_capture": (i32, [P, P, u64, P, u64, P, P]),
        "pbig_hip_batch_replay": (i32, [P, P]),
        "pbig_hip_batch_info": (i32, [P, ctypes.POINTER(u64), ctypes.POINTER(u32), ctypes.POINTER(u32)]),
        "pbig_hip_reference_cell_size_u32": (u32, [P, u32, i32, u32]),
        "pbig_hip_check_u32": (i32, [P, P, u32, i32, u32, u32, P, u64, ctypes.POINTER(u64), P]),
        "pbig_hip_check_host_u32": (i32, [P, P, u32, i32, u32, u32, P, u64, ctypes.POINTER(u64)]),
        "pbig_hip_check_host_u32_each": (i32, [P, P, u32, i32, u32, u32, PAIRS_FN, P, ctypes.POINTER(u64)]),
        "pbig_hip_rice_encode": (i32, [P, u64, u32, P, u64, ctypes.POINTER(u64), P, u64, ctypes.POINTER(u64), P,
                                       u64, ctypes.POINTER(u64)]),
        "pbig_hip_rice_decode": (i32, [P, u64, P, u64, P, u64, P, u64, ctypes.POINTER(u64)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def reference_cell_size(boxes: np.ndarray, range_: float) -> int:
    """The reference's decideCellSize (src/pbig/pbig.cu:1285-1357), via the C-ABI."""
    b = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 4)
    return int(load_library().pbig_hip_reference_cell_size(b.ctypes.data, b.shape[0], float(range_)))


def reference_cell_size_u32(boxes: np.ndarray, dims: int, range_: int) -> int:
    """The reference's decideCellSize for its unsigned build, via the C-ABI."""
    b = np.ascontiguousarray(boxes, dtype=np.uint32).reshape(-1, 2 * dims)
    return int(load_library().pbig_hip_reference_cell_size_u32(b.ctypes.data, b.shape[0], int(dims), int(range_)))


def rice_encode(pairs: np.ndarray, n_boxes: int):
    """(k, 2) uint32 pairs -> (words uint64, cell_start uint32, cell_recs
    uint32): the reference's compressed record stream and its cell arrays."""
    lib = load_library()
    p = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
    nw, nc, nr = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    rc = lib.pbig_hip_rice_encode(p.ctypes.data if len(p) else None, len(p), int(n_boxes), None, 0, ctypes.byref(nw),
                                  None, 0, ctypes.byref(nc), None, 0, ctypes.byref(nr))
    if rc not in (PBIG_OK, PBIG_E_OVERFLOW):
        raise PbigError(rc, "rice_encode: bad pairs")
    words = np.zeros(max(nw.value, 1), dtype=np.uint64)
    start = np.zeros(nc.value + 1, dtype=np.uint32)
    recs = np.zeros(max(nr.value, 1), dtype=np.uint32)
    rc = lib.pbig_hip_rice_encode(p.ctypes.data if len(p) else None, len(p), int(n_boxes), words.ctypes.data,
                                  len(words), ctypes.byref(nw), start.ctypes.data, len(start), ctypes.byref(nc),
                                  recs.ctypes.data, len(recs), ctypes.byref(nr))
    if rc != PBIG_OK:
        raise PbigError(rc, "rice_encode failed")
    return words[: nw.value], start, recs[: nr.value]


def rice_decode(words: np.ndarray, cell_start: np.ndarray, cell_recs: np.ndarray) -> np.ndarray:
    """The stream and its cell arrays -> (k, 2) uint32 pairs (lo, hi)."""
    lib = load_library()
    w = np.ascontiguousarray(words, dtype=np.uint64)
    st = np.ascontiguousarray(cell_start, dtype=np.uint32)
    rc_ = np.ascontiguousarray(cell_recs, dtype=np.uint32)
    n = ctypes.c_uint64()
    rc = lib.pbig_hip_rice_decode(w.ctypes.data if len(w) else None, len(w), st.ctypes.data, max(len(st) - 1, 0),
                                  rc_.ctypes.data, len(rc_), None, 0, ctypes.byref(n))
    if rc not in (PBIG_OK, PBIG_E_OVERFLOW):
        raise PbigError(rc, "rice_decode: malformed stream")
    out = np.empty((max(n.value, 1), 2), dtype=np.uint32)
    rc = lib.pbig_hip_rice_decode(w.ctypes.data if len(w) else None, len(w), st.ctypes.data, max(len(st) - 1, 0),
                                  rc_.ctypes.data, len(rc_), out.ctypes.data, n.value, ctypes.byref(n))
    if rc != PBIG_OK:
        raise PbigError(rc, "rice_decode failed")
    return out[: n.value]


def _ptr(t) -> int:
    """Device pointer of a torch tensor (torch is only memory plumbing here)."""
    return int(t.data_ptr())


NULL_STREAM = (1 << 64) - 1  # PBIG_HIP_NULL_STREAM


def _stream_ptr(stream) -> int | None:
    """None -> the context's stream; a torch stream (or raw handle) -> that
    stream, where handle 0 (torch's default stream) is HIP's null stream."""
    if stream is None:
        return None
    h = int(getattr(stream, "cuda_stream", stream))
    return h if h else NULL_STREAM


class Engine:
    """One HIP context of the engine (the reference's pbig_init .. pbig_finalize)."""

    def __init__(self, device: int = 0):
        self._lib = load_library()
        h = ctypes.c_void_p()
        rc = self._lib.pbig_hip_create(ctypes.byref(h), int(device))
        if rc != PBIG_OK:
            raise PbigError(rc, f"pbig_hip_create(device={device}) failed")
        self._h = h
        self.device = device

    # -- lifecycle ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.pbig_hip_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _raise(self, rc: int):
        msg = self._lib.pbig_hip_last_error(self._h)
        raise PbigError(rc, msg.decode() if msg else "")

    def last_error(self) -> str:
        msg = self._lib.pbig_hip_last_error(self._h)
        return msg.decode() if msg else ""

    def stats(self) -> dict:
        st = Stats()
        self._lib.pbig_hip_last_stats(self._h, ctypes.byref(st))
        return st.as_dict()

    def reserve(self, n_boxes: int, n_cells: int, n_memberships: int):
        rc = self._lib.pbig_hip_reserve(self._h, n_boxes, n_cells, n_memberships)
        if rc != PBIG_OK:
            self._raise(rc)

    def timing_start(self, max_calls: int):
        rc = self._lib.pbig_hip_timing_start(self._h, int(max_calls))
        if rc != PBIG_OK:
            self._raise(rc)

    def timing_read(self) -> tuple[dict, int, str]:
        """Summed device milliseconds per kernel over the recorded calls, the
        number of calls and the path they took."""
        arr = (ctypes.c_double * 8)()
        calls = ctypes.c_int(0)
        path = ctypes.c_int(0)
        rc = self._lib.pbig_hip_timing_read(self._h, arr, 8, ctypes.byref(calls), ctypes.byref(path))
        if rc != PBIG_OK:
            self._raise(rc)
        names = KERNELS.get(path.value, ())
        return {k: float(arr[i]) for i, k in enumerate(names)}, int(calls.value), PATH_NAMES.get(path.value, "")

    def configure(self, path: int = PATH_AUTO, bin_target: int = 0, levels_min_boxes: int = 0,
                  host_levels: bool = False):
        """Path choice and tuning (pbig_hip_configure): bin_target 0 = 2000;
        levels_min_boxes 0 = 2^20 (the automatic choice's smallest instance
        whose bins beyond LDS are re-binned on the device; 64 times that for
        an instance of mostly dense clusters); host_levels: bins beyond LDS
        always re-binned by the host, never by device levels."""
        cfg = Config(int(path), int(bin_target), int(levels_min_boxes), int(bool(host_levels)))
        rc = self._lib.pbig_hip_configure(self._h, ctypes.byref(cfg))
        if rc != PBIG_OK:
            self._raise(rc)

    # -- plain path ---------------------------------------------------------
    def check_device(self, boxes, range_: float, cell_size: float, pairs, stream=None,
                     allow_overflow: bool = False) -> int:
        """boxes: (n, 4) float32 device tensor; pairs: (cap, 2) int32/uint32
        device tensor.  Returns the number of pairs (synchronous)."""
        n = int(boxes.shape[0])
        cap = int(pairs.shape[0]) if pairs is not None else 0
        out = ctypes.c_uint64(0)
        rc = self._lib.pbig_hip_check_device(self._h, _ptr(boxes) if n else None, n, float(range_),
                                             float(cell_size), _ptr(pairs) if cap else None, cap,
                                             ctypes.byref(out), _stream_ptr(stream))
        if rc == PBIG_E_OVERFLOW and allow_overflow:
            return int(out.value)
        if rc != PBIG_OK:
            self._raise(rc)
        return int(out.value)

    def check_device_async(self, boxes, range_: float, cell_size: float, pairs, result, stream=None):
        """Enqueue only.  result: (2,) int64 device tensor <- {n_pairs, error}."""
        n = int(boxes.shape[0])
        cap = int(pairs.shape[0])
        rc = self._lib.pbig_hip_check_device_async(self._h, _ptr(boxes) if n else None, n, float(range_),
                                                   float(cell_size), _ptr(pairs), cap, _ptr(result),
                                                   _stream_ptr(stream))
        if rc != PBIG_OK:
            self._raise(rc)

    def check_device_noalloc(self, boxes, range_: float, cell_size: float, pairs, result, stream=None):
        """Enqueue only, with no device allocation: the library's fixed scratch
        array holds the workspace (up to about 2^20 boxes).  result: (2,)
        int64 device tensor <- {n_pairs, error}."""
        n = int(boxes.shape[0])
        cap = int(pairs.shape[0])
        rc = self._lib.pbig_hip_check_device_noalloc(self._h, _ptr(boxes) if n else None, n, float(range_),
                                                     float(cell_size), _ptr(pairs), cap, _ptr(result),
                                                     _stream_ptr(stream))
        if rc != PBIG_OK:
            self._raise(rc)

    # -- the reference's unsigned build (CType = unsigned int, DIMS 2 or 3) ---
    def check_u32(self, boxes, dims: int, range_: int, cell_size: int = 0, pairs=None, stream=None,
                  allow_overflow: bool = False) -> int:
        """boxes: (n, 2*dims) int32/uint32 device tensor {xmin,xmax,ymin,ymax[,zmin,zmax]};
        pairs: (cap, 2) device tensor, or None to count only.  Returns the pair count."""
        n = int(boxes.shape[0])
        cap = int(pairs.shape[0]) if pairs is not None else 0
        out = ctypes.c_uint64(0)
        rc = self._lib.pbig_hip_check_u32(self._h, _ptr(boxes) if n else None, n, int(dims), int(range_),
                                          int(cell_size), _ptr(pairs) if cap else None, cap, ctypes.byref(out),
                                          _stream_ptr(stream))
        if rc == PBIG_E_OVERFLOW and allow_overflow:
            return int(out.value)
        if rc != PBIG_OK:
            self._raise(rc)
        return int(out.value)

    def check_host_u32(self, boxes: np.ndarray, dims: int, range_: int, cell_size: int = 0,
                       capacity: int | None = None) -> np.ndarray:
        """boxes: (n, 2*dims) uint32 host array (rid = index).  Returns (k, 2) uint32 pairs."""
        b = np.ascontiguousarray(boxes, dtype=np.uint32).reshape(-1, 2 * dims)
        n = b.shape[0]
        rec = np.empty((n, 2 * dims + 1), dtype=np.uint32)
        rec[:, :2 * dims] = b
        rec[:, 2 * dims] = np.arange(n, dtype=np.uint32)
        cap = int(capacity) if capacity is not None else max(1024, 2 * n)
        while True:
            out = np.empty((max(cap, 1), 2), dtype=np.uint32)
            cnt = ctypes.c_uint64(0)
            rc = self._lib.pbig_hip_check_host_u32(self._h, rec.ctypes.data if n else None, n, int(dims),
                                                   int(range_), int(cell_size), out.ctypes.data, cap,
                                                   ctypes.byref(cnt))
            if rc == PBIG_E_OVERFLOW and capacity is None:
                cap = int(cnt.value)
                continue
            if rc != PBIG_OK:
                self._raise(rc)
            return out[: int(cnt.value)]

    def check_host(self, rects, range_: float, cell_size: float = 0.0, capacity: int | None = None):
        """rects: RectInfo records (RECT_DTYPE) or an (n, 4) float32 array
        (then rid = index).  Returns an (k, 2) uint32 numpy array of pairs."""
        rects = as_rects(rects)
        n = rects.shape[0]
        cap = int(capacity) if capacity is not None else max(1024, 2 * n)
        while True:
            out = np.empty((max(cap, 1), 2), dtype=np.uint32)
            cnt = ctypes.c_uint64(0)
            rc = self._lib.pbig_hip_check_host(self._h, rects.ctypes.data if n else None, n, float(range_),
                                               float(cell_size), out.ctypes.data, cap, ctypes.byref(cnt))
            if rc == PBIG_E_OVERFLOW and capacity is None:
                cap = int(cnt.value)
                continue
            if rc != PBIG_OK:
                self._raise(rc)
            return out[: int(cnt.value)]

    def _each(self, call, fn) -> int:
        errors = []

        def chunk(user, ptr, m):
            try:
                fn(np.ctypeslib.as_array(ptr, shape=(int(m), 2)) if m else np.zeros((0, 2), np.uint32))
            except BaseException as e:  # (ctypes drops exceptions raised in callbacks)
                errors.append(e)

        cb = PAIRS_FN(chunk)
        cnt = ctypes.c_uint64(0)
        rc = call(cb, ctypes.byref(cnt))
        if errors:
            raise errors[0]
        if rc != PBIG_OK:
            self._raise(rc)
        return int(cnt.value)

    def check_host_each(self, rects, range_: float, cell_size: float, fn) -> int:
        """pbig_hip_check_host_each: fn(pairs) on every chunk of the pairs, an
        (m, 2) uint32 view of the library's pinned memory valid during the
        call, while the next chunk is copied back.  Returns the total."""
        rects = as_rects(rects)
        n = rects.shape[0]
        return self._each(lambda cb, cnt: self._lib.pbig_hip_check_host_each(
            self._h, rects.ctypes.data if n else None, n, float(range_), float(cell_size), cb, None, cnt), fn)

    def check_host_u32_each(self, boxes: np.ndarray, dims: int, range_: int, cell_size: int, fn) -> int:
        """pbig_hip_check_host_u32_each on (n, 2*dims) uint32 boxes (rid = index)."""
        b = np.ascontiguousarray(boxes, dtype=np.uint32).reshape(-1, 2 * dims)
        n = b.shape[0]
        rec = np.empty((n, 2 * dims + 1), dtype=np.uint32)
        rec[:, :2 * dims] = b
        rec[:, 2 * dims] = np.arange(n, dtype=np.uint32)
        return self._each(lambda cb, cnt: self._lib.pbig_hip_check_host_u32_each(
            self._h, rec.ctypes.data if n else None, n, int(dims), int(range_), int(cell_size), cb, None, cnt), fn)


def as_rects(boxes) -> np.ndarray:
    """(n, 4) float32 boxes -> RectInfo records with rid = index."""
    if isinstance(boxes, np.ndarray) and boxes.dtype == RECT_DTYPE:
        return np.ascontiguousarray(boxes)
    b = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 4)
    r = np.empty(b.shape[0], dtype=RECT_DTYPE)
    r["xmin"], r["xmax"], r["ymin"], r["ymax"] = b[:, 0], b[:, 1], b[:, 2], b[:, 3]
    r["rid"] = np.arange(b.shape[0], dtype=np.uint32)
    return r


class Batch:
    """A planned set of independent instances checked by one launch sequence
    (pbig_hip_batch_*).  ``offsets``: (k+1,) box offsets of the instances in
    the concatenated box array; ``ranges``, ``cell_sizes``: per instance."""

    def __init__(self, engine: Engine, offsets, ranges, cell_sizes):
        offsets = np.asarray(offsets, dtype=np.int64)
        self._create(engine, offsets[:-1], np.diff(offsets), ranges, cell_sizes)

    @classmethod
    def from_instances(cls, engine: Engine, starts, counts, ranges, cell_sizes) -> "Batch":
        """Instance i = boxes [starts[i], starts[i] + counts[i]) of the box array."""
        b = cls.__new__(cls)
        b._create(engine, np.asarray(starts, dtype=np.int64), np.asarray(counts, dtype=np.int64), ranges,
                  cell_sizes)
        return b

    @staticmethod
    def _table(starts, counts, ranges, cell_sizes):
        starts = np.asarray(starts, dtype=np.int64)
        counts = np.asarray(counts, dtype=np.int64)
        k = starts.shape[0]
        ranges = np.broadcast_to(np.asarray(ranges, dtype=np.float64), (k,))
        cells = np.broadcast_to(np.asarray(cell_sizes, dtype=np.float32), (k,))
        arr = (Instance * max(k, 1))()
        for i in range(k):
            arr[i].box_offset = int(starts[i])
            arr[i].n_boxes = int(counts[i])
            arr[i].cell_size = float(cells[i])
            arr[i].range = float(ranges[i])
        return arr, k, (int(np.max(starts + counts)) if k else 0)

    def _create(self, engine, starts, counts, ranges, cell_sizes):
        self._eng = engine
        self._lib = engine._lib
        arr, k, end = self._table(starts, counts, ranges, cell_sizes)
        h = ctypes.c_void_p()
        self._h = None
        rc = self._lib.pbig_hip_batch_create(engine._h, arr, k, ctypes.byref(h))
        if rc != PBIG_OK:
            engine._raise(rc)
        self._h = h
        self.n_inst = k
        self.n_boxes_total = end

    def update(self, starts, counts, ranges, cell_sizes, stream=None):
        """Re-plan for new instances, reusing the device workspace."""
        arr, k, end = self._table(starts, counts, ranges, cell_sizes)
        rc = self._lib.pbig_hip_batch_update(self._h, arr, k, _stream_ptr(stream))
        if rc != PBIG_OK:
            self._eng._raise(rc)
        self.n_inst = k
        self.n_boxes_total = end

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pbig_hip_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        nb, nc, nbin = ctypes.c_uint64(), ctypes.c_uint32(), ctypes.c_uint32()
        self._lib.pbig_hip_batch_info(self._h, ctypes.byref(nb), ctypes.byref(nc), ctypes.byref(nbin))
        return {"n_boxes": int(nb.value), "n_chunks": int(nc.value), "n_bins": int(nbin.value)}

    def check(self, boxes, pairs, allow_overflow: bool = False, stream=None, finish: bool = False):
        """Synchronous.  Returns (total pairs, (k, 2) uint64 numpy array of
        per-instance (offset, count)).  finish=True completes the last
        ``launch`` (same arguments and stream) instead of starting over."""
        cap = int(pairs.shape[0]) if pairs is not None else 0
        io = np.zeros((max(self.n_inst, 1), 2), dtype=np.uint64)
        out = ctypes.c_uint64(0)
        fn = self._lib.pbig_hip_batch_finish if finish else self._lib.pbig_hip_batch_check
        rc = fn(self._h, _ptr(boxes) if self.n_boxes_total else None, int(boxes.shape[0]),
                _ptr(pairs) if cap else None, cap, io.ctypes.data, ctypes.byref(out), _stream_ptr(stream))
        if rc == PBIG_E_OVERFLOW and allow_overflow:
            return int(out.value), io[: self.n_inst]
        if rc != PBIG_OK:
            self._eng._raise(rc)
        return int(out.value), io[: self.n_inst]

    def unresolved(self, stream=None) -> np.ndarray:
        """After a ``launch`` whose error word is only the bins-beyond-LDS bit (8): a bool
        array, True for the instances that had a bin beyond LDS (their counts
        are short); the other instances' (offset, count) entries and pairs
        from that launch are complete.  Raises PbigError(PBIG_E_OVERFLOW)
        when the launch found more pairs than its capacity."""
        flags = np.zeros(max(self.n_inst, 1), dtype=np.uint8)
        n = ctypes.c_uint32(0)
        rc = self._lib.pbig_hip_batch_unresolved(self._h, flags.ctypes.data, ctypes.byref(n), _stream_ptr(stream))
        if rc != PBIG_OK:
            self._eng._raise(rc)
        return flags[: self.n_inst].astype(bool)

    def launch(self, boxes, pairs, inst_pairs=None, result=None, stream=None):
        """Enqueue only.  inst_pairs: (k, 2) int64 device tensor; result: (2,)
        int64 device tensor <- {total, error}."""
        rc = self._lib.pbig_hip_batch_launch(self._h, _ptr(boxes) if self.n_boxes_total else None,
                                             int(boxes.shape[0]), _ptr(pairs), int(pairs.shape[0]),
                                             _ptr(inst_pairs) if inst_pairs is not None else None,
                                             _ptr(result) if result is not None else None, _stream_ptr(stream))
        if rc != PBIG_OK:
            self._eng._raise(rc)

    def capture(self, boxes, pairs, inst_pairs=None, result=None):
        rc = self._lib.pbig_hip_batch_capture(self._h, _ptr(boxes) if self.n_boxes_total else None,
                                              int(boxes.shape[0]), _ptr(pairs), int(pairs.shape[0]),
                                              _ptr(inst_pairs) if inst_pairs is not None else None,
                                              _ptr(result) if result is not None else None)
        if rc != PBIG_OK:
            self._eng._raise(rc)

    def replay(self, stream=None):
        rc = self._lib.pbig_hip_batch_replay(self._h, _stream_ptr(stream))
        if rc != PBIG_OK:
            self._eng._raise(rc)
