"""Seeded synthetic box workloads (2-D float32, {xmin, xmax, ymin, ymax}).

Every value is produced with IEEE basic operations only (+, -, *, /, sqrt,
frexp/ldexp, rounding), so a given seed yields bit-identical boxes on any
host: the logarithm, exponential and cosine used for the Gaussian and
log-normal laws are evaluated here by range reduction and series rather than
by libm or numpy's SIMD kernels, whose last bit can differ between CPUs.

Random stream: SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom
number generators", OOPSLA 2014), counter-based: draw i of stream s is
mix(key(seed, s) + (i + 1) * 0x9E3779B97F4A7C15).  A float in [0, 1) is the
top 24 bits times 2**-24, which is exact in float32.

Workloads (sizes as named by the benchmark configurations):
  uniform(n, L, smin, smax)      centres U[0, L)^2, sides U[smin, smax]
  clustered(n, L, ...)           90% of centres from 64 isotropic Gaussian
                                 clusters (sigma = 0.5% of L), 10% uniform;
                                 log-normal sides (mu = ln 4, sigma = 1)
                                 clipped to [0.5, 4096]
  batched(k, ...)                k instances, n log-uniform in [64, 262144],
                                 L = 16 sqrt(n), sides U[1, 8]
  serving_requests(...)          Poisson arrivals, n log-uniform in
                                 [1e3, 1e6], 50/50 uniform / clustered
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_LN2 = 0.6931471805599453          # ln 2 rounded to double
_TWO_PI = 6.283185307179586        # 2*pi rounded to double
_HALF_PI = 1.5707963267948966


def _mix(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def _key(seed: int, stream: int) -> np.uint64:
    with np.errstate(over="ignore"):
        s = np.array([(seed * 1000003 + stream * 7919 + 1) & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64)
    return _mix(s)[0]


def raw64(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return _mix(_key(seed, stream) + i * _GOLDEN)


def u01_f32(seed: int, stream: int, n: int) -> np.ndarray:
    """Uniform float32 in [0, 1), exact multiples of 2**-24."""
    return (raw64(seed, stream, n) >> np.uint64(40)).astype(np.float32) * np.float32(2.0 ** -24)


def u01_f64(seed: int, stream: int, n: int) -> np.ndarray:
    """Uniform float64 in (0, 1): (top 53 bits + 0.5) * 2**-53, never 0."""
    return ((raw64(seed, stream, n) >> np.uint64(11)).astype(np.float64) + 0.5) * (2.0 ** -53)


# ---------------------------------------------------------------------------
# Deterministic elementary functions (float64, basic operations only)
# ---------------------------------------------------------------------------

def det_log(x: np.ndarray) -> np.ndarray:
    """Natural log for x > 0: x = m * 2**e, m in [sqrt(.5), sqrt(2)), then
    log m = 2 atanh(s), s = (m-1)/(m+1), |s| < 0.1716, series to s**25."""
    x = np.asarray(x, dtype=np.float64)
    m, e = np.frexp(x)                      # m in [0.5, 1)
    small = m < 0.7071067811865476
    m = np.where(small, m * 2.0, m)
    e = np.where(small, e - 1, e).astype(np.float64)
    s = (m - 1.0) / (m + 1.0)
    s2 = s * s
    acc = np.zeros_like(s)
    for k in range(25, 0, -2):
        acc = acc * s2 + 1.0 / k
    return 2.0 * s * acc + e * _LN2


def det_exp(x: np.ndarray) -> np.ndarray:
    """exp(x): x = k ln2 + r, |r| <= ln2/2, Taylor series to r**20."""
    x = np.asarray(x, dtype=np.float64)
    k = np.round(x / _LN2)
    r = x - k * _LN2
    acc = np.ones_like(r)
    for n in range(20, 0, -1):
        acc = 1.0 + acc * r / n
    return np.ldexp(acc, k.astype(np.int64))


def det_cos_2pi(v: np.ndarray) -> np.ndarray:
    """cos(2 pi v) for any v: reduce to t in [-1/2, 1/2], use symmetry to
    |theta| <= pi/2, Taylor series to theta**22."""
    v = np.asarray(v, dtype=np.float64)
    t = v - np.round(v)
    th = np.abs(t) * _TWO_PI                # [0, pi]
    flip = th > _HALF_PI
    th = np.where(flip, np.pi - th, th)
    t2 = th * th
    acc = np.ones_like(th)
    for n in range(22, 0, -2):
        acc = 1.0 - acc * t2 / (n * (n - 1))
    return np.where(flip, -acc, acc)


def std_normal(seed: int, stream: int, n: int) -> np.ndarray:
    """Box-Muller (one variate per uniform pair), deterministic math."""
    u = u01_f64(seed, stream, n)
    v = u01_f64(seed, stream + 1, n)
    return np.sqrt(-2.0 * det_log(u)) * det_cos_2pi(v)


# ---------------------------------------------------------------------------
# Workloads
# ---------------------------------------------------------------------------

def _assemble(cx, cy, w, h) -> np.ndarray:
    cx = cx.astype(np.float32)
    cy = cy.astype(np.float32)
    hw = np.float32(0.5) * w.astype(np.float32)
    hh = np.float32(0.5) * h.astype(np.float32)
    out = np.empty((cx.shape[0], 4), dtype=np.float32)
    out[:, 0] = cx - hw
    out[:, 1] = cx + hw
    out[:, 2] = cy - hh
    out[:, 3] = cy + hh
    return out


def uniform(n: int, L: float, smin: float, smax: float, seed: int) -> np.ndarray:
    """Centres U[0, L)^2, width and height independently U[smin, smax]."""
    L32 = np.float32(L)
    cx = u01_f32(seed, 0, n) * L32
    cy = u01_f32(seed, 1, n) * L32
    span = np.float32(smax - smin)
    w = np.float32(smin) + u01_f32(seed, 2, n) * span
    h = np.float32(smin) + u01_f32(seed, 3, n) * span
    return _assemble(cx, cy, w, h)


def clustered(n: int, L: float, seed: int, n_clusters: int = 64, sigma_frac: float = 0.005,
              clustered_frac: float = 0.9, ln_mu: float = 1.3862943611198906, ln_sigma: float = 1.0,
              smin: float = 0.5, smax: float = 4096.0) -> np.ndarray:
    """Skewed instance: clustered_frac of the centres from n_clusters isotropic
    Gaussians (centres U[0, L)^2, sigma = sigma_frac * L), the rest uniform;
    sides log-normal(ln_mu, ln_sigma) clipped to [smin, smax]."""
    sigma = sigma_frac * L
    ccx = u01_f64(seed, 10, n_clusters) * L
    ccy = u01_f64(seed, 11, n_clusters) * L
    pick = np.minimum((u01_f64(seed, 12, n) * n_clusters).astype(np.int64), n_clusters - 1)
    is_cl = u01_f64(seed, 13, n) < clustered_frac
    gx = ccx[pick] + sigma * std_normal(seed, 20, n)
    gy = ccy[pick] + sigma * std_normal(seed, 22, n)
    ux = u01_f64(seed, 14, n) * L
    uy = u01_f64(seed, 15, n) * L
    cx = np.where(is_cl, gx, ux)
    cy = np.where(is_cl, gy, uy)
    w = np.clip(det_exp(ln_mu + ln_sigma * std_normal(seed, 30, n)), smin, smax)
    h = np.clip(det_exp(ln_mu + ln_sigma * std_normal(seed, 32, n)), smin, smax)
    return _assemble(cx, cy, w, h)


def log_uniform_int(seed: int, stream: int, k: int, lo: float, hi: float) -> np.ndarray:
    """floor(exp(U[ln lo, ln hi])) with deterministic math, clipped to [lo, hi]."""
    u = u01_f64(seed, stream, k)
    a, b = det_log(np.array([lo, hi], dtype=np.float64))
    n = np.floor(det_exp(a + u * (b - a))).astype(np.int64)
    return np.clip(n, int(lo), int(hi))


def batched(k: int, seed: int, nmin: int = 64, nmax: int = 262144, smin: float = 1.0,
            smax: float = 8.0, size_scale: float = 1.0):
    """k independent uniform instances; returns (boxes, offsets[k+1], ranges[k]).
    Instance t has n_t log-uniform in [nmin, nmax] boxes (times size_scale),
    centres U[0, L_t)^2 with L_t = 16 sqrt(n_t), sides U[smin, smax]."""
    n = log_uniform_int(seed, 100, k, nmin, nmax)
    if size_scale != 1.0:
        n = np.maximum((n * size_scale).astype(np.int64), 2)
    offsets = np.zeros(k + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(n)
    ranges = (16.0 * np.sqrt(n.astype(np.float64))).astype(np.float32)
    boxes = np.empty((int(offsets[-1]), 4), dtype=np.float32)
    for t in range(k):
        boxes[offsets[t]:offsets[t + 1]] = uniform(int(n[t]), float(ranges[t]), smin, smax,
                                                    seed * 100003 + t + 1)
    return boxes, offsets, ranges


def serving_requests(count: int, seed: int, rate_per_s: float, nmin: int = 1000,
                     nmax: int = 1000000):
    """Poisson request stream: returns a list of dicts with arrival time (s),
    kind ('uniform' | 'clustered'), n, L.  The domain side scales with n so
    that box density matches the 4M-box uniform and 2M-box clustered
    instances: L = 65536 sqrt(n / 4194304) (uniform), 65536 sqrt(n / 2097152)
    (clustered)."""
    n = log_uniform_int(seed, 200, count, nmin, nmax)
    gaps = -det_log(u01_f64(seed, 201, count)) / rate_per_s
    t = np.cumsum(gaps)
    kind = u01_f64(seed, 202, count) < 0.5
    reqs = []
    for i in range(count):
        if kind[i]:
            L = 65536.0 * float(np.sqrt(n[i] / 4194304.0))
            reqs.append({"id": i, "t": float(t[i]), "kind": "uniform", "n": int(n[i]), "L": L,
                         "seed": seed * 7 + i})
        else:
            L = 65536.0 * float(np.sqrt(n[i] / 2097152.0))
            reqs.append({"id": i, "t": float(t[i]), "kind": "clustered", "n": int(n[i]), "L": L,
                         "seed": seed * 7 + i})
    return reqs


def request_boxes(req: dict) -> np.ndarray:
    if req["kind"] == "uniform":
        return uniform(req["n"], req["L"], 1.0, 32.0, req["seed"])
    return clustered(req["n"], req["L"], req["seed"])


# Named benchmark instances ----------------------------------------------------

SEEDS = {"c1": 1, "c2": 2, "c3": 3, "c4": 4, "c5": 5}


def c1_small_uniform():
    """16,384 rects, centres U[0,1024)^2, sides U[1,8]."""
    return uniform(16384, 1024.0, 1.0, 8.0, SEEDS["c1"]), 1024.0


def c2_large_uniform(n: int = 4194304):
    """4,194,304 rects, centres U[0,65536)^2, sides U[1,32]."""
    return uniform(n, 65536.0, 1.0, 32.0, SEEDS["c2"]), 65536.0


def c3_clustered(n: int = 2097152):
    """2,097,152 rects, 90% from 64 clusters (sigma 327.68), log-normal sides."""
    return clustered(n, 65536.0, SEEDS["c3"]), 65536.0
