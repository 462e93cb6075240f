"""Host logic of the serving loop: the pool's first-fit allocator (no GPU)."""
import random

from pbig_amd.serve import Pool


def test_pool_first_fit_and_merge():
    p = Pool(100)
    a = p.alloc(30)
    b = p.alloc(30)
    c = p.alloc(30)
    assert (a, b, c) == (0, 30, 60)
    assert p.alloc(20) == -1  # only 10 left
    p.release(b, 30)
    assert p.alloc(20) == 30  # first fit reuses the hole
    p.release(a, 30)
    p.release(30, 20)
    p.release(c, 30)
    assert p.free == [(0, 100)] and p.used == 0


def test_pool_random_never_overlaps():
    rng = random.Random(5)
    p = Pool(10_000)
    live = {}
    for _ in range(5000):
        if live and rng.random() < 0.45:
            s = rng.choice(list(live))
            p.release(s, live.pop(s))
        else:
            n = rng.randint(1, 900)
            s = p.alloc(n)
            if s >= 0:
                assert all(s + n <= t or u + live[u] <= s for u in live for t in [u])
                live[s] = n
        assert p.used == sum(live.values())
        assert sum(ln for _, ln in p.free) == p.capacity - p.used
    for s in list(live):
        p.release(s, live.pop(s))
    assert p.free == [(0, 10_000)]


def test_cell_side_floor():
    """The autotuner's cells are never narrower than the widest box of a
    bounded-size sample (with headroom), so boxes span at most two cells per
    axis; unbounded (log-normal) sides set no floor."""
    import numpy as np

    from pbig_amd import autotune, workloads

    b, L = workloads.c2_large_uniform(n=200000)
    fl = autotune.side_floor(b)
    assert 32.0 < fl < 36.0
    t = {"classes": [{"max_n": None, "boxes_per_cell": 0.25}], "bin_target": 3000}
    c = autotune.cell_sizes([0, b.shape[0]], [L], t, b)[0]
    assert c >= fl and c >= np.max(np.maximum(b[:, 1] - b[:, 0], b[:, 3] - b[:, 2]))
    bc, Lc = workloads.c3_clustered(n=100000)
    assert autotune.side_floor(bc) == 0.0
    assert autotune.cell_sizes([0, bc.shape[0]], [Lc], t, bc)[0] == autotune.cell_sizes([0, bc.shape[0]], [Lc], t)[0]
