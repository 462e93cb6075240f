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
