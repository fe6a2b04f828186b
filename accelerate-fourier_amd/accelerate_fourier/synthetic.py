"""Seeded synthetic signals: a splitmix64 stream per seed.

Uniform doubles take the top 53 bits of each 64-bit draw; normals come from
Box-Muller on pairs of uniforms.  The stream is counter based, so any slice of
it can be generated independently and reproduced in C from the same seed.
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed: int, count: int, start: int = 0) -> np.ndarray:
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, count: int, lo: float = -1.0, hi: float = 1.0, start: int = 0) -> np.ndarray:
    u = (splitmix64(seed, count, start) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


def normal(seed: int, count: int, start: int = 0) -> np.ndarray:
    m = (count + 1) // 2
    u = uniform(seed, 2 * m, 0.0, 1.0, 2 * start)
    u1 = np.maximum(u[0::2], 1e-300)
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * m)
    z[0::2] = r * np.cos(2 * np.pi * u2)
    z[1::2] = r * np.sin(2 * np.pi * u2)
    return z[:count]


def complex_uniform(seed: int, shape) -> np.ndarray:
    """Real parts then imaginary parts, each uniform in [-1, 1)."""
    n = int(np.prod(shape))
    v = uniform(seed, 2 * n)
    return (v[:n] + 1j * v[n:]).reshape(shape)


def complex_normal(seed: int, shape) -> np.ndarray:
    n = int(np.prod(shape))
    v = normal(seed, 2 * n)
    return (v[:n] + 1j * v[n:]).reshape(shape)


def integers(seed: int, count: int, lo: int, hi: int) -> np.ndarray:
    """Uniform integers in [lo, hi)."""
    u = uniform(seed, count, 0.0, 1.0)
    return (lo + np.floor(u * (hi - lo))).astype(np.int64)


def _primes_between(lo: int, hi: int) -> np.ndarray:
    """Primes p with lo <= p <= hi (sieve of Eratosthenes)."""
    s = np.ones(hi + 1, dtype=bool)
    s[:2] = False
    for p in range(2, int(hi ** 0.5) + 1):
        if s[p]:
            s[p * p::p] = False
    return np.nonzero(s[lo:])[0] + lo


def _smooth7_between(lo: int, hi: int) -> np.ndarray:
    out = []
    p2 = 1
    while p2 <= hi:
        p3 = p2
        while p3 <= hi:
            p5 = p3
            while p5 <= hi:
                p7 = p5
                while p7 <= hi:
                    if p7 >= lo:
                        out.append(p7)
                    p7 *= 7
                p5 *= 5
            p3 *= 3
        p2 *= 2
    return np.array(sorted(out), dtype=np.int64)


# The uneven-batch workload: 4096 rows in four length classes.
UNEVEN_CLASSES = ("pow2 2^6..2^16", "7-smooth 60..50000", "primes 257..65521", "primes 100003..1000003")
UNEVEN_COUNTS = (1638, 1229, 819, 410)


def uneven_lengths(seed: int = 2):
    """Lengths and class ids of the 4096-row uneven batch: per class, lengths drawn
    uniformly (with replacement) from the class's set, then all rows shuffled."""
    sets = (np.array([1 << k for k in range(6, 17)], dtype=np.int64), _smooth7_between(60, 50000),
            _primes_between(257, 65521), _primes_between(100003, 1000003))
    total = sum(UNEVEN_COUNTS)
    u = splitmix64(seed, 2 * total)
    lens, cls = [], []
    k = 0
    for c, (st, cnt) in enumerate(zip(sets, UNEVEN_COUNTS)):
        for _ in range(cnt):
            lens.append(int(st[int(u[k] % np.uint64(len(st)))]))
            cls.append(c)
            k += 1
    order = np.argsort(u[total:], kind="stable")
    return np.array(lens, dtype=np.int64)[order], np.array(cls, dtype=np.int64)[order]
