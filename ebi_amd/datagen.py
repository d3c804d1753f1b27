"""Seeded synthetic column generators for tests and benchmarks.

All randomness comes from a counter-based SplitMix64 stream (so every value
depends only on (seed, index) and columns are byte-reproducible across
machines), with Box-Muller for normals.  Decimal rounding is round-half-away
from zero, as Rust's f64::round used by the reference tests
(lemolatoon/ebi ebi/tests/aggregation.rs:14-16).
"""
from __future__ import annotations

import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """z_i = mix(seed' + (i + 1) * golden) for i < n, as uint64."""
    with np.errstate(over="ignore"):
        s = np.uint64((seed * 0x100000001B3 + stream * 0x632BE59BD9B4E019) & 0xFFFFFFFFFFFFFFFF)
        z = s + (np.arange(1, n + 1, dtype=np.uint64) * _GOLD)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, n: int, stream: int = 0) -> np.ndarray:
    return (splitmix64(seed, n, stream) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, n: int, stream: int = 0) -> np.ndarray:
    u1 = uniform01(seed, n, 2 * stream + 1)
    u2 = uniform01(seed, n, 2 * stream + 2)
    u1 = np.maximum(u1, 1e-300)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def round_half_away(x: np.ndarray, decimals: int) -> np.ndarray:
    s = 10.0 ** decimals
    y = x * s
    return np.copysign(np.floor(np.abs(y) + 0.5), y) / s


def _no_negzero(x: np.ndarray) -> np.ndarray:
    return x + 0.0  # -0.0 + 0.0 == +0.0


def random_walk(n: int, x0: float = 100.0, sigma: float = 0.1, decimals: int = 2, seed: int = 1) -> np.ndarray:
    return _no_negzero(round_half_away(x0 + np.cumsum(sigma * normal(seed, n)), decimals))


def uniform(n: int, lo: float = -1000.0, hi: float = 1000.0, decimals: int = 3, seed: int = 2) -> np.ndarray:
    return _no_negzero(round_half_away(lo + (hi - lo) * uniform01(seed, n), decimals))


def step_series(n: int, p: float = 0.01, decimals: int = 2, seed: int = 3) -> np.ndarray:
    """Piecewise-constant levels; run lengths ~ Geom(p); levels a random walk."""
    u = uniform01(seed, n, 7)
    change = u < p
    change[0] = True
    level_idx = np.cumsum(change) - 1
    steps = normal(seed, int(level_idx[-1]) + 1 if n else 0, 9)
    levels = round_half_away(100.0 + np.cumsum(steps), decimals)
    return _no_negzero(levels[level_idx])


def sinusoid(n: int, period: float = 1000.0, noise: float = 0.01, decimals: int = 4, seed: int = 4) -> np.ndarray:
    i = np.arange(n, dtype=np.float64)
    return _no_negzero(round_half_away(100.0 * np.sin(2.0 * np.pi * i / period) + noise * normal(seed, n), decimals))


# The four column generators of the plain-decode workload: (name, fn, decimals)
GENERATORS = {
    "walk2": (lambda n, seed=1: random_walk(n, 100.0, 0.1, 2, seed), 2),
    "unif3": (lambda n, seed=2: uniform(n, -1000.0, 1000.0, 3, seed), 3),
    "step2": (lambda n, seed=3: step_series(n, 0.01, 2, seed), 2),
    "sin4": (lambda n, seed=4: sinusoid(n, 1000.0, 0.01, 4, seed), 4),
}

CODECS = ["uncompressed", "rle", "gorilla", "chimp", "chimp128", "elf", "buff", "delta_sprintz"]


def xor_values() -> np.ndarray:
    """The adversarial XOR_VALUES sequence of the reference round-trip test
    (ebi/tests/round_trip.rs:352-376): flip one more mantissa/exponent bit per
    step, doubling (negated) instead when the flip gives NaN or infinity."""
    vals = np.zeros(66, dtype=np.float64)
    vals[0] = 3.3
    vals[1] = 3.3
    i = 1
    while i < 64:
        last = vals[i]
        nb = np.array([last]).view(np.uint64)[0] ^ np.uint64(1 << (i - 1))
        nxt = np.array([nb], dtype=np.uint64).view(np.float64)[0]
        if np.isnan(nxt) or np.isinf(nxt):
            vals[i] = last * -2.0
        else:
            vals[i + 1] = nxt
            i += 1
    return vals


def xor_values2() -> np.ndarray:
    """ebi/tests/round_trip.rs:382-390."""
    v = np.array([2.3, 3.3, 3.3])
    b = v[2:3].view(np.uint64).copy()
    for i in range(2):
        b ^= np.uint64(1 << i)
    v[2] = b.view(np.float64)[0]
    return v
