"""Seeded synthetic column generators for tests and benchmarks.

The generators run natively (ebi_synth_column in libebi_hip): a counter-based
SplitMix64 stream per (seed, stream), Box-Muller normals, and decimal rounding
half away from zero as Rust's f64::round used by the reference tests
(lemolatoon/ebi ebi/tests/aggregation.rs:14-16).
"""
from __future__ import annotations

import numpy as np

from . import synth as _synth


def round_half_away(x: np.ndarray, decimals: int) -> np.ndarray:
    s = 10.0 ** decimals
    y = np.asarray(x, dtype=np.float64) * s
    return np.copysign(np.floor(np.abs(y) + 0.5), y) / s + 0.0


def random_walk(n: int, x0: float = 100.0, sigma: float = 0.1, decimals: int = 2, seed: int = 1) -> np.ndarray:
    return _synth(0, n, seed, decimals, x0, sigma)


def uniform(n: int, lo: float = -1000.0, hi: float = 1000.0, decimals: int = 3, seed: int = 2) -> np.ndarray:
    return _synth(1, n, seed, decimals, lo, hi)


def step_series(n: int, p: float = 0.01, decimals: int = 2, seed: int = 3) -> np.ndarray:
    """Piecewise-constant levels; run lengths ~ Geom(p); levels a random walk."""
    return _synth(2, n, seed, decimals, p)


def sinusoid(n: int, period: float = 1000.0, noise: float = 0.01, decimals: int = 4, seed: int = 4) -> np.ndarray:
    return _synth(3, n, seed, decimals, period, noise)


# The four column generators of the plain-decode workload: (name, fn, decimals)
GENERATORS = {
    "walk2": (lambda n, seed=1: random_walk(n, 100.0, 0.1, 2, seed), 2),
    "unif3": (lambda n, seed=2: uniform(n, -1000.0, 1000.0, 3, seed), 3),
    "step2": (lambda n, seed=3: step_series(n, 0.01, 2, seed), 2),
    "sin4": (lambda n, seed=4: sinusoid(n, 1000.0, 0.01, 4, seed), 4),
}

# the eight codecs of the plain-decode and batched workloads
CODECS = ["uncompressed", "rle", "gorilla", "chimp", "chimp128", "elf", "buff", "delta_sprintz"]
# every codec the GPU path decodes (ElfOnChimp: the Elf wrapper over Chimp)
ALL_CODECS = CODECS + ["elf_on_chimp"]


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
