"""java.util.SplittableRandom, vectorised with numpy.

The reference seeds shot sampling and Kraus sampling with
`new SplittableRandom(seed)` (StatevectorEngine.java:81-85, :238-240) and draws
`nextDouble()` per shot (:245, :252) and `nextDouble(bound)` per Kraus choice
(:489).  The JDK class is SplitMix64 (Steele, Lea & Flood, OOPSLA 2014):
state += 0x9e3779b97f4a7c15; output = mix64(state) with the Stafford
variant-13 finaliser; nextDouble = (nextLong >>> 11) * 2^-53;
nextDouble(bound) = nextDouble() * bound, clamped below bound.
The host draws variates in reference order and uploads them; kernels never
generate their own on the parity path.
"""
from __future__ import annotations

import math

import numpy as np

GOLDEN_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_MASK = (1 << 64) - 1


def _mix64(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


class SplittableRandom:
    def __init__(self, seed: int):
        self.seed = int(seed) & _MASK

    def next_longs(self, count: int) -> np.ndarray:
        """The next `count` nextLong() values as uint64."""
        with np.errstate(over="ignore"):
            k = np.arange(1, count + 1, dtype=np.uint64)
            states = np.uint64(self.seed) + k * GOLDEN_GAMMA
        self.seed = (self.seed + count * 0x9E3779B97F4A7C15) & _MASK
        return _mix64(states)

    def next_doubles(self, count: int) -> np.ndarray:
        return (self.next_longs(count) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)

    def next_double(self) -> float:
        return float(self.next_doubles(1)[0])

    def next_double_bounded(self, bound: float) -> float:
        r = self.next_double() * bound
        if r >= bound:
            r = math.nextafter(bound, -math.inf)
        return r

    def next_long(self) -> int:
        v = int(self.next_longs(1)[0])
        return v - (1 << 64) if v >= (1 << 63) else v
