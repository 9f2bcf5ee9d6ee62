"""Seeded cube workloads: uniform random legal states, flip-heavy states.

Random numbers come from splitmix64 so the same seed gives the same cubes on
every machine.  A uniform state: Fisher-Yates corner and edge permutations,
edges at positions 10 and 11 swapped when the parities differ, 7 uniform
twists with the 8th making the sum 0 mod 3, 11 uniform flips with the 12th
making the sum even.
"""
from __future__ import annotations

from .cubie import CubieState, permutation_parity

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.x = seed & MASK64

    def next(self) -> int:
        self.x = (self.x + 0x9E3779B97F4A7C15) & MASK64
        z = self.x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def below(self, n: int) -> int:
        """Uniform integer in [0, n) (multiply-shift of the top 32 bits)."""
        return ((self.next() >> 32) * n) >> 32


def _shuffle(rng: SplitMix64, n: int) -> list[int]:
    p = list(range(n))
    for i in range(n - 1, 0, -1):
        j = rng.below(i + 1)
        p[i], p[j] = p[j], p[i]
    return p


def uniform_states(n: int, seed: int) -> list[CubieState]:
    rng = SplitMix64(seed)
    out = []
    for _ in range(n):
        cp = _shuffle(rng, 8)
        ep = _shuffle(rng, 12)
        if permutation_parity(cp) != permutation_parity(ep):
            ep[10], ep[11] = ep[11], ep[10]
        co = [rng.below(3) for _ in range(7)]
        co.append((-sum(co)) % 3)
        eo = [rng.below(2) for _ in range(11)]
        eo.append(sum(eo) % 2)
        out.append(CubieState(tuple(cp), tuple(co), tuple(ep), tuple(eo)))
    return out


def sample_indices(pool: int, n: int, seed: int) -> list[int]:
    """n distinct indices of range(pool), drawn with seed (partial Fisher-Yates)."""
    if n > pool:
        raise ValueError(f"cannot draw {n} of {pool}")
    rng = SplitMix64(seed)
    p = list(range(pool))
    for i in range(n):
        j = i + rng.below(pool - i)
        p[i], p[j] = p[j], p[i]
    return p[:n]


def flip_heavy_states(n: int, seed: int) -> list[CubieState]:
    """Corners and edge permutation solved, a random even-sized subset of the 12
    edges flipped in place; the last state is the superflip (all 12)."""
    rng = SplitMix64(seed)
    out = []
    for i in range(n):
        if i == n - 1:
            flips = [1] * 12
        else:
            while True:
                flips = [rng.below(2) for _ in range(12)]
                if sum(flips) % 2 == 0 and sum(flips) > 0:
                    break
        out.append(CubieState(tuple(range(8)), (0,) * 8, tuple(range(12)), tuple(flips)))
    return out
