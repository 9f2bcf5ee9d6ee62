"""Cubie model of the 3x3 cube (the reference's convention, src/cuda_cube_solver/cubie.py:1-51).

``cp[position]`` is the corner cubie at that position; applying ``b`` after
``a`` gives ``cp = a.cp[b.cp[i]]``.  The 18 half-turn-metric moves are derived
from the sticker geometry: a clockwise quarter turn of a face rotates its layer
by -90 degrees about the face normal.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Iterable

CORNER_NAMES = ("URF", "UFL", "ULB", "UBR", "DFR", "DLF", "DBL", "DRB")
EDGE_NAMES = ("UR", "UF", "UL", "UB", "DR", "DF", "DL", "DB", "FR", "FL", "BL", "BR")
FACE_NAMES = ("U", "R", "F", "D", "L", "B")
_NORMAL = {"U": (0, 1, 0), "R": (1, 0, 0), "F": (0, 0, 1), "D": (0, -1, 0), "L": (-1, 0, 0), "B": (0, 0, -1)}


@dataclass(frozen=True)
class CubieState:
    cp: tuple = tuple(range(8))
    co: tuple = (0,) * 8
    ep: tuple = tuple(range(12))
    eo: tuple = (0,) * 12

    def then(self, transform: "CubieState") -> "CubieState":
        return compose(self, transform)

    def inverse(self) -> "CubieState":
        cinv = [0] * 8
        einv = [0] * 12
        for pos, c in enumerate(self.cp):
            cinv[c] = pos
        for pos, e in enumerate(self.ep):
            einv[e] = pos
        return CubieState(tuple(cinv), tuple((-self.co[cinv[i]]) % 3 for i in range(8)),
                          tuple(einv), tuple(self.eo[einv[i]] for i in range(12)))

    def is_legal(self) -> bool:
        return (sorted(self.cp) == list(range(8)) and sorted(self.ep) == list(range(12))
                and all(0 <= v < 3 for v in self.co) and all(0 <= v < 2 for v in self.eo)
                and sum(self.co) % 3 == 0 and sum(self.eo) % 2 == 0
                and permutation_parity(self.cp) == permutation_parity(self.ep))

    def to_bytes(self) -> bytes:
        """The 40-byte C-ABI layout: cp[8], co[8], ep[12], eo[12]."""
        return bytes(self.cp) + bytes(self.co) + bytes(self.ep) + bytes(self.eo)

    @classmethod
    def from_bytes(cls, b: bytes) -> "CubieState":
        b = bytes(b)
        return cls(tuple(b[0:8]), tuple(b[8:16]), tuple(b[16:28]), tuple(b[28:40]))


IDENTITY = CubieState()


def compose(a: CubieState, b: CubieState) -> CubieState:
    """``b`` applied after ``a``."""
    return CubieState(
        tuple(a.cp[b.cp[i]] for i in range(8)),
        tuple((a.co[b.cp[i]] + b.co[i]) % 3 for i in range(8)),
        tuple(a.ep[b.ep[i]] for i in range(12)),
        tuple((a.eo[b.ep[i]] + b.eo[i]) % 2 for i in range(12)),
    )


def _add(*vs):
    return tuple(sum(c) for c in zip(*vs))


def _rot(n, v):
    """Rotate v by -90 degrees about the unit axis n: n (n.v) - n x v."""
    d = sum(a * b for a, b in zip(n, v))
    cr = (n[1] * v[2] - n[2] * v[1], n[2] * v[0] - n[0] * v[2], n[0] * v[1] - n[1] * v[0])
    return tuple(n[i] * d - cr[i] for i in range(3))


def _quarter_turn(face: str) -> CubieState:
    n = _NORMAL[face]
    cp, co, ep, eo = list(range(8)), [0] * 8, list(range(12)), [0] * 12
    for names, perm, ori, mod in ((CORNER_NAMES, cp, co, 3), (EDGE_NAMES, ep, eo, 2)):
        centers = [_add(*(_NORMAL[f] for f in name)) for name in names]
        for p, name in enumerate(names):
            if sum(a * b for a, b in zip(centers[p], n)) <= 0:
                continue
            q = centers.index(_rot(n, centers[p]))
            perm[q] = p  # position q receives the cubie from p
            face0 = _rot(n, _NORMAL[name[0]])  # where sticker 0 of that cubie lands
            ori[q] = [_NORMAL[f] for f in names[q]].index(face0) % mod
    return CubieState(tuple(cp), tuple(co), tuple(ep), tuple(eo))


def _power(t: CubieState, k: int) -> CubieState:
    r = IDENTITY
    for _ in range(k % 4):
        r = compose(r, t)
    return r


MOVE_NAMES_HTM: tuple = tuple(s for f in FACE_NAMES for s in (f, f"{f}2", f"{f}'"))
MOVES_HTM: tuple = tuple(_power(_quarter_turn(f), k) for f in FACE_NAMES for k in (1, 2, 3))


def apply_moves(state: CubieState, moves: Iterable[CubieState]) -> CubieState:
    for m in moves:
        state = compose(state, m)
    return state


def permutation_parity(p) -> int:
    return sum(p[i] > p[j] for i in range(len(p)) for j in range(i + 1, len(p))) & 1


def parse_moves(text: str) -> list[int]:
    """Move indices of a scramble such as "R U R' U2" (KeyError on an unknown move)."""
    index = {name: i for i, name in enumerate(MOVE_NAMES_HTM)}
    return [index[tok] for tok in text.split()]


def moves_text(moves: Iterable[int]) -> str:
    return " ".join(MOVE_NAMES_HTM[m] for m in moves)
