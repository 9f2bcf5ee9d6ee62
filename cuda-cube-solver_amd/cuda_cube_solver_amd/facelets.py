"""54-sticker facelet strings <-> cubie states.

The string format is the reference's and the kociemba package's
(src/cuda_cube_solver/facelets.py:1-41): faces in U, R, F, D, L, B order,
nine stickers each, read row by row as seen from outside the cube, with U
seen with B at the top, D with F at the top and the four side faces with U at
the top.  Any six symbols may be used; the centres say which face a symbol
belongs to.

Here the sticker positions of every corner and edge slot are derived from
that viewing convention and the cube's geometry (cubie.py's face normals)
rather than written out, and parsing is a table lookup of each slot's
sticker tuple.  The error messages and the order in which the checks run
are the reference's, since callers match on them.
"""
from __future__ import annotations

from .cubie import _NORMAL, CORNER_NAMES, EDGE_NAMES, FACE_NAMES, CubieState, permutation_parity

FACES = "".join(FACE_NAMES)  # "URFDLB": the order of the faces in the string

# How each face is drawn: the directions of a row step to the right and of a
# column step down, as seen from outside the cube.
_VIEW = {
    "U": ((1, 0, 0), (0, 0, 1)),    # B at the top
    "R": ((0, 0, -1), (0, -1, 0)),
    "F": ((1, 0, 0), (0, -1, 0)),
    "D": ((1, 0, 0), (0, 0, -1)),   # F at the top
    "L": ((0, 0, 1), (0, -1, 0)),
    "B": ((-1, 0, 0), (0, -1, 0)),
}


def _sticker_at(face: str, where: tuple) -> int:
    """String index of the sticker of `face` on the cubie centred at `where`."""
    right, down = _VIEW[face]
    n = _NORMAL[face]
    for k in range(9):
        r, c = divmod(k, 3)
        if tuple(n[i] + (c - 1) * right[i] + (r - 1) * down[i] for i in range(3)) == where:
            return 9 * FACES.index(face) + k
    raise AssertionError(f"no {face} sticker at {where}")


def _slot_stickers(names) -> tuple:
    """Per slot, the string indices of its stickers in the order of the slot's name
    (for corners: the U/D sticker first, then clockwise)."""
    out = []
    for name in names:
        centre = tuple(sum(_NORMAL[f][i] for f in name) for i in range(3))
        out.append(tuple(_sticker_at(f, centre) for f in name))
    return tuple(out)


CORNER_FACELETS = _slot_stickers(CORNER_NAMES)
EDGE_FACELETS = _slot_stickers(EDGE_NAMES)
CORNER_FACES = CORNER_NAMES
EDGE_FACES = EDGE_NAMES

# Sticker tuple read at a slot -> (cubie, orientation).  A cubie with
# orientation o shows its k-th name letter at the slot's sticker (k + o) mod n.
_CORNER_READ = {tuple(name[(k - o) % 3] for k in range(3)): (j, o)
                for j, name in enumerate(CORNER_NAMES) for o in range(3)}
_EDGE_READ = {tuple(name[(k - o) % 2] for k in range(2)): (j, o)
              for j, name in enumerate(EDGE_NAMES) for o in range(2)}


class FaceletError(ValueError):
    """The string does not describe a legal cube."""


def _stickers(text: str) -> str:
    return "".join(text.split())


def is_facelet_string(text: str) -> bool:
    s = _stickers(text)
    return len(s) == 54 and len(set(s)) == 6


def _faces_of(s: str) -> str:
    """The string as face letters, checked sticker by sticker and face by face."""
    if len(s) != 54:
        raise FaceletError(f"expected 54 stickers, got {len(s)}")
    centres = s[4::9]
    if len(set(centres)) != 6:
        raise FaceletError("the six centre stickers must all differ")
    table = {sym: face for sym, face in zip(centres, FACES)}
    for sym in s:
        if sym not in table:
            raise FaceletError(f"sticker {sym!r} matches no centre")
    faces = "".join(table[sym] for sym in s)
    for face in FACES:
        k = faces.count(face)
        if k != 9:
            raise FaceletError(f"face {face} has {k} stickers, expected 9")
    return faces


def parse_facelets(text: str) -> CubieState:
    """Cubie state of a 54-sticker string; FaceletError if it is not a legal cube."""
    faces = _faces_of(_stickers(text))
    cp, co = [], []
    for slot, idx in enumerate(CORNER_FACELETS):
        seen = tuple(faces[i] for i in idx)
        hit = _CORNER_READ.get(seen)
        if hit is None:
            if not set(seen) & {"U", "D"}:
                raise FaceletError(f"corner at {CORNER_NAMES[slot]} has no U or D sticker")
            raise FaceletError(f"corner at {CORNER_NAMES[slot]} has impossible colours {''.join(seen)}")
        cp.append(hit[0])
        co.append(hit[1])
    ep, eo = [], []
    for slot, idx in enumerate(EDGE_FACELETS):
        seen = tuple(faces[i] for i in idx)
        hit = _EDGE_READ.get(seen)
        if hit is None:
            raise FaceletError(f"edge at {EDGE_NAMES[slot]} has impossible colours {''.join(seen)}")
        ep.append(hit[0])
        eo.append(hit[1])
    if len(set(cp)) != 8:
        raise FaceletError("some corner appears twice")
    if len(set(ep)) != 12:
        raise FaceletError("some edge appears twice")
    if sum(co) % 3:
        raise FaceletError("a corner is twisted (corner orientations do not sum to 0 mod 3)")
    if sum(eo) % 2:
        raise FaceletError("an edge is flipped (edge orientations do not sum to 0 mod 2)")
    if permutation_parity(cp) != permutation_parity(ep):
        raise FaceletError("two pieces are swapped (corner and edge permutation parities differ)")
    return CubieState(tuple(cp), tuple(co), tuple(ep), tuple(eo))


def to_facelets(state: CubieState) -> str:
    """54-sticker string (face letters) of a cubie state."""
    out = [face for face in FACES for _ in range(9)]
    for slots, names, perm, ori, n in ((CORNER_FACELETS, CORNER_NAMES, state.cp, state.co, 3),
                                       (EDGE_FACELETS, EDGE_NAMES, state.ep, state.eo, 2)):
        for slot, idx in enumerate(slots):
            name = names[perm[slot]]
            for k in range(n):
                out[idx[(k + ori[slot]) % n]] = name[k]
    return "".join(out)
