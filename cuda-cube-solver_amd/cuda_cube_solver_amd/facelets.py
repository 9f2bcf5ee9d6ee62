"""54-sticker facelet strings <-> cubie states.

The string layout is the common one used by the reference
(src/cuda_cube_solver/facelets.py:1-41) and the kociemba package: faces in
U, R, F, D, L, B order, nine stickers each; any six symbols may be used, the
centres define the faces.
"""
from __future__ import annotations

from .cubie import CubieState, permutation_parity

FACES = "URFDLB"
# Sticker indices of each corner position (U/D sticker first, then clockwise) and edge
# position; this is the string format itself, identical to the reference's.
CORNER_FACELETS = ((8, 9, 20), (6, 18, 38), (0, 36, 47), (2, 45, 11),
                   (29, 26, 15), (27, 44, 24), (33, 53, 42), (35, 17, 51))
EDGE_FACELETS = ((5, 10), (7, 19), (3, 37), (1, 46), (32, 16), (28, 25),
                 (30, 43), (34, 52), (23, 12), (21, 41), (50, 39), (48, 14))
CORNER_FACES = ("URF", "UFL", "ULB", "UBR", "DFR", "DLF", "DBL", "DRB")
EDGE_FACES = ("UR", "UF", "UL", "UB", "DR", "DF", "DL", "DB", "FR", "FL", "BL", "BR")


class FaceletError(ValueError):
    """The string does not describe a legal cube."""


def is_facelet_string(text: str) -> bool:
    s = "".join(text.split())
    return len(s) == 54 and len(set(s)) == 6


def parse_facelets(text: str) -> CubieState:
    """Cubie state of a 54-sticker string; FaceletError if it is not a legal cube."""
    s = "".join(text.split())
    if len(s) != 54:
        raise FaceletError(f"expected 54 stickers, got {len(s)}")
    centres = [s[4 + 9 * f] for f in range(6)]
    if len(set(centres)) != 6:
        raise FaceletError("the six centre stickers must all differ")
    face_of = dict(zip(centres, FACES))
    bad = next((c for c in s if c not in face_of), None)
    if bad is not None:
        raise FaceletError(f"sticker {bad!r} matches no centre")
    faces = [face_of[c] for c in s]
    for f in FACES:
        k = faces.count(f)
        if k != 9:
            raise FaceletError(f"face {f} has {k} stickers, expected 9")
    cp, co = [], []
    for pos, slots in enumerate(CORNER_FACELETS):
        st = [faces[i] for i in slots]
        twist = next((k for k in range(3) if st[k] in "UD"), None)
        if twist is None:
            raise FaceletError(f"corner at {CORNER_FACES[pos]} has no U or D sticker")
        key = st[twist] + st[(twist + 1) % 3] + st[(twist + 2) % 3]
        if key not in CORNER_FACES:
            raise FaceletError(f"corner at {CORNER_FACES[pos]} has impossible colours {''.join(st)}")
        cp.append(CORNER_FACES.index(key))
        co.append(twist)
    ep, eo = [], []
    for pos, (a, b) in enumerate(EDGE_FACELETS):
        pair = faces[a] + faces[b]
        if pair in EDGE_FACES:
            ep.append(EDGE_FACES.index(pair))
            eo.append(0)
        elif pair[::-1] in EDGE_FACES:
            ep.append(EDGE_FACES.index(pair[::-1]))
            eo.append(1)
        else:
            raise FaceletError(f"edge at {EDGE_FACES[pos]} has impossible colours {pair}")
    if sorted(cp) != list(range(8)):
        raise FaceletError("some corner appears twice")
    if sorted(ep) != list(range(12)):
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
    out = list("".join(f * 9 for f in FACES))
    for pos, slots in enumerate(CORNER_FACELETS):
        name = CORNER_FACES[state.cp[pos]]
        for k in range(3):
            out[slots[(k + state.co[pos]) % 3]] = name[k]
    for pos, slots in enumerate(EDGE_FACELETS):
        name = EDGE_FACES[state.ep[pos]]
        for k in range(2):
            out[slots[(k + state.eo[pos]) % 2]] = name[k]
    return "".join(out)
