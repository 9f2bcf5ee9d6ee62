"""Table kinds and the ladder of table presets.

The names, files, sizes and published figures of the presets are the
reference's drop-in contract (src/cuda_cube_solver/kinds.py:115-225): a
``--table`` name, a ``prune-ud-*.bin`` file and a ``.json`` beside it mean the
same table here.  What differs is what a build needs: the BFS here keeps one
exact depth per entry in a nibble (the reference keeps depth mod 3 in two bits
plus its block table), so ``build_bytes`` is this library's own peak, and on a
small card the ladder can settle on a smaller preset than the reference would.
"""
from __future__ import annotations

import json
from dataclasses import dataclass
from pathlib import Path

N_CLASSES = 9930       # corner classes under the 16 UD symmetries
BLOCK_ENTRIES = 124    # entries per 32-byte block of the block format
B3_ENTRIES = 155       # entries per 32-byte block of the base-3 format

# Edge factor of each kind: E-slice positions (495) x the rest.  304 keeps the
# positions of the U-layer edges (70) and the E-edge flips (16); a group kind
# keeps the 8 non-E flips (256) and the E content up to its group, in `nec`
# classes (the reference's GROUP_KINDS, kinds.py:100-107; the class counts are
# checked against the library's own group closure in the tests).
_EDGE_FACTOR = {"304": 70 * 16, "s4": 256 * 8, "d4f": 256 * 12, "a4t": 256 * 16, "a4": 256 * 16,
                "d4t": 256 * 24, "d4": 256 * 24}
GROUP_KINDS = tuple(k for k in _EDGE_FACTOR if k != "304")


def n_edge(kind: str) -> int:
    try:
        return 495 * _EDGE_FACTOR[kind]
    except KeyError:
        raise ValueError(f"unknown table kind {kind!r}") from None


def _blocks(entries: int, per: int) -> int:
    return (entries + per - 1) // per


@dataclass(frozen=True)
class Preset:
    """A solving table: a kind in the block format (2-bit offsets) or in base 3."""

    kind: str
    format: str           # "block" or "b3"
    cubes_per_s: float    # the reference's figure: 1,000 random cubes on an RTX 5060 Ti
    superflip_s: float    # the reference's superflip time on that card
    auto: bool = True     # False: used only when named

    @property
    def name(self) -> str:
        return self.kind + ("-b3" if self.format == "b3" else "")

    @property
    def file(self) -> str:
        return "prune-ud-{}-{}.bin".format(self.kind, "b3" if self.format == "b3" else "blk")

    @property
    def entries(self) -> int:
        return N_CLASSES * n_edge(self.kind)

    @property
    def table_bytes(self) -> int:
        """Bytes of the solving table, on disk and in device memory."""
        return 32 * _blocks(self.entries, B3_ENTRIES if self.format == "b3" else BLOCK_ENTRIES)

    @property
    def build_bytes(self) -> int:
        """Peak device memory of building it here (ccs_tables_build): the
        exact-depth build table (a nibble per entry) and the solving table; a
        base-3 build also holds one minimum byte per 124-entry block."""
        nibbles = 4 * _blocks(self.entries, 8)
        extra = _blocks(self.entries, BLOCK_ENTRIES) if self.format == "b3" else 0
        return nibbles + self.table_bytes + extra


def _ladder(rows):
    return [Preset(kind, fmt, rate, sf, auto) for kind, fmt, rate, sf, auto in rows]


# Smallest first; each step up is faster on random cubes.  The last four are the
# robust variants (E-slice flips kept exactly), chosen only by name.
PRESETS = _ladder([
    ("304", "block", 6.0, 44.6, True),
    ("s4", "b3", 13.1, 6.2, True),
    ("d4f", "b3", 16.9, 77.7, True),
    ("a4t", "b3", 21.8, 27.1, True),
    ("a4t", "block", 24.1, 11.5, True),
    ("d4t", "b3", 35.2, 47.5, True),
    ("d4t", "block", 37.5, 16.0, True),
    ("a4", "b3", 20.7, 4.0, False),
    ("a4", "block", 22.2, 4.2, False),
    ("d4", "b3", 31.6, 3.2, False),
    ("d4", "block", 33.3, 3.2, False),
])
PRESET_BY_NAME = {p.name: p for p in PRESETS}

SOLVE_MARGIN = 400 << 20  # device memory a solve needs besides the table


def built(preset: Preset, tables: Path) -> bool:
    return (Path(tables) / preset.file).exists()


def choose(tables: Path, memory: int, build_memory: int | None = None,
           to_build: bool = False) -> tuple[Preset | None, Preset | None]:
    """(best, better): the fastest automatic preset whose table fits ``memory``
    bytes and is on disk or can be built in ``build_memory`` (default
    ``memory``; ``to_build`` ignores files on disk), and the fastest one that
    fits but cannot be built here, if that is a different one."""
    cap = memory if build_memory is None else build_memory
    ladder = [p for p in PRESETS if p.auto and p.table_bytes + SOLVE_MARGIN <= memory]

    def usable(p: Preset) -> bool:
        return (not to_build and built(p, tables)) or p.build_bytes + SOLVE_MARGIN <= cap

    best = next((p for p in reversed(ladder) if usable(p)), None)
    top = ladder[-1] if ladder else None
    return best, (top if top is not None and top is not best else None)


def pick(tables: Path, memory="auto", name: str | None = None,
         to_build: bool = False) -> tuple[Preset | None, Preset | None]:
    """The named preset, or ``choose`` for ``memory`` in GB or "auto" (the free
    device memory, which also caps a larger figure)."""
    if name:
        if name not in PRESET_BY_NAME:
            raise ValueError(f"unknown table {name!r}; choose from {', '.join(PRESET_BY_NAME)}")
        return PRESET_BY_NAME[name], None
    from . import _lib

    free = int(_lib.device_info().free_bytes)
    budget = free if memory == "auto" else min(free, int(float(memory) * 1e9))
    return choose(Path(tables), budget, free, to_build)


def describe(tables: Path) -> list[str]:
    """The `tables` listing: one row per preset (the reference's columns)."""
    head = "{:8} {:>8} {:>8} {:>10} {:>11}  file".format("preset", "VRAM", "cubes/s", "superflip", "build VRAM")
    rows = [head]
    for p in PRESETS:
        notes = ("" if built(p, tables) else "  (not built)") + ("" if p.auto else "  (only with --table)")
        rows.append("{:8} {:6.2f}GB {:8.1f} {:9.1f}s {:9.1f}GB  {}{}".format(
            p.name, p.table_bytes / 1e9, p.cubes_per_s, p.superflip_s, p.build_bytes / 1e9, p.file, notes))
    return rows


def write_meta(path: Path, preset: Preset, counts: list[int], seconds: float) -> None:
    """The ``.json`` beside a table file (the reference's layout)."""
    meta = {"preset": preset.name, "entries": preset.entries, "depth_counts": list(counts),
            "seconds": round(seconds, 1)}
    Path(path).write_text(json.dumps(meta, indent=2) + "\n")
