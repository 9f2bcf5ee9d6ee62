"""Table kinds and the memory ladder of table presets (names, files and sizes
of the reference's src/cuda_cube_solver/kinds.py:99-225, which are the
drop-in contract)."""
from __future__ import annotations

import json
from dataclasses import dataclass
from pathlib import Path

N_CLASSES = 9930
B3_ENTRIES = 155
BLOCK_ENTRIES = 124

# E-content classes of each group kind; 304 has no group (EdgeKind(3,4)).
_NEC = {"s4": 8, "d4f": 12, "a4t": 16, "a4": 16, "d4t": 24, "d4": 24}
GROUP_KINDS = tuple(_NEC)


def n_edge(kind: str) -> int:
    if kind == "304":
        return 495 * 70 * 16
    if kind not in _NEC:
        raise ValueError(f"unknown table kind {kind!r}")
    return 495 * 256 * _NEC[kind]


@dataclass(frozen=True)
class Preset:
    """A solving table: a kind in block format (2-bit offsets) or base 3."""

    kind: str
    format: str  # "block" or "b3"
    cubes_per_s: float  # the reference's 1,000-cube figure (RTX 5060 Ti)
    superflip_s: float
    auto: bool = True

    @property
    def name(self) -> str:
        return self.kind if self.format == "block" else f"{self.kind}-b3"

    @property
    def file(self) -> str:
        return f"prune-ud-{self.kind}-{'blk' if self.format == 'block' else 'b3'}.bin"

    @property
    def entries(self) -> int:
        return N_CLASSES * n_edge(self.kind)

    @property
    def table_bytes(self) -> int:
        per = BLOCK_ENTRIES if self.format == "block" else B3_ENTRIES
        return -(-self.entries // per) * 32

    @property
    def build_bytes(self) -> int:
        """Peak device memory to build it here: the exact-depth nibble table
        (4 bits per entry) and the solving table."""
        return (self.entries + 7) // 8 * 4 + self.table_bytes


PRESETS = [
    Preset("304", "block", 6.0, 44.6),
    Preset("s4", "b3", 13.1, 6.2),
    Preset("d4f", "b3", 16.9, 77.7),
    Preset("a4t", "b3", 21.8, 27.1),
    Preset("a4t", "block", 24.1, 11.5),
    Preset("d4t", "b3", 35.2, 47.5),
    Preset("d4t", "block", 37.5, 16.0),
    Preset("a4", "b3", 20.7, 4.0, auto=False),
    Preset("a4", "block", 22.2, 4.2, auto=False),
    Preset("d4", "b3", 31.6, 3.2, auto=False),
    Preset("d4", "block", 33.3, 3.2, auto=False),
]
PRESET_BY_NAME = {p.name: p for p in PRESETS}

SOLVE_MARGIN = 400 << 20  # device memory needed besides the table


def built(preset: Preset, tables: Path) -> bool:
    return (Path(tables) / preset.file).exists()


def choose(tables: Path, memory: int, build_memory: int | None = None,
           to_build: bool = False) -> tuple[Preset | None, Preset | None]:
    """The fastest preset whose table fits in ``memory`` bytes and is built or can be
    built in ``build_memory``, and the fastest one that fits but cannot be built
    here (None if not faster)."""
    build_memory = memory if build_memory is None else build_memory
    fits = [p for p in PRESETS if p.auto and p.table_bytes + SOLVE_MARGIN <= memory]
    usable = [p for p in fits if (built(p, tables) and not to_build) or p.build_bytes + SOLVE_MARGIN <= build_memory]
    best = usable[-1] if usable else None
    better = fits[-1] if fits and fits[-1] is not best else None
    return best, better


def pick(tables: Path, memory="auto", name: str | None = None,
         to_build: bool = False) -> tuple[Preset | None, Preset | None]:
    """``choose`` for a budget in GB or "auto" (the free device memory), or the preset ``name``."""
    if name:
        if name not in PRESET_BY_NAME:
            raise ValueError(f"unknown table {name!r}; choose from {', '.join(PRESET_BY_NAME)}")
        return PRESET_BY_NAME[name], None
    from . import _lib

    free = int(_lib.device_info().free_bytes)
    budget = free if memory == "auto" else min(free, int(float(memory) * 1e9))
    return choose(Path(tables), budget, free, to_build)


def describe(tables: Path) -> list[str]:
    rows = [f"{'preset':8} {'VRAM':>8} {'cubes/s':>8} {'superflip':>10} {'build VRAM':>11}  file"]
    for p in PRESETS:
        rows.append(f"{p.name:8} {p.table_bytes / 1e9:6.2f}GB {p.cubes_per_s:8.1f} {p.superflip_s:9.1f}s "
                    f"{p.build_bytes / 1e9:9.1f}GB  {p.file}{'' if built(p, tables) else '  (not built)'}"
                    f"{'' if p.auto else '  (only with --table)'}")
    return rows


def write_meta(path: Path, preset: Preset, counts: list[int], seconds: float) -> None:
    path.write_text(json.dumps({"preset": preset.name, "entries": preset.entries, "depth_counts": counts,
                                "seconds": round(seconds, 1)}, indent=2) + "\n")
