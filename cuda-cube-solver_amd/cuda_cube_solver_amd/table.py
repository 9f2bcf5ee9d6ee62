"""Table build and I/O (the reference's table.py:170-200 build_preset and
load_device), on libcubesolver's BFS kernels."""
from __future__ import annotations

import time
from pathlib import Path

from . import kinds
from .solver import Tables


def build_preset(out_dir: Path, preset: kinds.Preset, log=print) -> Path:
    """Build a preset's table on the GPU and write only its solving file
    (``.tmp`` first, then renamed) plus the ``.json`` metadata."""
    out_dir = Path(out_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    t = time.perf_counter()
    tables, counts = Tables.build(preset.kind, preset.format, log=log)
    try:
        path = out_dir / preset.file
        tables.save(preset.format, path)
    finally:
        tables.close()
    kinds.write_meta(path.with_suffix(".json"), preset, counts, time.perf_counter() - t)
    return path


def load(path: Path, kind: str, fmt: str) -> Tables:
    return Tables.open(Path(path), kind, fmt)
