"""Optimal HTM solver on a resident table (the reference's solver.py API:
open_solver, GpuSolver.solve, Solution, TableNotFound, NoTableFits).

The search runs in libcubesolver's HIP kernels; this module prepares cubes
and results.  ``mode="plain"`` is the reference's per-round structure
(solver.py:254-320), ``mode="batched"`` the persistent form that keeps every
cube's IDA* rounds on the device.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field
from pathlib import Path

from . import _lib, kinds
from .cubie import MOVE_NAMES_HTM, CubieState, parse_moves  # noqa: F401  (re-exported like the reference)


@dataclass
class Solution:
    moves: list[int]
    nodes: int
    seconds: float
    iteration_nodes: list[tuple[int, int]] = field(default_factory=list)  # (bound, nodes)

    def text(self) -> str:
        return " ".join(MOVE_NAMES_HTM[m] for m in self.moves)


class TableNotFound(FileNotFoundError):
    """A pruning table file is missing."""


class NoTableFits(RuntimeError):
    """No table preset fits in the device memory budget."""


def _table_path(directory: Path, filename: str) -> Path:
    """``directory / filename``, or TableNotFound with the command that builds
    it (the message of the reference's solver.py:82-88)."""
    path = directory / filename
    if path.exists():
        return path
    hint = ""
    stem = filename[len("prune-ud-"):-len(".bin")] if filename.startswith("prune-ud-") and filename.endswith(".bin") else ""
    for suffix, keep in (("-blk", ""), ("-b3", "-b3")):
        if stem.endswith(suffix) and stem[: -len(suffix)]:
            hint = (f"; build it with `cuda-cube-solver build-table --table {stem[: -len(suffix)]}{keep} "
                    f"--out {directory}`")
            break
    raise TableNotFound(f"{path} not found{hint}")


def open_solver(tables: Path = Path("tables"), memory="auto", table: str | None = None, build: bool = True,
                log=print) -> "GpuSolver":
    """A solver on the preset ``table``, or on the fastest one that fits
    ``memory`` (GB, or "auto" for the free device memory).  A missing table
    file is built first, unless ``build`` is False."""
    from . import table as table_io

    directory = Path(tables)
    if directory.suffix == ".bin":  # a table file was named: use its directory
        directory = directory.parent
    preset, _ = kinds.pick(directory, memory, table)
    if preset is None:
        raise NoTableFits("no table fits in that GPU memory (the smallest, 304, needs 1.8 GB)")
    target = directory / preset.file
    if not target.exists():
        if not build:
            raise TableNotFound(f"{target} not found")
        gb = 1e9
        log(f"building table {preset.name} ({preset.table_bytes / gb:.2f} GB, needs "
            f"{preset.build_bytes / gb:.1f} GB of device memory) in {directory}/")
        written = table_io.build_preset(directory, preset, log=lambda line: log("  " + line))
        log(f"wrote {written}")
    return GpuSolver(directory, kind=preset.kind, table_format=preset.format)


class Tables:
    """A resident table handle (device memory owned by the library)."""

    def __init__(self, handle: int, kind: str, fmt: str):
        self._h = ctypes.c_void_p(handle)
        self.kind = kind
        self.format = fmt

    @classmethod
    def open(cls, path: Path, kind: str, fmt: str) -> "Tables":
        h = ctypes.c_void_p()
        code = _lib.FORMAT_B3 if fmt == "b3" else _lib.FORMAT_BLOCK
        rc = _lib.lib().ccs_tables_open(str(path).encode(), kind.encode(), code, ctypes.byref(h))
        if rc == _lib.CCS_ENOENT:
            raise TableNotFound(_lib.lib().ccs_last_error().decode())
        if rc == _lib.CCS_EIO:
            raise ValueError(_lib.lib().ccs_last_error().decode())
        _lib.check(rc)
        return cls(h.value, kind, fmt)

    @classmethod
    def build(cls, kind: str, fmt: str, log=None, keep_depths: bool = False) -> tuple["Tables", list[int]]:
        """BFS on the device; fmt "block", "b3" or "both" (both kept resident).
        ``keep_depths`` also keeps the exact-depth build table (read("depths", ...))."""
        h = ctypes.c_void_p()
        counts = (ctypes.c_uint64 * 32)()
        nd = ctypes.c_int(32)
        cb = _lib.LOG_FN(lambda line, user: log(line.decode()) if log else None)
        code = {"block": 0, "b3": 1, "both": 2}[fmt] | (_lib.BUILD_KEEP_DEPTHS if keep_depths else 0)
        rc = _lib.lib().ccs_tables_build(kind.encode(), code, cb, None, counts, ctypes.byref(nd), ctypes.byref(h))
        if rc == _lib.CCS_EFAULT:
            raise AssertionError(_lib.lib().ccs_last_error().decode())
        _lib.check(rc)
        return cls(h.value, kind, "block" if fmt == "both" else fmt), [int(c) for c in counts[: nd.value]]

    @property
    def handle(self) -> ctypes.c_void_p:
        if not self._h:
            raise RuntimeError("tables handle is closed")
        return self._h

    def use(self, fmt: str) -> None:
        """Search with the resident block or b3 table (both are kept after build(..., "both"))."""
        _lib.check(_lib.lib().ccs_tables_use(self.handle, 1 if fmt == "b3" else 0))
        self.format = fmt

    @staticmethod
    def _code(fmt: str) -> int:
        return {"block": 0, "b3": 1, "depths": _lib.TABLE_DEPTHS}[fmt]

    def nbytes(self, fmt: str | None = None) -> int:
        return int(_lib.lib().ccs_tables_bytes(self.handle, self._code(fmt or self.format)))

    def read(self, fmt: str, offset: int, n: int) -> bytes:
        """n bytes at offset of the "block" or "b3" table, or of the kept "depths" build table."""
        buf = ctypes.create_string_buffer(n)
        _lib.check(_lib.lib().ccs_tables_read(self.handle, self._code(fmt), offset, n, buf))
        return buf.raw

    def save(self, fmt: str, path: Path) -> None:
        _lib.check(_lib.lib().ccs_tables_save(self.handle, 1 if fmt == "b3" else 0, str(path).encode()))

    def close(self) -> None:
        if self._h:
            _lib.lib().ccs_tables_close(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _pack(states) -> ctypes.Array:
    raw = b"".join(s.to_bytes() for s in states)
    return ctypes.create_string_buffer(raw, len(raw))


class GpuSolver:
    """Batched optimal solver over one resident table.

    ``mode``: "plain" (per-round launches and host synchronisation, the
    reference's step mode) or "batched" (one persistent launch per solve()).
    ``prefix_len``/``deeper_prefix_at`` split rounds into work items as the
    reference does (solver.py:119-123, 189-195)."""

    def __init__(self, tables: Path = Path("tables"), kind: str = "d4t", table_format: str = "block",
                 mode: str = "plain", prefix_len: int = 4, deeper_prefix_at=(5e7, 3e9), blocks_per_cu: int = 0,
                 poll_interval: int = 64, chunk_nodes: int = 0, table: Tables | None = None, max_bound: int = 0,
                 block: int = 0, probes_per_step: int = 0, read_mode: int = 0):
        if table is None:
            tables = Path(tables)
            if tables.suffix == ".bin":
                tables = tables.parent
            name = f"prune-ud-{kind}-{'b3' if table_format == 'b3' else 'blk'}.bin"
            table = Tables.open(_table_path(tables, name), kind, table_format)
        self.table = table
        self.kind = table.kind
        self.mode = mode
        self.prefix_len = prefix_len
        self.deeper_prefix_at = tuple(deeper_prefix_at)
        self.blocks_per_cu = blocks_per_cu
        self.poll_interval = poll_interval
        self.chunk_nodes = chunk_nodes
        self.max_bound = max_bound  # 0: no limit below the longest optimal solution (20)
        self.block = block  # lanes per workgroup (0: the largest the kernels are built for)
        self.probes_per_step = probes_per_step  # 1 or 2 table probes in flight per lane (0: 2)
        self.read_mode = read_mode  # a step's two table reads: 1 paired, 2 in turn (0: by form and batch size)
        self.last_stats: _lib.SolveStats | None = None

    def _opts(self, mode: str | None = None) -> _lib.SolveOpts:
        o = _lib.SolveOpts()
        o.mode = {"plain": 0, "batched": 1}[mode or self.mode]
        o.prefix_len = self.prefix_len
        o.deeper_at[0], o.deeper_at[1] = self.deeper_prefix_at
        o.block = self.block
        o.blocks_per_cu = self.blocks_per_cu
        o.max_bound = self.max_bound
        o.poll_interval = self.poll_interval
        o.chunk_nodes = self.chunk_nodes
        o.probes_per_step = self.probes_per_step
        o.read_mode = self.read_mode
        return o

    def solve(self, states: list[CubieState], log=None, mode: str | None = None) -> list[Solution]:
        n = len(states)
        if n == 0:
            return []
        res = (_lib.Result * n)()
        stats = _lib.SolveStats()
        t0 = time.perf_counter()
        rc = _lib.lib().ccs_solve(self.table.handle, _pack(states), n, ctypes.byref(self._opts(mode)), res,
                                  ctypes.byref(stats))
        if rc == _lib.CCS_EFAULT:
            raise AssertionError(_lib.lib().ccs_last_error().decode())
        _lib.check(rc)
        self.last_stats = stats
        out = []
        for r in res:
            hist = [(int(r.round_bound[j]), int(r.round_nodes[j])) for j in range(r.n_rounds)]
            out.append(Solution(list(r.moves[: r.length]), int(r.nodes), float(r.seconds), hist))
        if log:
            log(f"{time.perf_counter() - t0:8.2f}s solved {n}/{n}")
        return out
