"""Serving loop on a resident table: cubes are admitted as they arrive and
searched together on the device (libcubesolver's ccs_server_* entry points)."""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

from . import _lib
from .cubie import CubieState
from .solver import Tables


@dataclass
class Served:
    tag: int
    length: int
    moves: list[int]
    nodes: int
    queued_ms: float
    latency_ms: float


class Server:
    """``state_budget_bytes`` bounds the device memory of searches in flight
    (admission stops above it); ``max_inflight`` caps them further."""

    def __init__(self, tables: Tables, state_budget_bytes: int = 0, max_inflight: int = 0,
                 prefix_len: int = 4, deeper_prefix_at=(0.0, 0.0), blocks_per_cu: int = 0, chunk_nodes: int = 0,
                 read_mode: int = 0):
        o = _lib.ServerOpts()
        o.state_budget_bytes = state_budget_bytes
        o.max_inflight = max_inflight
        o.prefix_len = prefix_len
        o.deeper_at[0], o.deeper_at[1] = deeper_prefix_at
        o.blocks_per_cu = blocks_per_cu
        o.chunk_nodes = chunk_nodes
        o.read_mode = read_mode  # a step's two table reads: 1 paired, 2 in turn (0: the library's choice)
        self._h = ctypes.c_void_p()
        self.tables = tables
        _lib.check(_lib.lib().ccs_server_start(tables.handle, ctypes.byref(o), ctypes.byref(self._h)))

    def submit(self, cube: CubieState, tag: int) -> None:
        raw = ctypes.create_string_buffer(cube.to_bytes(), 40)
        _lib.check(_lib.lib().ccs_server_submit(self._h, raw, tag))

    def poll(self, max_results: int = 256, timeout_ms: int = 0) -> list[Served]:
        buf = (_lib.Served * max_results)()
        n = _lib.check(_lib.lib().ccs_server_poll(self._h, buf, max_results, timeout_ms))
        out = []
        for r in buf[:n]:
            if r.length == -2:
                raise AssertionError(f"served solution for tag {r.tag} does not solve the cube")
            out.append(Served(int(r.tag), int(r.length), list(r.moves[: max(r.length, 0)]), int(r.nodes),
                              float(r.queued_ms), float(r.latency_ms)))
        return out

    def collect(self, n: int, timeout_s: float = 60.0) -> dict[int, Served]:
        """Wait for n results; returns them by tag."""
        got: dict[int, Served] = {}
        deadline = time.monotonic() + timeout_s
        while len(got) < n:
            left = deadline - time.monotonic()
            if left <= 0:
                raise TimeoutError(f"{len(got)}/{n} results after {timeout_s}s")
            for r in self.poll(256, int(min(left, 0.5) * 1000) or 1):
                got[r.tag] = r
        return got

    def stop(self) -> None:
        if self._h:
            h, self._h = self._h, ctypes.c_void_p()
            _lib.check(_lib.lib().ccs_server_stop(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.stop()
