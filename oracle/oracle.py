"""ctypes wrapper of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

CUBE_DTYPE = np.dtype([("cp", np.uint8, 8), ("co", np.uint8, 8), ("ep", np.uint8, 12), ("eo", np.uint8, 12)])


class OKind(ctypes.Structure):
    _fields_ = [("is304", ctypes.c_int), ("nec", ctypes.c_int), ("n_edge", ctypes.c_uint64),
                ("n_entries", ctypes.c_uint64), ("gen_class", ctypes.c_uint16 * 384),
                ("gen_rep", (ctypes.c_uint16 * 384) * 2)]


class OSolveOpts(ctypes.Structure):
    _fields_ = [("prefix_len", ctypes.c_int), ("deeper_at", ctypes.c_double * 2), ("max_bound", ctypes.c_int)]


class OSolveResult(ctypes.Structure):
    _fields_ = [("length", ctypes.c_int), ("moves", ctypes.c_uint8 * 24), ("n_rounds", ctypes.c_int),
                ("round_bound", ctypes.c_uint32 * 24), ("round_nodes", ctypes.c_uint64 * 24),
                ("nodes", ctypes.c_uint64), ("root_h", ctypes.c_uint32 * 6)]


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        P = ctypes.c_void_p
        L.or_init.restype = ctypes.c_int
        L.or_table_index.restype = ctypes.c_uint64
        L.or_table_index.argtypes = [P, P]
        L.or_unrank.argtypes = [P, ctypes.c_uint64, P]
        L.or_corner_raw.restype = ctypes.c_uint32
        L.or_corner_raw.argtypes = [P]
        L.or_classidx.restype = P
        L.or_class_rep.restype = P
        L.or_class_stab.restype = P
        L.or_kind_init.argtypes = [P, ctypes.c_char_p]
        L.or_group_kind.argtypes = [P, P, ctypes.c_int]
        L.or_bfs.argtypes = [P, P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_histogram.argtypes = [P, ctypes.c_uint64, P]
        L.or_build_b3.argtypes = [P, P, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, P]
        L.or_lookup.restype = ctypes.c_uint32
        L.or_lookup.argtypes = [P, ctypes.c_int, ctypes.c_uint64]
        L.or_solve.argtypes = [P, P, ctypes.c_int, P, P, P]
        L.or_solve_many.argtypes = [P, P, ctypes.c_int, P, ctypes.c_int, P, P, ctypes.c_int, P]
        L.or_canonical_prefixes.restype = ctypes.c_int64
        L.or_canonical_prefixes.argtypes = [ctypes.c_int, P, ctypes.c_int64]
        L.or_set_deadline.argtypes = [ctypes.c_double]
        L.or_sym_conjugate.argtypes = [ctypes.c_int, P, P]
        L.or_sym_matrix.argtypes = [ctypes.c_int, P]
        L.or_axis_frames.argtypes = [P, P]
        L.or_compose.argtypes = [P, P, P]
        L.or_inverse.argtypes = [P, P]
        L.or_move.argtypes = [ctypes.c_int, P]
        L.or_init()
        _lib = L
    return _lib


def cubes(states) -> np.ndarray:
    """(cp, co, ep, eo) tuples or dicts -> structured array."""
    arr = np.zeros(len(states), dtype=CUBE_DTYPE)
    for i, s in enumerate(states):
        if isinstance(s, dict):
            s = (s["cp"], s["co"], s["ep"], s["eo"])
        arr[i] = tuple(np.asarray(x, dtype=np.uint8) for x in s)
    return arr


def ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def kind(name: str) -> OKind:
    k = OKind()
    rc = lib().or_kind_init(ctypes.byref(k), name.encode())
    if rc:
        raise ValueError(f"oracle: unknown kind {name!r} ({rc})")
    return k


def group_kind(gens) -> OKind:
    k = OKind()
    flat = np.array([list(p) + list(f) for p, f in gens], dtype=np.int32).reshape(-1)
    rc = lib().or_group_kind(ctypes.byref(k), ptr(flat), len(gens))
    if rc:
        raise ValueError("not normalised")
    return k


def n_classes() -> int:
    return lib().or_init()


def classidx() -> np.ndarray:
    p = lib().or_classidx()
    return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_uint32)), shape=(153090,)).copy()


def class_rep() -> np.ndarray:
    p = lib().or_class_rep()
    return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_uint32)), shape=(n_classes(),)).copy()


def class_stab() -> np.ndarray:
    p = lib().or_class_stab()
    return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_uint16)), shape=(n_classes(),)).copy()


def table_index(k: OKind, arr: np.ndarray) -> np.ndarray:
    L = lib()
    out = np.zeros(len(arr), dtype=np.uint64)
    base = ptr(arr)
    for i in range(len(arr)):
        out[i] = L.or_table_index(ctypes.byref(k), base + i * CUBE_DTYPE.itemsize)
    return out


def unrank(k: OKind, idx) -> np.ndarray:
    L = lib()
    idx = np.asarray(idx, dtype=np.uint64)
    out = np.zeros(len(idx), dtype=CUBE_DTYPE)
    base = ptr(out)
    for i, x in enumerate(idx):
        L.or_unrank(ctypes.byref(k), int(x), base + i * CUBE_DTYPE.itemsize)
    return out


def compose(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    out = np.zeros(1, dtype=CUBE_DTYPE)
    lib().or_compose(ptr(a), ptr(b), ptr(out))
    return out


def move(m: int) -> np.ndarray:
    out = np.zeros(1, dtype=CUBE_DTYPE)
    lib().or_move(m, ptr(out))
    return out


def sym_conjugate(sym48: int, a: np.ndarray) -> np.ndarray:
    out = np.zeros(1, dtype=CUBE_DTYPE)
    lib().or_sym_conjugate(sym48, ptr(a), ptr(out))
    return out


def sym_matrix(sym48: int) -> list[list[int]]:
    out = np.zeros(9, dtype=np.int32)
    lib().or_sym_matrix(sym48, ptr(out))
    return out.reshape(3, 3).tolist()


def axis_frames(a: np.ndarray) -> np.ndarray:
    out = np.zeros(6, dtype=CUBE_DTYPE)
    lib().or_axis_frames(ptr(a), ptr(out))
    return out


def canonical_prefixes(length: int) -> np.ndarray:
    L = lib()
    n = L.or_canonical_prefixes(length, None, 0)
    out = np.zeros((n, length), dtype=np.uint8)
    L.or_canonical_prefixes(length, ptr(out), n)
    return out


def bfs(k: OKind, n_threads: int = 0, verbose: bool = False):
    """(mod-3 table, block table, depth counts) of a kind, built on the CPU."""
    n_words = (k.n_entries + 15) // 16
    n_blocks = -(-k.n_entries // 124)
    mod3 = np.empty(n_words, dtype=np.uint32)
    blk = np.empty(n_blocks * 8, dtype=np.uint32)
    counts = np.zeros(32, dtype=np.uint64)
    nd = lib().or_bfs(ctypes.byref(k), ptr(mod3), ptr(blk), ptr(counts), 32, n_threads, int(verbose))
    if nd < 0:
        raise RuntimeError(f"oracle bfs failed ({nd})")
    return mod3, blk, [int(c) for c in counts[:nd]]


def histogram(mod3: np.ndarray, n_entries: int) -> list[int]:
    out = np.zeros(4, dtype=np.uint64)
    lib().or_histogram(ptr(mod3), n_entries, ptr(out))
    return [int(c) for c in out]


def build_b3(mod3: np.ndarray, blk: np.ndarray, n_entries: int) -> np.ndarray:
    nb = -(-n_entries // 155)
    out = np.empty(nb * 8, dtype=np.uint32)
    lib().or_build_b3(ptr(mod3), ptr(blk), 0, nb, n_entries, ptr(out))
    return out


def lookup(table: np.ndarray, fmt: int, idx: int) -> int:
    return lib().or_lookup(ptr(table), fmt, int(idx))


def solve_many(k: OKind, table: np.ndarray, fmt: int, arr: np.ndarray, prefix_len: int = 4,
               deeper_at=(5e7, 3e9), n_threads: int = 0, max_bound: int = 24):
    """Optimal solutions with the reference's step-mode pruning and node counting."""
    opts = OSolveOpts(prefix_len, (ctypes.c_double * 2)(*deeper_at), max_bound)
    res = (OSolveResult * len(arr))()
    secs = np.zeros(len(arr), dtype=np.float64)
    bad = lib().or_solve_many(ctypes.byref(k), ptr(table), fmt, ptr(arr), len(arr), ctypes.byref(opts),
                              ctypes.cast(res, ctypes.c_void_p), n_threads, ptr(secs))
    out = []
    for i, r in enumerate(res):
        out.append({"length": r.length, "moves": list(r.moves[: max(r.length, 0)]),
                    "rounds": [(int(r.round_bound[j]), int(r.round_nodes[j])) for j in range(r.n_rounds)],
                    "nodes": int(r.nodes), "root_h": list(r.root_h), "seconds": float(secs[i])})
    return out, bad


def timed_nodes(k: OKind, table: np.ndarray, fmt: int, arr: np.ndarray, seconds: float, n_threads: int):
    """Nodes the oracle visits on these cubes (one per thread) within a wall-clock budget."""
    import time

    lib().or_set_deadline(seconds)
    t = time.perf_counter()
    try:
        res, _ = solve_many(k, table, fmt, arr, n_threads=n_threads)
    finally:
        lib().or_set_deadline(0.0)
    return sum(r["nodes"] for r in res), time.perf_counter() - t, res


def complete_round_nodes(result: dict) -> int:
    """Nodes in rounds whose bound is below the solution length (hard_survey.py:51-52)."""
    return sum(n for b, n in result["rounds"] if b < result["length"])


def threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
