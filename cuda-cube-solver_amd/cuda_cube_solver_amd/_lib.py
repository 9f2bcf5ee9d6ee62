"""ctypes binding of libcubesolver.so (declared in include/cubesolver.h).

The library is built in-tree (``make -C cuda-cube-solver_amd``); importing a
function that needs it without the library raises LibraryMissing, there is no
other implementation to fall back to.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

PKG_ROOT = Path(__file__).resolve().parents[1]
LIB_PATH = Path(os.environ.get("CCS_LIBRARY", PKG_ROOT / "libcubesolver.so"))

CCS_OK = 0
CCS_ENOENT = -2
CCS_EIO = -5
CCS_ENOMEM = -12
CCS_EFAULT = -14
CCS_EBUSY = -16
CCS_EINVAL = -22
CCS_EDEVICE = -100
NOALLOC_MAX_CUBES = 256
FORMAT_BLOCK = 0
FORMAT_B3 = 1
TABLE_DEPTHS = 2
BUILD_KEEP_DEPTHS = 4


class LibraryMissing(RuntimeError):
    """libcubesolver.so was not built (run `make -C cuda-cube-solver_amd`)."""


class CcsError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class DeviceInfo(ctypes.Structure):
    _fields_ = [("free_bytes", ctypes.c_uint64), ("total_bytes", ctypes.c_uint64),
                ("compute_units", ctypes.c_int32), ("device", ctypes.c_int32),
                ("name", ctypes.c_char * 64), ("arch", ctypes.c_char * 32)]


class KindInfo(ctypes.Structure):
    _fields_ = [("n_edge", ctypes.c_uint64), ("n_entries", ctypes.c_uint64),
                ("n_classes", ctypes.c_int32), ("nec", ctypes.c_int32)]


class SolveOpts(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("prefix_len", ctypes.c_int32), ("deeper_at", ctypes.c_double * 2),
                ("block", ctypes.c_int32), ("blocks_per_cu", ctypes.c_int32), ("max_bound", ctypes.c_int32),
                ("poll_interval", ctypes.c_int32), ("chunk_nodes", ctypes.c_int32),
                ("probes_per_step", ctypes.c_int32), ("read_mode", ctypes.c_int32)]


class Result(ctypes.Structure):
    _fields_ = [("length", ctypes.c_int32), ("moves", ctypes.c_uint8 * 24), ("n_rounds", ctypes.c_uint32),
                ("round_bound", ctypes.c_uint32 * 24), ("round_nodes", ctypes.c_uint64 * 24),
                ("nodes", ctypes.c_uint64), ("seconds", ctypes.c_double)]


class SolveStats(ctypes.Structure):
    _fields_ = [("probes", ctypes.c_uint64), ("nodes", ctypes.c_uint64), ("kernel_ms", ctypes.c_double),
                ("launches", ctypes.c_uint32), ("wall_ms", ctypes.c_double), ("diag_built", ctypes.c_int32),
                ("diag", ctypes.c_uint64 * 32)]


class ServerOpts(ctypes.Structure):
    _fields_ = [("state_budget_bytes", ctypes.c_uint64), ("max_inflight", ctypes.c_int32),
                ("prefix_len", ctypes.c_int32), ("deeper_at", ctypes.c_double * 2),
                ("blocks_per_cu", ctypes.c_int32), ("chunk_nodes", ctypes.c_int32), ("read_mode", ctypes.c_int32)]


class Served(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_uint64), ("length", ctypes.c_int32), ("moves", ctypes.c_uint8 * 24),
                ("nodes", ctypes.c_uint64), ("queued_ms", ctypes.c_double), ("latency_ms", ctypes.c_double)]


LOG_FN = ctypes.CFUNCTYPE(None, ctypes.c_char_p, ctypes.c_void_p)

# name -> (restype, argtypes); every symbol include/cubesolver.h declares
P, I, U64, CP = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_char_p
SIGNATURES = {
    "ccs_version": (I, []),
    "ccs_last_error": (CP, []),
    "ccs_device_info": (I, [P]),
    "ccs_kind_info": (I, [CP, P]),
    "ccs_tables_open": (I, [CP, CP, I, P]),
    "ccs_tables_build": (I, [CP, I, LOG_FN, P, P, P, P]),
    "ccs_tables_save": (I, [P, I, CP]),
    "ccs_tables_use": (I, [P, I]),
    "ccs_tables_bytes": (U64, [P, I]),
    "ccs_tables_read": (I, [P, I, U64, U64, P]),
    "ccs_tables_close": (None, [P]),
    "ccs_probe": (I, [P, P, I, I, P, P, P]),
    "ccs_unrank_check": (I, [P, P, I, P]),
    "ccs_solve": (I, [P, P, I, P, P, P]),
    "ccs_solve_noalloc": (I, [P, P, I, P, P, P]),
    "ccs_server_start": (I, [P, P, P]),
    "ccs_server_submit": (I, [P, P, U64]),
    "ccs_server_poll": (I, [P, P, I, I]),
    "ccs_server_stop": (I, [P]),
    "ccs_gather_bench": (I, [P, I, U64, I, P, P]),
    "ccs_index_host": (I, [CP, P, I, P]),
}

_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise LibraryMissing(f"{LIB_PATH} not found; build it with `make -C {PKG_ROOT}`")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> int:
    if rc < 0:
        raise CcsError(rc, lib().ccs_last_error().decode(errors="replace"))
    return rc


def device_info() -> DeviceInfo:
    d = DeviceInfo()
    check(lib().ccs_device_info(ctypes.byref(d)))
    return d


def kind_info(kind: str) -> KindInfo:
    k = KindInfo()
    check(lib().ccs_kind_info(kind.encode(), ctypes.byref(k)))
    return k
