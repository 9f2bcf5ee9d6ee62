"""MI355X-native optimal Rubik's cube solver (half-turn metric).

A drop-in for the reference's solving path (ozimchuk/cuda-cube-solver): same
table names, on-disk table layout, CLI and Python API, with the BFS and the
IDA* search in hand-written HIP kernels for gfx950 behind a C ABI
(libcubesolver.so, include/cubesolver.h) loaded through ctypes.
"""
__version__ = "0.1.0"
