"""GPU test of the search diagnostics build (libcubesolver_diag.so, the
kernels compiled with -DCCS_SEARCH_STATS): the counters add up and the
solutions are the same as the product library's."""
import json
import os
import subprocess
import sys
from pathlib import Path

import pytest

from cuda_cube_solver_amd import states
from cuda_cube_solver_amd.solver import GpuSolver, Tables

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]
DIAG_LIB = ROOT / "cuda-cube-solver_amd" / "libcubesolver_diag.so"

SCRIPT = r"""
import json, sys
sys.path.insert(0, sys.argv[1])
from cuda_cube_solver_amd import states
from cuda_cube_solver_amd.solver import GpuSolver, Tables
tb, _ = Tables.build("304", "block")
solver = GpuSolver(table=tb, mode="batched")
sols = solver.solve(states.uniform_states(48, 101))
st = solver.last_stats
print(json.dumps({"lengths": [len(s.moves) for s in sols], "probes": st.probes, "nodes": st.nodes,
                  "built": st.diag_built, "diag": list(st.diag)}))
"""


def test_diagnostics_build_counts_add_up():
    assert DIAG_LIB.exists(), "libcubesolver_diag.so not built"
    env = dict(os.environ, CCS_LIBRARY=str(DIAG_LIB))
    out = subprocess.run([sys.executable, "-c", SCRIPT, str(ROOT / "cuda-cube-solver_amd")], env=env,
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    r = json.loads(out.stdout.strip().splitlines()[-1])
    d = r["diag"]
    assert r["built"] == 1
    assert sum(d[0:6]) == r["probes"] > 0                     # every probe booked to its frame
    assert all(d[6 + i] <= d[i] for i in range(6))            # a probe prunes at most once
    assert sum(d[14:23]) == d[12] > 0                         # the histogram covers every step
    assert all(c > 0 for c in d[23:27])
    assert 0 < d[27] <= d[23] and 0 < d[28] <= d[12]        # passes with an item start
    assert d[29] >= d[12]                                    # at least one pass iteration per step
    assert 0 < d[13] < r["nodes"]                             # frame updates: pushes and pops
    # the product library, same cubes: same lengths, no diagnostics
    tb, _ = Tables.build("304", "block")
    try:
        solver = GpuSolver(table=tb, mode="batched")
        sols = solver.solve(states.uniform_states(48, 101))
        assert [len(s.moves) for s in sols] == r["lengths"]
        assert solver.last_stats.diag_built == 0 and not any(solver.last_stats.diag)
    finally:
        tb.close()
