"""Optimal lengths and complete-round node counts of seeded workload states,
computed by the CPU oracle (the reference's search rules) over oracle-built
tables, written to seeded_results.json for the GPU parity tests.

    python tests/golden/make_seeded_fixtures.py TABLE_DIR

TABLE_DIR holds prune-ud-<kind>-blk.bin files written by make_table_hashes.py.
"""
from __future__ import annotations

import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1] / "oracle"))
sys.path.insert(0, str(HERE.parents[1] / "cuda-cube-solver_amd"))
import oracle as O  # noqa: E402
from cuda_cube_solver_amd import states  # noqa: E402

# (name, kind, seed, count): C1 is 64 uniform states (seed 1) on 304; C3's
# workload is seed 3 on d4t (the first 64 here); C2's seed 2 on d4t.
SETS = [("c1_304_seed1", "304", 1, 64), ("c3_d4t_seed3_first64", "d4t", 3, 64),
        ("c2_d4t_seed2_first64", "d4t", 2, 64)]


def main() -> None:
    table_dir = Path(sys.argv[1])
    only = set(sys.argv[2:])
    out_path = HERE / "seeded_results.json"
    out = json.loads(out_path.read_text()) if out_path.exists() else {}
    for name, kind, seed, count in SETS:
        if only and name not in only:
            continue
        table = np.memmap(table_dir / f"prune-ud-{kind}-blk.bin", dtype=np.uint8, mode="r")
        sts = states.uniform_states(count, seed)
        arr = O.cubes([(s.cp, s.co, s.ep, s.eo) for s in sts])
        t = time.time()
        res, bad = O.solve_many(O.kind(kind), np.asarray(table), 0, arr)
        if bad:
            raise SystemExit(f"{name}: {bad} cubes unsolved")
        out[name] = {"kind": kind, "seed": seed, "count": count,
                     "lengths": [r["length"] for r in res],
                     "complete_round_nodes": [O.complete_round_nodes(r) for r in res],
                     "oracle_seconds": round(time.time() - t, 1)}
        print(name, out[name]["lengths"], out[name]["oracle_seconds"], flush=True)
        out_path.write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
