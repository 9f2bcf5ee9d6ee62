"""Build solving tables with the CPU oracle's BFS and record their SHA-256.

    python tests/golden/make_table_hashes.py 304 [d4t ...] [--keep DIR]

For each kind: the oracle BFS (a restatement of the reference's mod-3 + block
BFS) builds the tables on the CPU; the script checks its depth counts against
``reference_results.json`` (the reference's research/subgroups/results.jsonl)
and the histogram self-check, then writes the SHA-256 of the block file and of
the b3 file into ``table_sha256.json``.  ``--keep DIR`` also writes the table
files (``prune-ud-<kind>-blk.bin`` / ``-b3.bin``) there.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1] / "oracle"))
import oracle as O  # noqa: E402


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("kinds", nargs="+")
    ap.add_argument("--keep", type=Path)
    ap.add_argument("--threads", type=int, default=0)
    args = ap.parse_args()
    ref = json.loads((HERE / "reference_results.json").read_text())["depth_counts"]
    out_path = HERE / "table_sha256.json"
    hashes = json.loads(out_path.read_text()) if out_path.exists() else {}
    for name in args.kinds:
        k = O.kind(name)
        t = time.time()
        mod3, blk, counts = O.bfs(k, n_threads=args.threads, verbose=True)
        secs = time.time() - t
        if name in ref and counts != ref[name]["depth_counts"]:
            raise SystemExit(f"{name}: depth counts {counts} != reference {ref[name]['depth_counts']}")
        hist = O.histogram(mod3, k.n_entries)
        want = [sum(counts[d] for d in range(r, len(counts), 3)) for r in range(3)] + [0]
        if hist != want:
            raise SystemExit(f"{name}: histogram {hist} != {want}")
        b3 = O.build_b3(mod3, blk, k.n_entries)
        entry = {"entries": int(k.n_entries), "depth_counts": counts,
                 "blk_bytes": int(blk.nbytes), "blk_sha256": hashlib.sha256(blk.tobytes()).hexdigest(),
                 "b3_bytes": int(b3.nbytes), "b3_sha256": hashlib.sha256(b3.tobytes()).hexdigest(),
                 "oracle_bfs_seconds": round(secs, 1)}
        hashes[name] = entry
        print(name, json.dumps(entry), flush=True)
        if args.keep:
            args.keep.mkdir(parents=True, exist_ok=True)
            blk.tofile(args.keep / f"prune-ud-{name}-blk.bin")
            b3.tofile(args.keep / f"prune-ud-{name}-b3.bin")
        out_path.write_text(json.dumps(hashes, indent=1) + "\n")
        del mod3, blk, b3


if __name__ == "__main__":
    main()
