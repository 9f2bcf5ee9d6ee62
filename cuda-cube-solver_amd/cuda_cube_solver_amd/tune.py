"""Autotuner for the persistent (batched) search form.

Solves a seeded sample of cubes with each setting of the knobs and keeps the
fastest: nodes per fetched chunk (node budget), prefix length, the per-round
prefix policy (the node counts at which a round's items get one and two moves
longer), then, at the best of those, waves per CU (lanes per workgroup; the
LUTs and the lanes' stacks fill the LDS, so one workgroup runs per CU) and
table probes in flight per lane per step, and whether a step's two table reads
are issued together or in turn.  Every setting gives the same
optimal lengths; the tuner checks that too.

    python -m cuda_cube_solver_amd.tune --table d4t --cubes 2048 --seed 33 [--out knobs.json]
"""
from __future__ import annotations

import argparse
import itertools
import json
import sys
import time

from . import states
from .solver import GpuSolver, Tables

CHUNK_NODES = (65536, 262144, 1048576)
BLOCKS_PER_CU = (0,)         # 0: as many as the occupancy calculator allows (one workgroup per CU)
PREFIX_LEN = (4, 5)
DEEPER_AT = ((5e7, 3e9), (1e7, 1e9))
BLOCK = (768, 512)           # lanes per workgroup: 12 or 8 waves per CU
PROBES_PER_STEP = (2, 1)
READ_MODE = (1, 2)           # a step's two table reads paired or in turn


def tune(tables: Tables, cubes, log=None) -> dict:
    """Best knobs for solving `cubes` on `tables` in the batched form."""
    trials, lengths = [], None

    def run(chunk, bpc, plen, deeper, block=BLOCK[0], probes=PROBES_PER_STEP[0], reads=READ_MODE[0]):
        nonlocal lengths
        solver = GpuSolver(table=tables, mode="batched", prefix_len=plen, deeper_prefix_at=deeper,
                           blocks_per_cu=bpc, chunk_nodes=chunk, block=block, probes_per_step=probes,
                           read_mode=reads)
        t = time.perf_counter()
        sols = solver.solve(cubes)
        dt = time.perf_counter() - t
        got = [len(s.moves) for s in sols]
        if lengths is None:
            lengths = got
        elif got != lengths:
            raise AssertionError(f"knobs {chunk, bpc, plen, deeper, block, probes, reads} changed the optimal lengths")
        rec = {"chunk_nodes": chunk, "blocks_per_cu": bpc, "prefix_len": plen, "deeper_prefix_at": list(deeper),
               "block": block, "probes_per_step": probes, "read_mode": reads, "cubes_per_s": round(len(cubes) / dt, 2)}
        trials.append(rec)
        if log:
            log(f"tune: {rec}")
        return rec

    for chunk, bpc, plen in itertools.product(CHUNK_NODES, BLOCKS_PER_CU, PREFIX_LEN):
        run(chunk, bpc, plen, DEEPER_AT[0])
    best = max(trials, key=lambda r: r["cubes_per_s"])
    for deeper in DEEPER_AT[1:]:
        run(best["chunk_nodes"], best["blocks_per_cu"], best["prefix_len"], deeper)
    best = max(trials, key=lambda r: r["cubes_per_s"])
    keep = (best["chunk_nodes"], best["blocks_per_cu"], best["prefix_len"], tuple(best["deeper_prefix_at"]))
    for block in BLOCK[1:]:
        run(*keep, block=block)
    for probes in PROBES_PER_STEP[1:]:
        run(*keep, probes=probes)
    for reads in READ_MODE[1:]:
        run(*keep, reads=reads)
    best = max(trials, key=lambda r: r["cubes_per_s"])
    return {"best": best, "trials": trials, "sample": len(cubes)}


def knobs(best: dict) -> dict:
    """GpuSolver keyword arguments of a tuner result."""
    return {"chunk_nodes": best["chunk_nodes"], "blocks_per_cu": best["blocks_per_cu"],
            "prefix_len": best["prefix_len"], "deeper_prefix_at": tuple(best["deeper_prefix_at"]),
            "block": best.get("block", 0), "probes_per_step": best.get("probes_per_step", 0),
            "read_mode": best.get("read_mode", 0)}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="cuda_cube_solver_amd.tune")
    ap.add_argument("--table", default="d4t")
    ap.add_argument("--cubes", type=int, default=2048)
    ap.add_argument("--seed", type=int, default=33)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    tb, _ = Tables.build(args.table, "block")
    try:
        res = tune(tb, states.uniform_states(args.cubes, args.seed),
                   log=lambda m: print(m, file=sys.stderr, flush=True))
    finally:
        tb.close()
    text = json.dumps(res, indent=1)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text + "\n")
    print(json.dumps(res["best"]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
