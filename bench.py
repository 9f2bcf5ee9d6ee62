"""Benchmark: cubes/s on the d4t table, persistent batched form vs plain one-cube-at-a-time.

One step = one batched solve of the workload (4,096 seeded uniform random
states, seed 3, on d4t, one persistent launch).  Before it: the d4t tables
(block and base-3) are built from scratch on the GPU, reported per layer as
`table_build`, and the autotuner picks the persistent form's knobs on a
separate 2,048-state sample (seed 33: large enough that its final rounds are
ordered by item keys as the workload's are); neither is part of the step.
After it:
the first 128 states solved one at a time in the plain form, the random-probe
ceiling of the memory system, a Poisson serving run at 50% and 80% of the
measured capacity (seed 5), and the CPU oracle on a bounded sample.  Prints
one JSON line.

    python bench.py [--gpus 1] [--steps 1] [--warmup 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "cuda-cube-solver_amd"))

HBM_PEAK_GBS = 8000.0
BYTES_PER_PROBE = 32  # algorithmic: one 32-byte table block per lookup
# HBM bytes moved per probe by the search kernel, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE (KiB per dispatch) scaled by kernels of known size in the same run
# (FETCH_SIZE x2: it books each 128-byte read request as 64 B; WRITE_SIZE x1):
# profiles/search_kernel_d4t_pmc.md.  A random 32-byte block costs one 128-byte line.
FETCHED_BYTES_PER_PROBE = 128.0
WRITTEN_BYTES_PER_PROBE = 0.059


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def pct(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = max(0, min(len(xs) - 1, int(round(q / 100.0 * len(xs) + 0.5)) - 1))  # nearest rank
    return xs[k]


def build_summary(lines: list[str], counts: list[int], total_s: float) -> dict:
    """The table build from the library's log: one BFS layer per depth
    ("depth d fwd|bwd found n (t s)"), then the block and base-3 encoders
    ("encoded <fmt> table, <bytes> bytes (t s)").  A top-down layer expands the
    previous depth's entries, 18 moves each; a bottom-up layer tests every
    still-unvisited entry until one of its 18 neighbours sits at the previous
    depth."""
    import re

    layers, enc = [], {}
    remaining = sum(counts) - counts[0]
    for s in lines:
        m = re.match(r"depth\s+(\d+) (fwd|bwd) found\s+(\d+) \(\s*([\d.]+)s\)", s)
        if m:
            d, how, found, sec = int(m[1]), m[2], int(m[3]), float(m[4])
            row = {"depth": d, "direction": "top-down" if how == "fwd" else "bottom-up", "found": found,
                   "seconds": sec}
            if how == "fwd":
                row["expanded"] = counts[d - 1]
                row["neighbours_per_s"] = round(18 * counts[d - 1] / sec) if sec > 0 else None
            else:
                row["tested"] = remaining
            layers.append(row)
            remaining -= found
        m = re.match(r"encoded (\w+) table, (\d+) bytes \(\s*([\d.]+)s\)", s)
        if m:
            enc[m[1]] = {"bytes": int(m[2]), "seconds": float(m[3]),
                         "written_gb_s": round(int(m[2]) / float(m[3]) / 1e9, 1) if float(m[3]) > 0 else None}
    bfs_s = sum(r["seconds"] for r in layers)
    return {"entries": sum(counts), "depth_counts": counts, "seconds": round(total_s, 2),
            "bfs_seconds": round(bfs_s, 2), "entries_per_s": round(sum(counts) / bfs_s) if bfs_s else None,
            "layers": layers, "encode": enc}


def gather_ceiling(tb) -> dict:
    """Random-probe ceiling of the memory system on this table: 2^32 reads of a
    random 32-byte block's minimum word plus one offset word (the search's
    access pattern, four independent probes per lane), and a streaming read of
    the whole table, timed with HIP events in the library."""
    import ctypes

    from cuda_cube_solver_amd import _lib

    out = {}
    for mode, n in ((0, 1 << 32), (2, 0)):
        ms, cs = ctypes.c_double(), ctypes.c_uint64()
        _lib.check(_lib.lib().ccs_gather_bench(tb.handle, mode, n, 16, ctypes.byref(ms), ctypes.byref(cs)))
        if mode == 0:
            out["random_probes_g_per_s"] = round(n / ms.value / 1e6, 2)
            out["random_probe_algorithmic_gb_s"] = round(n * BYTES_PER_PROBE / ms.value / 1e6, 1)
        else:
            out["stream_gb_s"] = round(tb.nbytes("block") / ms.value / 1e6, 1)
    return out


def cpu_baseline(tb, cubes, nodes_per_cube: float, seconds_budget: float):
    """The oracle (C restatement of the reference's search) on the host cores,
    over the same GPU-built table: one workload cube per thread for a fixed
    wall-clock budget.  Its node rate divided by the workload's nodes per
    cube (the same node definition on both sides) gives cubes/s."""
    import numpy as np

    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    table = np.frombuffer(tb.read("block", 0, tb.nbytes("block")), dtype=np.uint8)
    k = O.kind("d4t")
    threads = min(os.cpu_count() or 1, 16)
    arr = O.cubes([(s.cp, s.co, s.ep, s.eo) for s in cubes[:threads]])
    nodes, dt, _ = O.timed_nodes(k, table, 0, arr, seconds_budget, threads)
    rate = nodes / dt
    return {"value": round(rate / nodes_per_cube, 4), "unit": "cubes/s", "cores": threads, "kind": "port",
            "sample": f"first {threads} workload states, one per thread, {dt:.1f} s wall: {nodes / 1e6:.0f} M nodes "
                      f"= {rate / 1e6:.2f} M nodes/s, divided by the workload's {nodes_per_cube / 1e6:.1f} M nodes/cube"}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cubes", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--plain", type=int, default=128, help="cubes solved one at a time for the plain figure")
    ap.add_argument("--table", default="d4t")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-gather", action="store_true", help="skip the random-probe ceiling measurement")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner; use --chunk-nodes etc.")
    ap.add_argument("--tune-cubes", type=int, default=2048, help="autotuner sample (seeded uniform states, seed 33)")
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--prefix-len", type=int, default=4)
    ap.add_argument("--chunk-nodes", type=int, default=0, help="persistent form: nodes per fetched chunk (0: library default)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--serve-seconds", type=float, default=15.0, help="arrival window per offered load")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: cubes are independent, so "
              f"several GPUs would each hold a copy of the table and take their share of the cubes",
              file=sys.stderr)
        sys.exit(2)

    from cuda_cube_solver_amd import _lib, states
    from cuda_cube_solver_amd.solver import GpuSolver, Tables

    dev = _lib.device_info()
    log(f"device {dev.name.decode()} {dev.arch.decode()} {dev.compute_units} CUs")
    t = time.perf_counter()
    build_log: list[str] = []
    # block and base-3 tables from scratch; searches use the block table
    tb, counts = Tables.build(args.table, "both", log=lambda s: (build_log.append(s), log(s)))
    build_s = time.perf_counter() - t
    log(f"built {args.table} in {build_s:.1f}s, depth counts {counts}")
    cubes = states.uniform_states(args.cubes, args.seed)
    knobs = {"chunk_nodes": args.chunk_nodes, "blocks_per_cu": args.blocks_per_cu, "prefix_len": args.prefix_len,
             "deeper_prefix_at": (5e7, 3e9)}
    tuned = None
    if not args.no_tune:
        from cuda_cube_solver_amd import tune

        t = time.perf_counter()
        tuned = tune.tune(tb, states.uniform_states(args.tune_cubes, 33), log=log)
        knobs = tune.knobs(tuned["best"])
        log(f"autotuner: {len(tuned['trials'])} settings in {time.perf_counter() - t:.1f}s, best {tuned['best']}")
    solver = GpuSolver(table=tb, mode="batched", **knobs)

    for w in range(args.warmup):
        solver.solve(cubes[: min(512, len(cubes))], mode="batched")
        log(f"warmup {w + 1}/{args.warmup} done")
    step_s, probes, kernel_ms, lat, nodes_total = [], 0, 0.0, [], 0
    lengths = None
    for k in range(args.steps):
        t = time.perf_counter()
        sols = solver.solve(cubes, mode="batched")
        step_s.append(time.perf_counter() - t)
        st = solver.last_stats
        probes += st.probes
        nodes_total += st.nodes
        kernel_ms += st.kernel_ms
        lat = [s.seconds * 1e3 for s in sols]
        lengths = [len(s.moves) for s in sols]
        log(f"step {k + 1}/{args.steps}: {step_s[-1]:.2f}s, {st.nodes / 1e9:.2f} G nodes, "
            f"{st.probes / 1e9:.2f} G probes")
    total_s = sum(step_s)
    # value: the device time of the search launches (HIP events on the library's
    # stream), inputs already in HBM; the end-to-end rate adds host root
    # preparation and the 40-byte-per-cube copies over PCIe
    value = args.cubes * args.steps / (kernel_ms / 1e3)
    end_to_end = args.cubes * args.steps / total_s

    # plain: one cube at a time, per-round launches (the reference's form at batch 1)
    plain_ms = []
    plain_len = []
    for s in cubes[: args.plain]:
        t = time.perf_counter()
        sol = solver.solve([s], mode="plain")[0]
        plain_ms.append((time.perf_counter() - t) * 1e3)
        plain_len.append(len(sol.moves))
    plain_value = len(plain_ms) / (sum(plain_ms) / 1e3) if plain_ms else None
    same = plain_len == lengths[: len(plain_len)]
    if plain_ms:
        log(f"plain {len(plain_ms)} cubes: {plain_value:.2f} cubes/s, lengths identical to batched: {same}")

    achieved = probes * BYTES_PER_PROBE / (kernel_ms / 1e3) / 1e9
    line = {
        "metric": "cubes/s (d4t, seeded random states) batched vs plain on 1 GPU; p50/p99 solve ms",
        "value": round(value, 3),
        "unit": "cubes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(kernel_ms / args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8/u32 integer",
        "data": "synthetic: seeded uniform random cube states (splitmix64), d4t table built by the GPU BFS",
        "config": {"workload": f"{args.cubes} seeded uniform random states (seed {args.seed}) on d4t, one "
                               f"persistent batched launch per step", "table": args.table,
                   "knobs": {k: (list(v) if isinstance(v, tuple) else v) for k, v in knobs.items()},
                   "knobs_from": (f"autotuner, {len(tuned['trials'])} settings on {tuned['sample']} seed-33 states"
                                  if tuned else "command line")},
        "batched_p50_ms": round(pct(lat, 50), 2),
        "batched_p99_ms": round(pct(lat, 99), 2),
        "plain_cubes_per_s": round(plain_value, 3) if plain_ms else None,
        "plain_p50_ms": round(pct(plain_ms, 50), 2) if plain_ms else None,
        "plain_p99_ms": round(pct(plain_ms, 99), 2) if plain_ms else None,
        "plain_cubes": len(plain_ms),
        "end_to_end_cubes_per_s": round(end_to_end, 3),
        "end_to_end_ms_per_step": round(total_s / args.steps * 1e3, 1),
        "batched_over_plain": round(end_to_end / plain_value, 2) if plain_ms else None,
        "lengths_identical": same,
        "length_histogram": {str(L): lengths.count(L) for L in sorted(set(lengths))},
        "probes_per_cube": round(probes / (args.cubes * args.steps)),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": round(probes / args.steps * (FETCHED_BYTES_PER_PROBE + WRITTEN_BYTES_PER_PROBE))},
    }
    line["table_build"] = build_summary(build_log, counts, build_s)
    # The measurements after the step are reported beside the metric; one that
    # fails is reported as an error and does not cost the line.
    def extra(name, fn):
        try:
            return fn()
        except Exception as e:  # noqa: BLE001
            log(f"{name} failed: {type(e).__name__}: {e}")
            return {"error": f"{type(e).__name__}: {e}"}

    if not args.no_gather:
        line["probe_ceiling"] = extra("probe ceiling", lambda: gather_ceiling(tb))
    if not args.no_serve:
        from cuda_cube_solver_amd import serve_bench

        hard = [c for c, L in zip(cubes, lengths) if L >= 18]
        line["serving"] = []
        for frac in (0.5, 0.8):
            qps = frac * end_to_end

            def one():
                r = serve_bench.run(tb, qps, max(16, int(qps * args.serve_seconds)), seed=5, hard_pool=hard,
                                    log=log)
                r.pop("lengths")
                return r

            r = extra(f"serving at {qps:.0f} QPS", one)
            r["load"] = f"{frac:.0%} of the batched capacity"
            line["serving"].append(r)
    if not args.no_cpu:
        line["cpu_baseline"] = extra("CPU baseline", lambda: cpu_baseline(
            tb, cubes, nodes_total / (args.cubes * args.steps), 20.0))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
