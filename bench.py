"""Benchmark: cubes/s on the d4t table, persistent batched form vs plain one-cube-at-a-time.

One step = one batched solve of the workload (4,096 seeded uniform random
states, seed 3, on d4t, one persistent launch).  Before it: the d4t tables
(block and base-3) are built from scratch on the GPU, reported per layer as
`table_build`, and the autotuner picks the persistent form's knobs on a
separate 2,048-state sample (seed 33: large enough that its final rounds are
ordered by item keys as the workload's are); neither is part of the step.
After it: every workload state solved one at a time in the plain form (all
lengths compared with the batched ones), the random-probe ceiling of the
memory system, the hard tail (256 length >= 18 positions plus 64 flip-heavy
states on d4, the ten distance-20 positions and the superflip each alone),
the cold and warm load time of the d4t file, and a Poisson serving run at 50%
and 80% of the measured capacity (seed 5).  The CPU baseline (the oracle's
complete solves of the first 64 workload states on the host cores) runs in a
thread beside the batched steps, whose time is device time.  Prints one JSON
line.

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
FETCHED_BYTES_PER_PROBE = 128.1
WRITTEN_BYTES_PER_PROBE = 0.087


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
    """Random-probe ceiling of the memory system on this table: 2^32 probes of
    random 32-byte blocks, four independent probes per lane, read as the block
    layout's probe reads (the minimum word plus one offset word) and as the
    search reads its half-block copy (one 16-byte load); and a streaming read
    of the whole table.  Timed with HIP events in the library."""
    import ctypes

    from cuda_cube_solver_amd import _lib

    out = {}
    for mode, n in ((0, 1 << 32), (3, 1 << 32), (2, 0)):
        ms, cs = ctypes.c_double(), ctypes.c_uint64()
        _lib.check(_lib.lib().ccs_gather_bench(tb.handle, mode, n, 16, ctypes.byref(ms), ctypes.byref(cs)))
        if mode == 0:
            out["random_probes_g_per_s"] = round(n / ms.value / 1e6, 2)
            out["random_probe_algorithmic_gb_s"] = round(n * BYTES_PER_PROBE / ms.value / 1e6, 1)
        elif mode == 3:
            out["random_probes_one_load_g_per_s"] = round(n / ms.value / 1e6, 2)
        else:
            out["stream_gb_s"] = round(tb.nbytes("block") / ms.value / 1e6, 1)
    return out


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads() -> int:
    """Cores this process may use: the affinity mask, capped by OMP_NUM_THREADS
    when it is set (the share of the host that a one-GPU job gets)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(n, int(cap))) if cap and cap.isdigit() else n


class CpuBaseline:
    """The oracle (C restatement of the reference's step-mode search) solving
    the first n workload states completely on the host cores, one cube per
    thread, over the same GPU-built table.  Runs in a thread (the oracle's C
    call releases the GIL) while the GPU measures device time; two cores stay
    free for the GPU host thread."""

    def __init__(self, tb, cubes, n_threads: int):
        import threading

        import numpy as np

        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle as O

        self.n_threads = n_threads
        self.table = np.frombuffer(tb.read("block", 0, tb.nbytes("block")), dtype=np.uint8)
        self.arr = O.cubes([(s.cp, s.co, s.ep, s.eo) for s in cubes])
        self.kind = O.kind("d4t")
        self.result = None
        self.error = None
        self.thread = threading.Thread(target=self._run, daemon=True)
        self.thread.start()

    def _run(self):
        import oracle as O

        try:
            t = time.perf_counter()
            res, bad = O.solve_many(self.kind, self.table, 0, self.arr, n_threads=self.n_threads)
            wall = time.perf_counter() - t
            self.result = (res, bad, wall)
        except Exception as e:  # noqa: BLE001
            self.error = e

    def finish(self, gpu_lengths) -> dict:
        t = time.perf_counter()
        while self.thread.is_alive():
            self.thread.join(timeout=30.0)
            if self.thread.is_alive():
                log(f"CPU baseline still running ({time.perf_counter() - t:.0f} s waited)")
        if self.error is not None:
            raise self.error
        res, bad, wall = self.result
        secs = [r["seconds"] * 1e3 for r in res]
        nodes = sum(r["nodes"] for r in res)
        n = len(res)
        return {"value": round(n / wall, 4), "unit": "cubes/s", "cores": self.n_threads, "kind": "port",
                "cpu": cpu_model(),
                "sample": f"first {n} workload states (seed 3) on d4t, complete solves, one cube per thread "
                          f"on {self.n_threads} threads: {wall:.1f} s wall, {nodes / 1e9:.2f} G nodes "
                          f"({nodes / wall / 1e6:.1f} M nodes/s)",
                "p50_ms": round(pct(secs, 50), 1), "p99_ms": round(pct(secs, 99), 1),
                "unsolved": bad,
                "lengths_equal_gpu": [r["length"] for r in res] == list(gpu_lengths[:n])}


def hard_tail(tb_d4t, log) -> dict:
    """Hard positions: C4 on d4 (256 of the reference's length >= 18 positions
    of data/random1000.txt drawn with seed 4, plus 64 flip-heavy states, one
    persistent launch), the ten distance-20 positions and the superflip each
    solved alone in the persistent form on d4t and d4.  Lengths are checked
    against the reference's recorded ones."""
    from cuda_cube_solver_amd import states
    from cuda_cube_solver_amd.cubie import MOVES_HTM, CubieState, apply_moves, parse_moves
    from cuda_cube_solver_amd.solver import GpuSolver, Tables

    golden = json.loads((ROOT / "tests" / "golden" / "reference_results.json").read_text())

    def scrambled(texts):
        return [apply_moves(CubieState(), [MOVES_HTM[m] for m in parse_moves(t)]) for t in texts]

    ref = golden["random1000"]
    pool = [i for i, r in enumerate(ref["results"]) if r["len"] >= 18]
    pick = [pool[i] for i in states.sample_indices(len(pool), 256, 4)]
    c4 = scrambled([ref["scrambles"][i] for i in pick]) + states.flip_heavy_states(64, 4)
    d20 = scrambled(golden["distance20_random10"]["scrambles"])
    sf = states.flip_heavy_states(2, 1)[-1]
    out = {}
    td4, _ = Tables.build("d4", "block")
    try:
        for name, tb in (("d4", td4), ("d4t", tb_d4t)):
            solver = GpuSolver(table=tb, mode="batched")
            if name == "d4":
                sols = solver.solve(c4)
                st = solver.last_stats
                lens = [len(s.moves) for s in sols]
                out["c4_d4"] = {"cubes": len(c4), "cubes_per_s": round(len(c4) / (st.kernel_ms / 1e3), 2),
                                "kernel_s": round(st.kernel_ms / 1e3, 2),
                                "lengths_match_reference": lens[:256] == [ref["results"][i]["len"] for i in pick],
                                "superflip_len": lens[-1]}
                log(f"C4 on d4: {out['c4_d4']}")
            times, ok = [], True
            for c in d20:
                t = time.perf_counter()
                ok &= len(solver.solve([c])[0].moves) == 20
                times.append(time.perf_counter() - t)
            t = time.perf_counter()
            sol = solver.solve([sf])[0]
            sf_s = time.perf_counter() - t
            out[f"alone_{name}"] = {"distance20_median_s": round(sorted(times)[len(times) // 2], 3),
                                    "distance20_max_s": round(max(times), 3), "distance20_all_20": ok,
                                    "superflip_s": round(sf_s, 2), "superflip_len": len(sol.moves)}
            log(f"alone on {name}: {out[f'alone_{name}']}")
    finally:
        td4.close()
    out["reference"] = ("RTX 5060 Ti (docs/benchmarks.md:75-90): distance-20 median 1.02 s on d4t, 1.00 s on "
                        "d4; superflip 16.0 s on d4t, 3.2 s on d4")
    return out


def table_load(tb, log) -> dict:
    """ccs_tables_open on the d4t block file written by ccs_tables_save: once
    with the file dropped from the page cache (cold), once warm.  The
    reference quotes 1-5 s per load (README.md:152)."""
    import tempfile

    from cuda_cube_solver_amd.solver import Tables

    d = Path(tempfile.mkdtemp(prefix="ccs_bench_"))
    path = d / "prune-ud-d4t-blk.bin"
    try:
        t = time.perf_counter()
        tb.save("block", path)
        save_s = time.perf_counter() - t
        fd = os.open(path, os.O_RDONLY)
        try:
            os.fsync(fd)
            os.posix_fadvise(fd, 0, 0, os.POSIX_FADV_DONTNEED)
        finally:
            os.close(fd)
        out = {"bytes": path.stat().st_size, "save_s": round(save_s, 2)}
        for name in ("cold", "warm"):
            t = time.perf_counter()
            t2 = Tables.open(path, "d4t", "block")
            sec = time.perf_counter() - t
            t2.close()
            out[f"{name}_s"] = round(sec, 2)
            out[f"{name}_gb_s"] = round(out["bytes"] / sec / 1e9, 2)
        log(f"table load: {out}")
        return out
    finally:
        if path.exists():
            path.unlink()
        d.rmdir()


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cubes", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--plain", type=int, default=-1,
                    help="workload cubes solved one at a time for the plain figure (-1: all of them)")
    ap.add_argument("--table", default="d4t")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-cubes", type=int, default=64, help="workload states the CPU baseline solves")
    ap.add_argument("--no-hard", action="store_true", help="skip the hard-tail measurements")
    ap.add_argument("--no-load", action="store_true", help="skip the table load timing")
    ap.add_argument("--no-gather", action="store_true", help="skip the random-probe ceiling measurement")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner; use --chunk-nodes etc.")
    ap.add_argument("--tune-cubes", type=int, default=2048, help="autotuner sample (seeded uniform states, seed 33)")
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--prefix-len", type=int, default=4)
    ap.add_argument("--block", type=int, default=0, help="lanes per search workgroup (0: library default)")
    ap.add_argument("--probes-per-step", type=int, default=0, help="table probes per lane per step, 1 or 2 (0: 2)")
    ap.add_argument("--read-mode", type=int, default=0,
                    help="a step's two table reads: 1 paired, 2 in turn (0: the library's choice by batch size)")
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
    cpu = None
    if not args.no_cpu:
        threads = max(1, host_threads() - 2)
        log(f"CPU baseline: {args.cpu_cubes} states on {threads} threads ({cpu_model()}), beside the GPU steps")
        cpu = CpuBaseline(tb, cubes[: args.cpu_cubes], threads)
    knobs = {"chunk_nodes": args.chunk_nodes, "blocks_per_cu": args.blocks_per_cu, "prefix_len": args.prefix_len,
             "deeper_prefix_at": (5e7, 3e9), "block": args.block, "probes_per_step": args.probes_per_step,
             "read_mode": args.read_mode}
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

    # the CPU baseline shares the host: it ends before the host-timed plain leg
    cpu_line = None
    if cpu is not None:
        try:
            cpu_line = cpu.finish(lengths)
        except Exception as e:  # noqa: BLE001
            cpu_line = {"error": f"{type(e).__name__}: {e}"}
        log(f"CPU baseline: {cpu_line}")
    # plain: one cube at a time, per-round launches (the reference's form at batch 1)
    plain_ms = []
    plain_len = []
    n_plain = len(cubes) if args.plain < 0 else args.plain
    for s in cubes[:n_plain]:
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
        "lengths_compared": len(plain_len),
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
    if not args.no_hard:
        line["hard_tail"] = extra("hard tail", lambda: hard_tail(tb, log))
    if not args.no_load:
        line["table_load"] = extra("table load", lambda: table_load(tb, log))
    if not args.no_serve:
        from cuda_cube_solver_amd import serve_bench

        hard = [c for c, L in zip(cubes, lengths) if L >= 18]
        line["serving"] = []
        for frac in (0.5, 0.8):
            qps = frac * end_to_end

            def one():
                r = serve_bench.run(tb, qps, max(16, int(qps * args.serve_seconds)), seed=5, hard_pool=hard,
                                    log=log)
                lens = r.pop("lengths")
                r["superflips"] = sum(1 for k, c in serve_bench.request_mix(r["requests"], 5, hard)
                                      if k == "flip" and c.eo == (1,) * 12)
                r["all_solved"] = min(lens) >= 0
                return r

            r = extra(f"serving at {qps:.0f} QPS", one)
            r["load"] = f"{frac:.0%} of the batched capacity"
            line["serving"].append(r)
    if cpu_line is not None:
        line["cpu_baseline"] = cpu_line
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
