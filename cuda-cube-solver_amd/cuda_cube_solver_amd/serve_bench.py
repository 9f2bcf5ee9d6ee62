"""Serving benchmark: Poisson arrivals of mixed cube requests on a resident table.

Request mix (seeded): 80% uniform random states, 15% from a pool of hard
states (optimal length >= 18), 5% flip-heavy states.  Requests are submitted
at their arrival times while earlier ones are still on the device; the
server admits them as search slots free up.  Reports p50/p99 latency
(submission -> result) and the achieved cubes/s.
"""
from __future__ import annotations

import math
import threading
import time

from . import states
from .server import Server
from .solver import Tables


def request_mix(n: int, seed: int, hard_pool: list) -> list:
    """n (kind, cube) requests: 80% uniform, 15% from hard_pool, 5% flip-heavy
    (a seeded list whose last state is the superflip, all 12 edges flipped)."""
    rng = states.SplitMix64(seed)
    uni = states.uniform_states(n, seed + 1000)
    flips = states.flip_heavy_states(max(n // 20 + 2, 2), seed + 2000)
    out = []
    for i in range(n):
        r = rng.below(100)
        if r < 80 or not hard_pool:
            out.append(("uniform", uni[i]))
        elif r < 95:
            out.append(("hard", hard_pool[rng.below(len(hard_pool))]))
        else:
            out.append(("flip", flips[rng.below(len(flips))]))
    return out


def poisson_times(n: int, qps: float, seed: int) -> list:
    rng = states.SplitMix64(seed)
    t, out = 0.0, []
    for _ in range(n):
        u = (rng.next() >> 11) / float(1 << 53)
        t += -math.log(1.0 - u) / qps
        out.append(t)
    return out


def pct(xs, q):
    xs = sorted(xs)
    k = max(0, min(len(xs) - 1, int(math.ceil(q / 100.0 * len(xs))) - 1))  # nearest rank
    return xs[k]


def run(tables: Tables, qps: float, n: int, seed: int = 5, hard_pool: list | None = None,
        state_budget_bytes: int = 0, log=None, **server_kw) -> dict:
    """Offer n requests at `qps` (Poisson arrivals); `server_kw` goes to Server."""
    reqs = request_mix(n, seed, hard_pool or [])
    when = poisson_times(n, qps, seed + 1)
    results = {}
    with Server(tables, state_budget_bytes=state_budget_bytes, **server_kw) as srv:
        done = threading.Event()

        quit = threading.Event()

        def collector():
            while len(results) < n and not quit.is_set():
                for r in srv.poll(256, 200):
                    results[r.tag] = r
            done.set()

        th = threading.Thread(target=collector, daemon=True)
        th.start()
        t0 = time.perf_counter()
        for i, ((kind, cube), t_arr) in enumerate(zip(reqs, when)):
            delay = t_arr - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            srv.submit(cube, tag=i)
        window = time.perf_counter() - t0
        ok = done.wait(timeout=max(120.0, 4 * n / qps))
        quit.set()
        th.join(timeout=5)  # the collector must not poll a stopped server
        if not ok:
            raise TimeoutError(f"{len(results)}/{n} served")
        wall = time.perf_counter() - t0
    lat = [results[i].latency_ms for i in range(n)]
    by_kind = {}
    for i, (kind, _) in enumerate(reqs):
        by_kind.setdefault(kind, []).append(results[i].latency_ms)
    # completions while requests were still arriving, per second of that window
    in_window = sum(1 for i in range(n) if when[i] * 1e3 + results[i].latency_ms <= window * 1e3)
    out = {"qps_offered": round(qps, 2), "requests": n, "arrival_window_s": round(window, 2),
           "served_in_window_per_s": round(in_window / window, 2), "served_overall_per_s": round(n / wall, 2),
           "p50_ms": round(pct(lat, 50), 2), "p99_ms": round(pct(lat, 99), 2),
           "max_ms": round(max(lat), 2),
           "p50_queued_ms": round(pct([results[i].queued_ms for i in range(n)], 50), 2),
           "p50_ms_by_kind": {k: round(pct(v, 50), 2) for k, v in by_kind.items()},
           "mix": {k: len(v) for k, v in by_kind.items()},
           "lengths": [results[i].length for i in range(n)]}
    if log:
        log(f"serving at {qps:.1f} QPS: {out['served_in_window_per_s']} cubes/s in the window, "
            f"p50 {out['p50_ms']} ms, p99 {out['p99_ms']} ms, max {out['max_ms']} ms")
    return out
