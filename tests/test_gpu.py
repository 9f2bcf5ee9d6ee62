"""GPU parity tests: the HIP kernels through the C ABI against the CPU oracle
and the reference's recorded results.  Run on an MI355X: pytest -m gpu."""
import ctypes
import hashlib

import numpy as np
import pytest

import oracle as O
from cuda_cube_solver_amd import _lib, cubie, states
from cuda_cube_solver_amd.cubie import MOVES_HTM, CubieState, apply_moves, parse_moves
from cuda_cube_solver_amd.solver import GpuSolver, Tables

pytestmark = pytest.mark.gpu


def _scrambled(texts):
    return [apply_moves(CubieState(), [MOVES_HTM[m] for m in parse_moves(t)]) for t in texts]


def _arr(sts):
    return O.cubes([(s.cp, s.co, s.ep, s.eo) for s in sts])


def _sha(tb, fmt):
    h = hashlib.sha256()
    n = tb.nbytes(fmt)
    step = 1 << 28
    for off in range(0, n, step):
        h.update(tb.read(fmt, off, min(step, n - off)))
    return h.hexdigest()


def _complete_nodes(sol):
    return sum(n for b, n in sol.iteration_nodes if b < len(sol.moves))


@pytest.fixture(scope="module")
def t304(reference_results):
    tb, counts = Tables.build("304", "both")
    assert counts == reference_results["depth_counts"]["304"]["depth_counts"]
    yield tb
    tb.close()


@pytest.fixture(scope="module")
def td4t(reference_results):
    tb, counts = Tables.build("d4t", "both")
    assert counts == reference_results["depth_counts"]["d4t"]["depth_counts"]
    yield tb
    tb.close()


@pytest.fixture(scope="module")
def blk304(t304):
    return np.frombuffer(t304.read("block", 0, t304.nbytes("block")), dtype=np.uint8)


def test_304_tables_are_bit_exact(t304, table_hashes):
    want = table_hashes["304"]
    assert t304.nbytes("block") == want["blk_bytes"]
    assert _sha(t304, "block") == want["blk_sha256"]
    assert t304.nbytes("b3") == want["b3_bytes"]
    assert _sha(t304, "b3") == want["b3_sha256"]


def test_d4t_tables_are_bit_exact(td4t, table_hashes):
    want = table_hashes["d4t"]
    assert td4t.nbytes("block") == want["blk_bytes"] == 7793525280
    assert _sha(td4t, "block") == want["blk_sha256"]
    assert td4t.nbytes("b3") == want["b3_bytes"] == 6234820224
    assert _sha(td4t, "b3") == want["b3_sha256"]


def _probe(tb, sts, move):
    n = len(sts)
    idx = np.zeros(n, np.uint64)
    lb = np.zeros(n, np.uint32)
    out = ctypes.create_string_buffer(40 * n)
    buf = ctypes.create_string_buffer(b"".join(s.to_bytes() for s in sts))
    _lib.check(_lib.lib().ccs_probe(tb.handle, buf, n, move, idx.ctypes.data, lb.ctypes.data, out))
    return idx, lb, [CubieState.from_bytes(out.raw[40 * i:40 * i + 40]) for i in range(n)]


@pytest.fixture(scope="module")
def b3_304(t304):
    return np.frombuffer(t304.read("b3", 0, t304.nbytes("b3")), dtype=np.uint8)


@pytest.mark.parametrize("fmt", ["block", "b3"])
@pytest.mark.parametrize("move", [-1] + list(range(18)) + [-2 - m for m in range(18)])
def test_probe_moves_index_and_bound_match_oracle(t304, blk304, b3_304, move, fmt):
    sts = states.uniform_states(300, 21) + states.flip_heavy_states(20, 22)
    t304.use(fmt)
    try:
        idx, lb, moved = _probe(t304, sts, move)
    finally:
        t304.use("block")
    for s, got in zip(sts, moved):
        if move >= 0:
            want = s.then(MOVES_HTM[move])
        elif move <= -2:
            want = MOVES_HTM[-2 - move].then(s)
        else:
            want = s
        assert got == want
    k = O.kind("304")
    want_idx = O.table_index(k, _arr(moved))
    assert (idx == want_idx).all()
    table, code = (blk304, 0) if fmt == "block" else (b3_304, 1)
    assert [int(v) for v in lb] == [O.lookup(table, code, int(i)) for i in want_idx]


def test_unrank_inverts_the_index_on_device(t304, td4t):
    for tb, name in ((t304, "304"), (td4t, "d4t")):
        n = _lib.kind_info(name).n_entries
        rng = np.random.default_rng(7)
        idx = np.concatenate([[0, n - 1], rng.integers(0, n, 5000)]).astype(np.uint64)
        out = np.zeros(len(idx), np.uint64)
        _lib.check(_lib.lib().ccs_unrank_check(tb.handle, idx.ctypes.data, len(idx), out.ctypes.data))
        assert (out == idx).all()


def test_reference_lengths_and_complete_round_nodes_d4t(td4t, reference_results):
    """The reference's first 100 random positions on d4t: optimal lengths as
    recorded, and nodes in completed rounds summing to the reference's own
    figure (research/subgroups/hard.jsonl, g1-27 random_nodes * 100)."""
    ref = reference_results["random1000"]
    cubes = _scrambled(ref["scrambles"][:100])
    solver = GpuSolver(table=td4t, mode="plain")
    sols = solver.solve(cubes)
    assert [len(s.moves) for s in sols] == [r["len"] for r in ref["results"][:100]]
    total = sum(_complete_nodes(s) for s in sols)
    assert total == round(reference_results["complete_round_nodes"]["d4t"]["random_nodes"] * 100)
    assert reference_results["complete_round_nodes"]["d4t"]["random_len"] == [len(s.moves) for s in sols]


def test_batched_matches_plain(td4t, reference_results):
    cubes = _scrambled(reference_results["random1000"]["scrambles"][100:164])
    solver = GpuSolver(table=td4t)
    plain = solver.solve(cubes, mode="plain")
    batched = solver.solve(cubes, mode="batched")
    assert [len(s.moves) for s in batched] == [len(s.moves) for s in plain]
    # same rounds, same items, same node counts in every completed round
    assert [_complete_nodes(s) for s in batched] == [_complete_nodes(s) for s in plain]


@pytest.mark.parametrize("mode", ["plain", "batched"])
def test_read_modes_give_the_reference_nodes(td4t, reference_results, mode):
    """A step's two table reads paired (1) or in turn (2), and one probe per
    step: the reference's lengths and completed-round node total on its first
    100 random positions (reads change the timing, never the values)."""
    ref = reference_results["random1000"]
    cubes = _scrambled(ref["scrambles"][:100])
    want = round(reference_results["complete_round_nodes"]["d4t"]["random_nodes"] * 100)
    for reads, probes in ((1, 2), (2, 2), (1, 1)):
        sols = GpuSolver(table=td4t, mode=mode, read_mode=reads, probes_per_step=probes).solve(cubes)
        assert [len(s.moves) for s in sols] == [r["len"] for r in ref["results"][:100]], (reads, probes)
        assert sum(_complete_nodes(s) for s in sols) == want, (reads, probes)


def test_b3_table_gives_the_same_lengths(td4t, reference_results):
    ref = reference_results["random1000"]
    cubes = _scrambled(ref["scrambles"][200:232])
    td4t.use("b3")
    try:
        sols = GpuSolver(table=td4t, mode="batched").solve(cubes)
    finally:
        td4t.use("block")
    assert [len(s.moves) for s in sols] == [r["len"] for r in ref["results"][200:232]]


def test_distance_20_positions(td4t, reference_results):
    d20 = reference_results["distance20_random10"]
    cubes = _scrambled(d20["scrambles"])
    sols = GpuSolver(table=td4t, mode="batched").solve(cubes)
    assert [len(s.moves) for s in sols] == [r["len"] for r in d20["results"]] == [20] * 10


def test_edge_cases(t304):
    solver = GpuSolver(table=t304, mode="batched")
    one = [CubieState().then(MOVES_HTM[m]) for m in range(18)]
    cubes = [CubieState()] + one + [apply_moves(CubieState(), [MOVES_HTM[m] for m in (0, 3)])]
    for mode in ("plain", "batched"):
        sols = solver.solve(cubes, mode=mode)
        assert [len(s.moves) for s in sols] == [0] + [1] * 18 + [2]
        for s, c in zip(sols, cubes):
            assert apply_moves(c, [MOVES_HTM[m] for m in s.moves]) == CubieState()
    assert solver.solve([]) == []
    bad = CubieState(cp=(1, 0) + tuple(range(2, 8)))
    with pytest.raises(_lib.CcsError, match="not a legal cube"):
        solver.solve([bad])


def test_flip_heavy_states(td4t):
    """Flip-heavy states (corners and edge permutation solved, an even subset of
    edges flipped in place): both search forms agree and every solution solves."""
    sts = states.flip_heavy_states(4, 31)[:-1]
    plain = GpuSolver(table=td4t, mode="plain").solve(sts)
    batched = GpuSolver(table=td4t, mode="batched").solve(sts)
    assert [len(s.moves) for s in plain] == [len(s.moves) for s in batched]
    for s, c in zip(batched, sts):
        assert apply_moves(c, [MOVES_HTM[m] for m in s.moves]) == CubieState()


def test_serving_results_equal_alone_and_in_the_mix(td4t):
    from cuda_cube_solver_amd.server import Server

    mix = states.uniform_states(48, 51) + states.flip_heavy_states(4, 52)[:-1]
    alone = GpuSolver(table=td4t, mode="batched").solve(mix)
    with Server(td4t, max_inflight=16) as srv:
        for i, c in enumerate(mix):
            srv.submit(c, tag=i)
        got = srv.collect(len(mix), timeout_s=120)
    assert sorted(got) == list(range(len(mix)))
    for i, c in enumerate(mix):
        r = got[i]
        assert r.length == len(alone[i].moves)
        assert apply_moves(c, [MOVES_HTM[m] for m in r.moves]) == CubieState()


def test_max_bound_below_the_optimum_is_a_clean_error(td4t, reference_results):
    """A bound limit below a cube's optimal length ends both forms with the
    reference's assertion (not solved), and the solver keeps working."""
    ref = reference_results["random1000"]
    i = next(k for k, r in enumerate(ref["results"]) if r["len"] == 17)
    cubes = _scrambled([ref["scrambles"][i]] + ref["scrambles"][300:303])
    for mode in ("plain", "batched"):
        with pytest.raises(AssertionError, match="not solved within max_bound"):
            GpuSolver(table=td4t, mode=mode, max_bound=16).solve(cubes)
    sols = GpuSolver(table=td4t, mode="batched").solve(cubes)
    assert [len(s.moves) for s in sols] == [17] + [r["len"] for r in ref["results"][300:303]]


def test_root_bound_above_max_bound_is_refused_in_both_forms(td4t):
    """A cube whose root bound already exceeds max_bound is given up without a
    round in both forms (a first round at the root bound would otherwise return
    a solution longer than the limit)."""
    rng = states.SplitMix64(71)
    short = [apply_moves(CubieState(), [MOVES_HTM[rng.below(18)] for _ in range(5)]) for _ in range(24)]
    sols = GpuSolver(table=td4t, mode="plain").solve(short)
    # cubes solved in their first round: root bound == optimal length
    tight = [(c, len(s.moves)) for c, s in zip(short, sols) if s.iteration_nodes and
             s.iteration_nodes[0][0] == len(s.moves) and len(s.moves) >= 2]
    assert tight, "no short scramble solved in its first round"
    cube, length = tight[0]
    for mode in ("plain", "batched"):
        with pytest.raises(AssertionError, match="not solved within max_bound"):
            GpuSolver(table=td4t, mode=mode, max_bound=length - 1).solve([cube])
        got = GpuSolver(table=td4t, mode=mode, max_bound=length).solve([cube])
        assert len(got[0].moves) == length


def test_server_stops_promptly_while_busy(td4t):
    """Stopping a server with hard cubes still in flight returns within
    seconds (pending and unfinished cubes are dropped); a new server on the
    same table then serves normally."""
    import time

    from cuda_cube_solver_amd.server import Server

    busy = states.flip_heavy_states(32, 61) + states.uniform_states(96, 62)
    srv = Server(td4t)
    for i, c in enumerate(busy):
        srv.submit(c, tag=i)
    time.sleep(0.3)
    t = time.perf_counter()
    srv.stop()
    assert time.perf_counter() - t < 10.0
    probe = states.uniform_states(4, 63)
    alone = GpuSolver(table=td4t, mode="batched").solve(probe)
    with Server(td4t) as again:
        for i, c in enumerate(probe):
            again.submit(c, tag=i)
        got = again.collect(len(probe), timeout_s=120)
    assert [got[i].length for i in range(len(probe))] == [len(s.moves) for s in alone]


def test_gather_bench_modes(td4t):
    """The random-probe ceiling benchmark runs in every mode on the resident
    d4t table (two 4-byte reads, the same nontemporal, one 16-byte half-block
    load, a streaming pass) and refuses an unknown mode."""
    for mode, n in ((0, 1 << 22), (1, 1 << 22), (3, 1 << 22), (2, 0)):
        ms, cs = ctypes.c_double(), ctypes.c_uint64()
        _lib.check(_lib.lib().ccs_gather_bench(td4t.handle, mode, n, 8, ctypes.byref(ms), ctypes.byref(cs)))
        assert ms.value > 0, mode
    ms, cs = ctypes.c_double(), ctypes.c_uint64()
    assert _lib.lib().ccs_gather_bench(td4t.handle, 4, 1, 8, ctypes.byref(ms), ctypes.byref(cs)) == _lib.CCS_EINVAL
