"""GPU test of ccs_solve_noalloc, the allocation-free persistent solve: its
results equal ccs_solve's persistent form on the same cubes (lengths, every
round's bound and node count), on two calls in a row, the second with fewer
cubes, so that state left in the shared device array by one call cannot leak
into the next.  It runs last in the suite (it needs only a table)."""
import ctypes

import pytest

from cuda_cube_solver_amd import _lib, states
from cuda_cube_solver_amd.cubie import MOVES_HTM, CubieState, apply_moves
from cuda_cube_solver_amd.solver import GpuSolver, Tables

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def td4t(reference_results):
    tb, counts = Tables.build("d4t", "block")
    assert counts == reference_results["depth_counts"]["d4t"]["depth_counts"]
    yield tb
    tb.close()


def _noalloc(tb, cubes):
    n = len(cubes)
    res = (_lib.Result * max(n, 1))()
    stats = _lib.SolveStats()
    raw = b"".join(c.to_bytes() for c in cubes)
    buf = ctypes.create_string_buffer(raw, max(len(raw), 1))
    rc = _lib.lib().ccs_solve_noalloc(tb.handle, buf, n, None, res, ctypes.byref(stats))
    return rc, res[:n], stats


def _rounds(r):
    return [(int(r.round_bound[j]), int(r.round_nodes[j])) for j in range(r.n_rounds)]


def test_noalloc_equals_the_persistent_solve(td4t):
    first = states.uniform_states(96, 91) + states.flip_heavy_states(4, 92)[:-1] + [CubieState()]
    second = states.uniform_states(24, 93)
    ref = GpuSolver(table=td4t, mode="batched")
    for cubes in (first, second):
        want = ref.solve(cubes)
        rc, got, stats = _noalloc(td4t, cubes)
        _lib.check(rc)
        assert [r.length for r in got] == [len(s.moves) for s in want]
        # completed rounds: same bounds and nodes (the final round's partial count
        # depends on when the solution turned up)
        for r, s in zip(got, want):
            L = r.length
            assert [x for x in _rounds(r) if x[0] < L] == [x for x in s.iteration_nodes if x[0] < L]
        for c, r in zip(cubes, got):
            assert apply_moves(c, [MOVES_HTM[m] for m in r.moves[: r.length]]) == CubieState()
        assert stats.probes > 0 and stats.launches == 1 and stats.kernel_ms > 0


def test_noalloc_limits(td4t):
    too_many = states.uniform_states(_lib.NOALLOC_MAX_CUBES + 1, 94)
    rc, _, _ = _noalloc(td4t, too_many)
    assert rc == _lib.CCS_EINVAL and "at most" in _lib.lib().ccs_last_error().decode()
    rc, _, _ = _noalloc(td4t, [])
    assert rc == _lib.CCS_OK
    # still correct after the refusals
    cubes = states.uniform_states(8, 95)
    rc, got, _ = _noalloc(td4t, cubes)
    _lib.check(rc)
    assert [r.length for r in got] == [len(s.moves) for s in GpuSolver(table=td4t, mode="batched").solve(cubes)]


def test_block_layout_search_equals_the_half_block_search(reference_results):
    """On a fresh handle the allocation-free solve searches the block table
    itself (it never makes the half-block copy); the first ccs_solve then
    makes the copy and searches that.  Both give the same lengths and the same
    completed-round nodes: the half-block layout holds the block format's
    values bit for bit."""
    tb, counts = Tables.build("304", "block")
    try:
        assert counts == reference_results["depth_counts"]["304"]["depth_counts"]
        cubes = states.uniform_states(48, 96)
        rc, got, _ = _noalloc(tb, cubes)  # block layout
        _lib.check(rc)
        want = GpuSolver(table=tb, mode="batched").solve(cubes)  # half-block layout
        assert [r.length for r in got] == [len(s.moves) for s in want]
        for r, s in zip(got, want):
            L = r.length
            assert [x for x in _rounds(r) if x[0] < L] == [x for x in s.iteration_nodes if x[0] < L]
    finally:
        tb.close()
