"""GPU tests of the table formats: the base-3 encoder against what its
definition fixes, and per-probe lookups in both formats on d4t and d4.

The base-3 format (kernels.cu:408-423, 1825-1869) has no golden vectors in
the reference, and any admissible table gives the same optimal lengths, so a
wrong cut choice or digit rule would not show in solve results.  Here each
sampled base-3 block is checked against the rules themselves, with the
values the reference encodes computed from the block table plus the depth
mod 3 of each entry (kernels.cu:1825-1835) -- never from an encoder:
  * header minimum = the least value of the block;
  * the cut is the lowest c in 0..3 with the largest sum of encoded bounds;
  * every digit follows the encoding rule for that cut;
  * the decoded bound never exceeds the exact BFS depth.
The exact depths come from the GPU build table (kept resident for the test),
whose depth counts equal the reference's and whose block encoding is checked
here entry by entry as well.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from cuda_cube_solver_amd import _lib, states
from cuda_cube_solver_amd.cubie import MOVES_HTM, CubieState
from cuda_cube_solver_amd.solver import Tables

pytestmark = pytest.mark.gpu

BLK, B3 = 124, 155
POW3 = np.array([1, 3, 9, 27, 81], np.int64)


@pytest.fixture(scope="module")
def td4t_depths(reference_results):
    tb, counts = Tables.build("d4t", "both", keep_depths=True)
    assert counts == reference_results["depth_counts"]["d4t"]["depth_counts"]
    assert tb.nbytes("depths") == 4 * ((30199910400 + 7) // 8)
    yield tb
    tb.close()


@pytest.fixture(scope="module")
def t304_depths(reference_results):
    tb, counts = Tables.build("304", "both", keep_depths=True)
    assert counts == reference_results["depth_counts"]["304"]["depth_counts"]
    yield tb
    tb.close()


def _depths(tb, first, n):
    """Exact depths of entries first .. first + n - 1 from the kept build table."""
    w0, w1 = first // 8, (first + n + 7) // 8
    words = np.frombuffer(tb.read("depths", 4 * w0, 4 * (w1 - w0)), dtype=np.uint32)
    nib = ((words[:, None] >> (4 * np.arange(8, dtype=np.uint32))) & 15).reshape(-1)
    return nib[first - 8 * w0: first - 8 * w0 + n].astype(np.int64)


def _check_b3_run(tb, n_entries, b_first, n_blocks):
    """Check base-3 blocks b_first .. b_first + n_blocks - 1 (within the table)."""
    e0 = b_first * B3
    e1 = min((b_first + n_blocks) * B3, n_entries)
    ne = e1 - e0
    depth = _depths(tb, e0, ne)
    assert (depth < 15).all(), "unvisited entry in the build table"
    # block table over the same entries: minimum and 2-bit offsets
    k0, k1 = e0 // BLK, (e1 - 1) // BLK + 1
    blk = np.frombuffer(tb.read("block", 32 * k0, 32 * (k1 - k0)), dtype=np.uint8).reshape(-1, 32)
    bits = np.unpackbits(blk[:, 1:], axis=1, bitorder="little").reshape(-1, 124, 2)
    off_all = (bits[:, :, 0] + 2 * bits[:, :, 1]).astype(np.int64)
    mn_all = blk[:, 0].astype(np.int64)
    e = np.arange(e0, e1)
    kb = e // BLK - k0
    off = off_all[kb, e % BLK]
    m124 = mn_all[kb]
    # the block format is min(d - block minimum, 3) over the block's least depth
    assert (off == np.minimum(depth - m124, 3)).all()
    assert (m124 <= depth).all()
    # the reference's encoded value: the block value; a saturated offset steps up
    # from minimum + 3 to the depth's residue mod 3 (kernels.cu:1825-1835)
    value = np.where(off < 3, m124 + off, m124 + 3 + ((depth % 3) - (m124 + 3)) % 3)
    assert (value <= depth).all()
    nb = (e1 - e0 + B3 - 1) // B3
    pad = nb * B3 - ne
    v = np.concatenate([value, np.full(pad, -1)]).reshape(nb, B3)
    d = np.concatenate([depth, np.full(pad, 99)]).reshape(nb, B3)
    valid = v >= 0
    big = np.int64(1 << 20)
    m = np.where(valid, v, big).min(axis=1)
    raw = np.frombuffer(tb.read("b3", 32 * b_first, 32 * nb), dtype=np.uint8).reshape(nb, 32).astype(np.int64)
    assert (raw[:, 0] & 63 == m).all(), "header minimum"
    # sum of encoded bounds for every cut; the chosen cut is the lowest maximiser
    x = v - m[:, None]
    scores = np.stack([np.where(valid, np.where(x <= c, m[:, None], np.where(x == c + 1, m[:, None] + c + 1,
                                                                           m[:, None] + c + 2)), 0).sum(axis=1)
                       for c in range(4)], axis=1)
    want_cut = scores.argmax(axis=1)  # argmax returns the first (lowest) maximiser
    cut = raw[:, 0] >> 6
    assert (cut == want_cut).all(), f"cut differs in {int((cut != want_cut).sum())} blocks"
    # digits: five per byte, little-endian base 3, entry 5q + k at digit k of byte 1 + q
    assert (raw[:, 1:] < 243).all()
    digits = ((raw[:, 1:, None] // POW3) % 3).reshape(nb, B3)
    want = np.where(~valid, 0, np.where(x <= cut[:, None], 0, np.where(x == cut[:, None] + 1, 1, 2)))
    assert (digits == want).all(), "digit rule"
    decoded = np.where(digits == 0, m[:, None], m[:, None] + cut[:, None] + digits)
    assert (np.where(valid, decoded <= d, True)).all(), "a decoded bound exceeds the exact depth"
    shared_best = (scores == scores.max(axis=1, keepdims=True)).sum(axis=1) > 1
    return int(shared_best.sum())


@pytest.mark.parametrize("kind", ["304", "d4t"])
def test_b3_encoder_follows_its_definition(kind, t304_depths, td4t_depths):
    """Runs of consecutive base-3 blocks at seeded places, plus the first and
    the (partial) last block of the table."""
    tb = t304_depths if kind == "304" else td4t_depths
    n_entries = _lib.kind_info(kind).n_entries
    nb = (n_entries + B3 - 1) // B3
    rng = np.random.default_rng(2024)
    run = 512
    starts = [0, nb - run] + [int(s) for s in rng.integers(0, nb - run, 40)]
    ties = 0
    for s in starts:
        ties += _check_b3_run(tb, n_entries, s, run)
    # the tie rule was exercised: blocks whose best score is shared by several cuts
    assert ties > 0


def _lookup_blocks(tb, fmt, idx):
    """Lower bound of each index read through the oracle's decoder, fetching
    only the 32-byte blocks that hold them."""
    per = BLK if fmt == "block" else B3
    code = 0 if fmt == "block" else 1
    out = []
    cache = {}
    for i in idx:
        b = int(i) // per
        if b not in cache:
            cache[b] = np.frombuffer(tb.read(fmt, 32 * b, 32), dtype=np.uint8).copy()
        out.append(O.lookup(cache[b], code, int(i) % per))
    return out


def _probe(tb, sts, move):
    n = len(sts)
    idx = np.zeros(n, np.uint64)
    lb = np.zeros(n, np.uint32)
    buf = ctypes.create_string_buffer(b"".join(s.to_bytes() for s in sts))
    _lib.check(_lib.lib().ccs_probe(tb.handle, buf, n, move, idx.ctypes.data, lb.ctypes.data, None))
    return idx, lb


@pytest.fixture(scope="module")
def td4_both(reference_results):
    tb, counts = Tables.build("d4", "both")
    assert counts == reference_results["depth_counts"]["d4"]["depth_counts"]
    yield tb
    tb.close()


@pytest.mark.parametrize("kind", ["d4t", "d4"])
@pytest.mark.parametrize("fmt", ["block", "b3"])
def test_probe_index_and_bound_in_both_formats(kind, fmt, td4t_depths, td4_both):
    """The device's table index and decoded bound of moved states equal the
    oracle's index and decoder, per probe, in the block and base-3 formats."""
    tb = td4t_depths if kind == "d4t" else td4_both
    k = O.kind(kind)
    sts = states.uniform_states(96, 81) + states.flip_heavy_states(8, 82)
    tb.use(fmt)
    try:
        for move in [-1, 0, 4, 8, 10, 14, 17, -2, -7, -19]:
            idx, lb = _probe(tb, sts, move)
            if move >= 0:
                moved = [s.then(MOVES_HTM[move]) for s in sts]
            elif move <= -2:
                moved = [MOVES_HTM[-2 - move].then(s) for s in sts]
            else:
                moved = sts
            want = O.table_index(k, O.cubes([(s.cp, s.co, s.ep, s.eo) for s in moved]))
            assert (idx == want).all(), f"index, move {move}"
            assert [int(v) for v in lb] == _lookup_blocks(tb, fmt, want), f"bound, move {move}"
    finally:
        tb.use("block")


def test_d4_b3_table_is_bit_exact(td4_both, table_hashes):
    """d4-b3 bytes equal the oracle BFS's (tests/golden/table_sha256.json)."""
    import hashlib

    want = table_hashes["d4"]
    assert td4_both.nbytes("b3") == want["b3_bytes"]
    h = hashlib.sha256()
    for off in range(0, want["b3_bytes"], 1 << 28):
        h.update(td4_both.read("b3", off, min(1 << 28, want["b3_bytes"] - off)))
    assert h.hexdigest() == want["b3_sha256"]


def test_b3_only_build_matches_the_both_build(td4t_depths, table_hashes):
    """A base-3-only build (block minima from the build table, no block table
    held) gives the same bytes as the base-3 table of a full build."""
    import hashlib

    tb, _ = Tables.build("d4t", "b3")
    try:
        assert tb.nbytes("block") == 0
        h = hashlib.sha256()
        n = tb.nbytes("b3")
        for off in range(0, n, 1 << 28):
            h.update(tb.read("b3", off, min(1 << 28, n - off)))
        assert h.hexdigest() == table_hashes["d4t"]["b3_sha256"]
    finally:
        tb.close()


@pytest.mark.parametrize("kind", ["s4", "d4f", "a4t", "a4"])
def test_other_presets_are_bit_exact(kind, table_hashes, reference_results):
    """The remaining presets' kinds (s4-b3, d4f-b3, a4t, a4t-b3, a4, a4-b3):
    both files equal the oracle BFS's SHA-256 and the depth counts equal the
    reference's (research/subgroups/results.jsonl)."""
    import hashlib

    if kind not in table_hashes:
        pytest.skip(f"no oracle hashes recorded for {kind}")
    want = table_hashes[kind]
    tb, counts = Tables.build(kind, "both")
    try:
        assert counts == want["depth_counts"] == reference_results["depth_counts"][kind]["depth_counts"]
        for fmt, key in (("block", "blk"), ("b3", "b3")):
            n = tb.nbytes(fmt)
            assert n == want[f"{key}_bytes"]
            h = hashlib.sha256()
            for off in range(0, n, 1 << 28):
                h.update(tb.read(fmt, off, min(1 << 28, n - off)))
            assert h.hexdigest() == want[f"{key}_sha256"], f"{kind} {fmt}"
    finally:
        tb.close()
