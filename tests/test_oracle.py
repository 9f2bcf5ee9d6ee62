"""The CPU oracle against the reference's own vectors and against brute force."""
import hashlib
from itertools import product

import numpy as np
import pytest

import oracle as O


def test_corner_classes_match_reference(index_vectors):
    cc = index_vectors["corner_classes"]
    assert O.n_classes() == cc["n"] == 9930
    assert hashlib.sha256(O.classidx().astype("<u4").tobytes()).hexdigest() == cc["classidx_sha256"]
    assert hashlib.sha256(O.class_rep().astype("<u4").tobytes()).hexdigest() == cc["rep_sha256"]
    assert hashlib.sha256(O.class_stab().astype("<u2").tobytes()).hexdigest() == cc["stab_sha256"]


def test_symmetry_order_matches_reference(index_vectors):
    for i in range(48):
        assert O.sym_matrix(i) == index_vectors["sym48_matrices"][i]


def test_moves_match_reference(index_vectors):
    for m, want in enumerate(index_vectors["moves"]):
        got = O.move(m)[0]
        assert [list(got[f]) for f in ("cp", "co", "ep", "eo")] == [want[f] for f in ("cp", "co", "ep", "eo")]


@pytest.mark.parametrize("name", ["304", "d4t", "d4", "s4", "d4f", "a4t", "a4"])
def test_table_index_matches_reference(index_vectors, name):
    e = index_vectors["kinds"][name]
    k = O.kind(name)
    assert k.n_edge == e["n_edge"]
    if name != "304":
        assert k.nec == e["nec"]
        assert list(k.gen_class) == e["gen_class"]
        assert [list(r) for r in k.gen_rep] == e["gen_rep"]
    got = O.table_index(k, O.cubes(index_vectors["states"]))
    assert (got == np.array(e["table_index"], dtype=np.uint64)).all()


@pytest.mark.parametrize("name", ["304", "d4t", "d4"])
def test_unrank_inverts_the_index(name):
    k = O.kind(name)
    rng = np.random.default_rng(5)
    idx = np.concatenate([[0, k.n_entries - 1], rng.integers(0, k.n_entries, 500)]).astype(np.uint64)
    back = O.table_index(k, O.unrank(k, idx))
    assert (back == idx).all()


def test_group_kind_must_be_normalised():
    with pytest.raises(ValueError, match="not normalised"):
        O.group_kind([((1, 0, 2, 3), (0, 0, 0, 0))])
    with pytest.raises(ValueError, match="not normalised"):
        O.group_kind([((0, 1, 2, 3), (1, 1, 0, 0))])


def test_canonical_prefixes_and_frames(index_vectors):
    for plen in (4, 5):
        p = O.canonical_prefixes(plen)
        assert len(p) == index_vectors[f"prefixes{plen}"]["n"]
        assert hashlib.sha256(p.tobytes()).hexdigest() == index_vectors[f"prefixes{plen}"]["sha256"]
    st = O.cubes(index_vectors["states"][:20])
    for i, fr in enumerate(index_vectors["frames"]):
        assert (O.axis_frames(st[i:i + 1]) == O.cubes(fr)).all()


def _canonical_distances(max_depth):
    """Exact distances of all states within max_depth moves, by BFS over canonical sequences."""
    solved = O.cubes([(range(8), [0] * 8, range(12), [0] * 12)])
    dist = {solved.tobytes(): 0}
    frontier = [solved]
    for d in range(1, max_depth + 1):
        nxt = []
        for s in frontier:
            for m in range(18):
                t = O.compose(s, O.move(m))
                key = t.tobytes()
                if key not in dist:
                    dist[key] = d
                    nxt.append(t)
        frontier = nxt
    return dist


def test_search_with_a_zero_table_is_iterative_deepening():
    """With every bound 0 the pruning reduces to iterative deepening over
    canonical sequences: lengths must equal brute-force BFS distances."""
    dist = _canonical_distances(4)
    rng = np.random.default_rng(3)
    keys = list(dist)
    picks = [keys[i] for i in rng.choice(len(keys), 40, replace=False)]
    arr = np.frombuffer(b"".join(picks), dtype=O.CUBE_DTYPE).copy()
    k = O.kind("304")
    zero = np.zeros(-(-k.n_entries // 124) * 32, dtype=np.uint8)
    res, bad = O.solve_many(k, zero, 0, arr, n_threads=4)
    assert bad == 0
    for r, key in zip(res, picks):
        assert r["length"] == dist[key]
        s = np.frombuffer(key, dtype=O.CUBE_DTYPE).copy()
        for m in r["moves"]:
            s = O.compose(s, O.move(m))
        assert s.tobytes() == O.cubes([(range(8), [0] * 8, range(12), [0] * 12)]).tobytes()


def test_block_and_b3_decoding_closed_forms():
    """Lookups decode the formats kernels.cu:396-434 define."""
    blk = np.zeros(64, dtype=np.uint8)
    words = blk.view(np.uint32)
    words[0] = 7 | (2 << 8) | (3 << 30)     # min 7, entry 0 offset 2, entry 11 offset 3
    words[1] = 1 << 0                       # entry 12 offset 1
    assert O.lookup(blk, 0, 0) == 9
    assert O.lookup(blk, 0, 11) == 10
    assert O.lookup(blk, 0, 12) == 8
    assert O.lookup(blk, 0, 5) == 7
    b3 = np.zeros(32, dtype=np.uint8)
    b3[0] = 9 | (2 << 6)                    # m 9, cut 2
    b3[1] = 1 + 2 * 3 + 0 * 9 + 1 * 27      # digits of entries 0..4: 1, 2, 0, 1, 0
    assert [O.lookup(b3, 1, j) for j in range(5)] == [12, 13, 9, 12, 9]
