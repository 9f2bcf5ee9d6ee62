"""Host-side logic of the product (no GPU): the Python mirror of the
reference's interface, the C-ABI library's exports and the LUT arithmetic
the kernels use, evaluated on the CPU."""
import ctypes
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from cuda_cube_solver_amd import _lib, cubie, facelets, kinds, states
from cuda_cube_solver_amd.cubie import MOVES_HTM, CubieState, apply_moves

ROOT = Path(__file__).resolve().parents[1]
SOLVED = "UUUUUUUUURRRRRRRRRFFFFFFFFFDDDDDDDDDLLLLLLLLLBBBBBBBBB"


def _state(d):
    return CubieState(tuple(d["cp"]), tuple(d["co"]), tuple(d["ep"]), tuple(d["eo"]))


def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "cubesolver.h").read_text()
    declared = set(re.findall(r"\b(ccs_[a-z0-9_]+)\s*\(", header))
    assert declared == set(_lib.SIGNATURES), declared ^ set(_lib.SIGNATURES)
    L = _lib.lib()
    missing = [name for name in declared if not hasattr(L, name)]
    assert not missing
    assert L.ccs_version() >= 100


def test_moves_match_reference(index_vectors):
    for m, want in zip(MOVES_HTM, index_vectors["moves"]):
        assert [list(getattr(m, f)) for f in ("cp", "co", "ep", "eo")] == [want[f] for f in ("cp", "co", "ep", "eo")]


@pytest.mark.parametrize("name", ["304", "d4t", "d4", "s4", "d4f", "a4t", "a4"])
def test_kernel_lut_arithmetic_matches_reference_index(index_vectors, name):
    """The kernels' v_perm/LDS index arithmetic (evaluated on the host) gives
    the reference's table index for every kind."""
    arr = O.cubes(index_vectors["states"])
    out = np.zeros(len(arr), dtype=np.uint64)
    _lib.check(_lib.lib().ccs_index_host(name.encode(), arr.ctypes.data, len(arr), out.ctypes.data))
    assert (out == np.array(index_vectors["kinds"][name]["table_index"], dtype=np.uint64)).all()


def test_corner_class_tables_agree_with_classidx():
    """The search reads corner classes from small LDS tables (twist entry,
    placement under a symmetry, rank within a symmetric orbit).  The library
    checks them against the classidx array for all 153,090 raw corner
    coordinates before its first use and refuses to run otherwise; the
    reference's classidx is pinned by its SHA-256 in test_oracle."""
    assert _lib.lib().ccs_index_host(b"d4t", None, 0, None) == 0, _lib.lib().ccs_last_error()


def test_kernel_lut_arithmetic_matches_oracle_on_seeded_states():
    sts = states.uniform_states(300, 99) + states.flip_heavy_states(20, 4)
    arr = O.cubes([(s.cp, s.co, s.ep, s.eo) for s in sts])
    for name in ("304", "d4t", "d4"):
        out = np.zeros(len(arr), dtype=np.uint64)
        _lib.check(_lib.lib().ccs_index_host(name.encode(), arr.ctypes.data, len(arr), out.ctypes.data))
        assert (out == O.table_index(O.kind(name), arr)).all()


def test_kernel_combination_ranks_cover_every_placement():
    """Every placement of the four D corners (the corner coordinate's 8C4 rank)
    and of the four U edges among the non-slice positions (304's 8C4 rank),
    with parities fixed up by swapping two other cubies: kernel arithmetic
    equals the oracle's index."""
    from itertools import combinations, permutations

    sts = []
    for k, pos in enumerate(combinations(range(8), 4)):
        cp = [0] * 8
        d_iter, u_iter = iter((4, 5, 6, 7)), iter((0, 1, 2, 3))
        for i in range(8):
            cp[i] = next(d_iter) if i in pos else next(u_iter)
        non_e = [0, 1, 2, 3, 4, 5, 6, 7]  # positions of the non-slice edges (8..11 are the E slice)
        upos = list(combinations(non_e, 4))[k]
        ep = [0] * 12
        ui, di = iter((0, 1, 2, 3)), iter((4, 5, 6, 7))
        for i in non_e:
            ep[i] = next(ui) if i in upos else next(di)
        ep[8:] = [8, 9, 10, 11]
        if cubie.permutation_parity(cp) != cubie.permutation_parity(ep):
            ep[10], ep[11] = ep[11], ep[10]
        sts.append(CubieState(tuple(cp), (0,) * 8, tuple(ep), (0,) * 12))
    assert all(s.is_legal() for s in sts) and len(sts) == 70
    arr = O.cubes([(s.cp, s.co, s.ep, s.eo) for s in sts])
    for name in ("304", "d4t"):
        out = np.zeros(len(arr), dtype=np.uint64)
        _lib.check(_lib.lib().ccs_index_host(name.encode(), arr.ctypes.data, len(arr), out.ctypes.data))
        assert (out == O.table_index(O.kind(name), arr)).all()


def test_kind_info():
    for name in ("304", "d4t", "d4", "s4", "d4f", "a4t", "a4"):
        k = _lib.kind_info(name)
        assert k.n_edge == kinds.n_edge(name) == O.kind(name).n_edge
        assert k.n_entries == 9930 * k.n_edge
    with pytest.raises(_lib.CcsError):
        _lib.kind_info("nope")


def test_recorded_table_hashes_come_from_complete_builds(table_hashes, reference_results, index_vectors):
    """Every kind's recorded SHA-256 comes from an oracle BFS whose depth
    counts equal the reference's (research/subgroups/results.jsonl), and its
    file sizes are the presets' (kinds.py:115-135)."""
    sizes = {(p["kind"], p["format"]): p["table_bytes"] for p in index_vectors["presets"]}
    assert set(table_hashes) == set(reference_results["depth_counts"])
    for kind, want in table_hashes.items():
        ref = reference_results["depth_counts"][kind]["depth_counts"]
        assert want["depth_counts"] == ref, kind
        assert sum(want["depth_counts"]) == want["entries"] == 9930 * _lib.kind_info(kind).n_edge
        for fmt, key in (("block", "blk"), ("b3", "b3")):
            if (kind, fmt) in sizes:
                assert want[f"{key}_bytes"] == sizes[(kind, fmt)], (kind, fmt)


def test_presets_match_reference(index_vectors):
    for want in index_vectors["presets"]:
        p = kinds.PRESET_BY_NAME[want["name"]]
        assert (p.kind, p.format, p.file, p.entries, p.table_bytes, p.auto) == \
            (want["kind"], want["format"], want["file"], want["entries"], want["table_bytes"], want["auto"])
    assert [p.name for p in kinds.PRESETS] == [w["name"] for w in index_vectors["presets"]]


def test_preset_choice(tmp_path):
    """The reference's ladder and rules (kinds.py:184-211), with this library's
    build memory: the exact-depth build table takes 4 bits per entry, so a
    16 GB card holds d4t but cannot build it here (the reference's 2-bit build
    can: tests/test_cuda_cube_solver.py:169-182 expects d4t).  On an MI355X
    every preset builds."""
    names = lambda pair: tuple(p.name if p else None for p in pair)
    assert names(kinds.choose(tmp_path, int(200e9))) == ("d4t", None)
    assert names(kinds.choose(tmp_path, int(16e9))) == ("a4t", "d4t")
    assert names(kinds.choose(tmp_path, int(8e9), int(16e9))) == ("a4t", "d4t-b3")
    (tmp_path / "prune-ud-d4t-blk.bin").touch()
    assert names(kinds.choose(tmp_path, int(16e9))) == ("d4t", None)
    assert names(kinds.choose(tmp_path, int(16e9), to_build=True)) == ("a4t", "d4t")
    (tmp_path / "prune-ud-d4t-b3.bin").touch()
    assert names(kinds.choose(tmp_path, int(8e9))) == ("d4t-b3", None)
    assert names(kinds.choose(tmp_path, int(1e9))) == (None, None)
    with pytest.raises(ValueError, match="unknown table"):
        kinds.pick(tmp_path, name="nope")


def test_b3_build_memory_at_the_budget_edge(tmp_path):
    """A base-3 build holds the nibble build table, the base-3 table and one
    minimum byte per 124-entry block (no block table): a budget one byte short
    of that falls back to the next preset down."""
    s4 = kinds.PRESET_BY_NAME["s4-b3"]
    n = s4.entries
    assert s4.build_bytes == 4 * ((n + 7) // 8) + 32 * ((n + 154) // 155) + (n + 123) // 124
    need = s4.build_bytes + kinds.SOLVE_MARGIN
    names = lambda pair: tuple(p.name if p else None for p in pair)
    assert names(kinds.choose(tmp_path, need))[0] == "s4-b3"
    assert names(kinds.choose(tmp_path, need - 1))[0] == "304"
    # the block preset of a kind needs its block table instead
    d4t = kinds.PRESET_BY_NAME["d4t"]
    assert d4t.build_bytes == 4 * ((d4t.entries + 7) // 8) + d4t.table_bytes


def test_cli_tables_listing(tmp_path):
    env = {"PYTHONPATH": str(ROOT / "cuda-cube-solver_amd")}
    out = subprocess.run([sys.executable, "-c",
                          "from cuda_cube_solver_amd import kinds; print('\\n'.join(kinds.describe(%r)))" % str(tmp_path)],
                         capture_output=True, text=True, env=env)
    lines = out.stdout.splitlines()
    assert lines[0].split() == ["preset", "VRAM", "cubes/s", "superflip", "build", "VRAM", "file"]
    assert lines[1].startswith("304        1.42GB      6.0      44.6s")
    assert lines[1].endswith("prune-ud-304-blk.bin  (not built)")
    assert lines[-1].endswith("prune-ud-d4-blk.bin  (not built)  (only with --table)")


def test_facelets_match_reference(index_vectors):
    sts = [_state(d) for d in index_vectors["states"][:50]]
    for s, f in zip(sts, index_vectors["facelets"]):
        assert facelets.to_facelets(s) == f
        assert facelets.parse_facelets(f) == s
    assert facelets.parse_facelets(SOLVED) == CubieState()


def _swap(s, *idx):
    n = len(idx) // 2
    out = list(s)
    for t, src in zip(idx[:n], idx[n:]):
        out[t] = s[src]
    return "".join(out)


@pytest.mark.parametrize("mutate, message", [
    (lambda s: s[:53], "54 stickers"),
    (lambda s: s.replace("U", "R", 1), "has 8 stickers"),
    (lambda s: _swap(s, 8, 9, 20, 9, 20, 8), "twisted"),
    (lambda s: _swap(s, 5, 10, 10, 5), "flipped"),
    (lambda s: _swap(s, 5, 7, 10, 19, 7, 5, 19, 10), "swapped"),
])
def test_illegal_facelets(mutate, message):
    with pytest.raises(facelets.FaceletError, match=message):
        facelets.parse_facelets(mutate(SOLVED))


def test_colour_letters():
    s = apply_moves(CubieState(), [MOVES_HTM[m] for m in cubie.parse_moves("R U2 F' L D B2")])
    col = facelets.to_facelets(s).translate(str.maketrans("URFDLB", "WRGYOB"))
    assert facelets.is_facelet_string(col) and facelets.parse_facelets(col) == s


def test_seeded_states_are_legal_and_deterministic():
    a = states.uniform_states(200, 3)
    assert a == states.uniform_states(200, 3) and a != states.uniform_states(200, 4)
    assert all(s.is_legal() for s in a)
    # roughly uniform twist of corner 0 and flip of edge 0
    assert 40 < sum(s.co[0] == 1 for s in a) < 100
    f = states.flip_heavy_states(16, 4)
    assert all(s.is_legal() and s.cp == tuple(range(8)) and s.ep == tuple(range(12)) for s in f)
    assert f[-1].eo == (1,) * 12


def test_cli_reports_bad_lines_in_order(tmp_path):
    """Input errors print `<n> error <reason>` (no table needed when every line is bad)."""
    env = {"PYTHONPATH": str(ROOT / "cuda-cube-solver_amd")}
    out = subprocess.run([sys.executable, "-m", "cuda_cube_solver_amd", "solve", "--tables", str(tmp_path)],
                         input="R X\n\n# comment\n" + SOLVED.replace("U", "R", 1) + "\n", capture_output=True, text=True,
                         env=env)
    assert out.returncode == 0, out.stderr
    assert out.stdout.splitlines() == ["1 error unknown move 'X'", "4 error face U has 8 stickers, expected 9"]


def test_cli_unknown_table(tmp_path):
    env = {"PYTHONPATH": str(ROOT / "cuda-cube-solver_amd")}
    out = subprocess.run([sys.executable, "-m", "cuda_cube_solver_amd", "solve", "--table", "nope"],
                         input="R U\n", capture_output=True, text=True, env=env)
    assert out.returncode != 0 and "unknown table 'nope'" in out.stderr


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus 2 exits with a message naming the
    single-GPU scope, before it touches a device."""
    import subprocess
    import sys
    from pathlib import Path

    root = Path(__file__).resolve().parent.parent
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
