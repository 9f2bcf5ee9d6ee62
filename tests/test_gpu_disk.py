"""GPU tests of the on-disk path and the command line: a table built by
`build-table` equals the oracle's bytes and solves through `solve --no-build`;
a b3 table survives save and open byte for byte."""
import hashlib
import json
import os
import subprocess
import sys
from pathlib import Path

import pytest

from cuda_cube_solver_amd.solver import Tables

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]


def _file_sha(path: Path) -> str:
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for block in iter(lambda: f.read(1 << 24), b""):
            h.update(block)
    return h.hexdigest()


def _cli(args, stdin=""):
    env = dict(os.environ, PYTHONPATH=str(ROOT / "cuda-cube-solver_amd"))
    return subprocess.run([sys.executable, "-m", "cuda_cube_solver_amd", *args], input=stdin, capture_output=True,
                          text=True, env=env, timeout=600)


def test_build_table_then_solve_from_disk(tmp_path, table_hashes, reference_results):
    out = _cli(["build-table", "--table", "304", "--out", str(tmp_path)])
    assert out.returncode == 0, out.stderr
    path = tmp_path / "prune-ud-304-blk.bin"
    assert out.stdout.strip() == str(path)
    assert path.stat().st_size == table_hashes["304"]["blk_bytes"]
    assert _file_sha(path) == table_hashes["304"]["blk_sha256"]
    meta = json.loads(path.with_suffix(".json").read_text())
    assert meta["preset"] == "304" and meta["depth_counts"] == table_hashes["304"]["depth_counts"]
    assert not list(tmp_path.glob("*.tmp"))

    ref = reference_results["random1000"]
    facelets = "DRLUUBFBRBLURRLRUBLRDDFDLFUFUFFDBRDUBRUFLLFDDBFLUBLRBD"  # the kociemba README example
    lines = ref["scrambles"][:5] + ["R X", facelets]
    out = _cli(["solve", "--tables", str(tmp_path), "--table", "304", "--no-build"], "\n".join(lines) + "\n")
    assert out.returncode == 0, out.stderr
    rows = out.stdout.splitlines()
    assert len(rows) == 7
    for k in range(5):
        n, length = rows[k].split()[:2]
        assert (int(n), int(length)) == (k + 1, ref["results"][k]["len"])
    assert rows[5] == "6 error unknown move 'X'"
    n, length, *moves = rows[6].split()
    assert n == "7" and int(length) == len(moves) <= 20
    from cuda_cube_solver_amd.cubie import MOVES_HTM, CubieState, apply_moves, parse_moves
    from cuda_cube_solver_amd.facelets import parse_facelets

    cube = parse_facelets(facelets)
    assert apply_moves(cube, [MOVES_HTM[m] for m in parse_moves(" ".join(moves))]) == CubieState()


def test_b3_table_save_and_open(tmp_path, table_hashes):
    tb, _ = Tables.build("304", "b3")
    try:
        path = tmp_path / "prune-ud-304-b3.bin"
        tb.save("b3", path)
    finally:
        tb.close()
    assert path.stat().st_size == table_hashes["304"]["b3_bytes"]
    assert _file_sha(path) == table_hashes["304"]["b3_sha256"]
    reopened = Tables.open(path, "304", "b3")
    try:
        h = hashlib.sha256()
        n = reopened.nbytes("b3")
        for off in range(0, n, 1 << 28):
            h.update(reopened.read("b3", off, min(1 << 28, n - off)))
        assert h.hexdigest() == table_hashes["304"]["b3_sha256"]
    finally:
        reopened.close()
