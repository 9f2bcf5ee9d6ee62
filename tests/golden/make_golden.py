"""Write the golden fixtures under tests/golden/ (run in the build container only).

Two kinds of fixture come out of this script:

* ``reference_results.json`` -- data the reference already holds: its 1,000
  random scrambles (``data/random1000.txt``) with the optimal lengths it
  recorded (``data/cuda-cube-solver_1000.txt``), the distance-20 sets and the
  superflip, the BFS depth counts of ``research/subgroups/results.jsonl`` and
  the complete-round node counts of ``research/subgroups/hard.jsonl``.  These
  are read as text/JSON, nothing is executed.

* ``index_vectors.json`` -- input/output vectors of the functions that fix the
  table index numbering (symmetry order, corner classes, group-class
  numbering, coordinates, canonical prefixes, search constants, axis frames).
  No reference-held fixture pins these, so they are taken by importing the
  reference's pure-Python host modules (``cubie``, ``geometry``, ``coords``,
  ``kinds``, ``solver.canonical_prefixes``/``search_const``/``frames``,
  ``facelets``) from ``/root/reference/src``.  Only numbers are written.

Usage: ``python tests/golden/make_golden.py`` (needs /root/reference).
"""
from __future__ import annotations

import hashlib
import json
import random
import sys
from pathlib import Path

import numpy as np

REF = Path("/root/reference")
HERE = Path(__file__).resolve().parent


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def reference_results() -> dict:
    data = REF / "data"

    def lines(p: Path) -> list[str]:
        return [l.strip() for l in p.read_text().splitlines() if l.strip()]

    def results(p: Path) -> list[dict]:
        out = []
        for l in lines(p):
            parts = l.split()
            out.append({"i": int(parts[0]), "len": int(parts[2]), "moves": " ".join(parts[3:])})
        return out

    res = {
        "random1000": {"scrambles": lines(data / "random1000.txt"),
                       "results": results(data / "cuda-cube-solver_1000.txt")},
        "first100_d4t": results(data / "single_cube" / "cuda-cube-solver_first100.txt"),
        "first100_d4": results(data / "d4" / "first100.txt"),
        "distance20_random10": {"scrambles": lines(data / "distance20" / "random10.moves.txt"),
                                "results": results(data / "distance20" / "cuda-cube-solver_random10.txt")},
        "superflip": {"scrambles": lines(data / "distance20" / "superflip.txt"),
                      "results": results(data / "distance20" / "cuda-cube-solver_superflip.txt")},
    }
    # Depth counts and complete-round node counts, keyed by survey name; the
    # presets map to survey kinds as the reference's kinds.py comment states
    # (s4 = g1-49, d4f = g1-43, a4t = g1-31, d4t = g1-27, d4 = g1-26, a4 = g1-30)
    # and 304 = g2-49 (the EdgeKind(3, 4) table).
    names = {"g1-49": "s4", "g1-43": "d4f", "g1-31": "a4t", "g1-27": "d4t", "g1-26": "d4", "g1-30": "a4",
             "g2-49": "304"}
    depth = {}
    for l in (REF / "research" / "subgroups" / "results.jsonl").read_text().splitlines():
        d = json.loads(l)
        if d["kind"] in names:
            depth[names[d["kind"]]] = {"entries": d["entries"], "depth_counts": d["depth_counts"]}
    hard = {}
    for l in (REF / "research" / "subgroups" / "hard.jsonl").read_text().splitlines():
        d = json.loads(l)
        if d["kind"] in names:
            hard[names[d["kind"]]] = {k: d[k] for k in ("random_nodes", "d20_nodes", "superflip_nodes",
                                                        "random_len", "d20_len", "superflip_len")}
    res["depth_counts"] = depth
    res["complete_round_nodes"] = hard
    return res


def index_vectors() -> dict:
    sys.path.insert(0, str(REF / "src"))
    from cuda_cube_solver import coords, cubie, facelets, geometry, kinds, solver

    out: dict = {}
    out["sym48_matrices"] = [s.matrix.tolist() for s in geometry.all_symmetries()]
    out["sym16_matrices"] = [s.matrix.tolist() for s in geometry.ud_symmetries()]
    cs, es = coords.ud_actions()
    out["sym16_corner"] = [{"perm": c.perm.tolist(), "shift": c.shift.tolist(), "reflect": c.reflect} for c in cs]
    out["sym16_edge"] = [{"perm": e.perm.tolist(), "flip": e.flip.tolist()} for e in es]
    out["axis_move_map"] = geometry.move_map(geometry.axis_rotation())
    out["moves"] = [{"cp": m.cp, "co": m.co, "ep": m.ep, "eo": m.eo} for m in cubie.MOVES_HTM]

    cc = coords.corner_classes()
    out["corner_classes"] = {
        "n": int(len(cc.rep)),
        "classidx_sha256": sha(cc.classidx.astype("<u4")),
        "rep_sha256": sha(cc.rep.astype("<u4")),
        "stab_sha256": sha(cc.stab.astype("<u2")),
        "classidx_head": cc.classidx[:64].tolist(),
        "rep_head": cc.rep[:64].tolist(),
        "stab_head": cc.stab[:64].tolist(),
    }

    rng = random.Random(20261203)

    def rand_state(n_moves: int) -> cubie.CubieState:
        return cubie.apply_moves(cubie.CubieState(), [cubie.MOVES_HTM[rng.randrange(18)] for _ in range(n_moves)])

    states = [rand_state(rng.randrange(1, 40)) for _ in range(300)]
    out["states"] = [{"cp": s.cp, "co": s.co, "ep": s.ep, "eo": s.eo} for s in states]
    packed = [coords.pack(s) for s in states]
    out["corner_raw"] = [int(coords.corner_raw(c)) for c, _ in packed]

    kind_out = {}
    for name in ("304", "d4t", "d4", "s4", "d4f", "a4t", "a4"):
        k = kinds.get(name)
        entry: dict = {"n_edge": int(k.n_edge)}
        if name == "304":
            idx = [int(coords.table_index(c, e, k)) for c, e in packed]
        else:
            entry["nec"] = int(k.nec)
            entry["gen_class"] = k.gen[0].astype(int).tolist()
            entry["gen_rep"] = k.gen[1].astype(int).tolist()
            entry["group"] = sorted([list(p), list(f)] for p, f in k.group)
            idx = []
            # Host-side GroupKind index (the reference computes it only on the GPU): the
            # E content class after the corner symmetry, exactly as kinds.GroupKind numbers it.
            for c, e in packed:
                p = int(cc.classidx[int(coords.corner_raw(c))])
                cls, s = p >> 4, p & 15
                ce = coords.conj_edges(e, es[s])
                emask = ecount = ocount = 0
                eorder = eflip = nflip = 0
                for i in range(12):
                    b = int(ce[i])
                    if b & 8:
                        emask |= 1 << i
                        eorder |= (b & 3) << (2 * ecount)
                        eflip |= (b >> 4) << ecount
                        ecount += 1
                    else:
                        nflip |= (b >> 4) << ocount
                        ocount += 1
                content = int(coords.EORDER_RANK[eorder]) * 16 + eflip
                gc = int(k.gen[0][content])
                idx.append(cls * k.n_edge + (int(coords.SLICE_RANK[emask]) * 256 + nflip) * k.nec + gc)
        entry["table_index"] = idx
        kind_out[name] = entry
    out["kinds"] = kind_out

    for plen in (4, 5):
        pre = solver.canonical_prefixes(plen)
        out[f"prefixes{plen}"] = {"n": int(len(pre)), "sha256": sha(pre.astype(np.uint8)),
                                  "head": pre[:8].tolist(), "tail": pre[-8:].tolist()}
    out["search_const_hex"] = solver.search_const().hex()
    out["frames"] = []
    for s in states[:20]:
        fr = solver.frames(s)
        out["frames"].append([{"cp": f.cp, "co": f.co, "ep": f.ep, "eo": f.eo} for f in fr])
    out["facelets"] = [facelets.to_facelets(s) for s in states[:50]]
    presets = []
    for p in kinds.PRESETS:
        presets.append({"name": p.name, "kind": p.kind, "format": p.format, "file": p.file,
                        "entries": p.entries, "table_bytes": p.table_bytes, "build_bytes": p.build_bytes,
                        "cubes_per_s": p.cubes_per_s, "superflip_s": p.superflip_s, "auto": p.auto})
    out["presets"] = presets
    return out


def main() -> None:
    (HERE / "reference_results.json").write_text(json.dumps(reference_results()) + "\n")
    (HERE / "index_vectors.json").write_text(json.dumps(index_vectors()) + "\n")
    print("wrote", HERE / "reference_results.json", HERE / "index_vectors.json")


if __name__ == "__main__":
    main()
