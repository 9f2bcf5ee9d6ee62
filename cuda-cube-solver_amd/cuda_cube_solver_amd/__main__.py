"""Command line: build the pruning tables, or solve cubes read from stdin.

Same commands, flags and output as the reference's CLI
(src/cuda_cube_solver/__main__.py:34-146).
"""
from __future__ import annotations

import argparse
import sys
from pathlib import Path

SOLVE_EPILOG = """\
input (one cube per line, from stdin):
  a scramble in face-turn notation, e.g.  R U R' U2 F'   (the cube that the
      scramble produces from solved), or
  a 54-sticker facelet string in U R F D L B face order, e.g.
      DRLUUBFBRBLURRLRUBLRDDFDLFUFUFFDBRDUBRUFLLFDDBFLUBLRBD
      (any six symbols; the centres define the faces, so colour letters work)
  blank lines and lines starting with # are skipped.

output (one line per cube, in input order):
  <line number> <length> <moves>     an optimal solution in the half-turn
                                     metric; applying <moves> to the cube
                                     solves it
  <line number> error <message>      invalid input (the other lines still run)
"""

PROG = "cuda-cube-solver"
_NO_TABLE = f"{PROG}: no table fits in that GPU memory (the smallest, 304, needs 1.8 GB)"
MEMORY_HELP = ("GPU memory for the table: 'auto' (default: the free device memory) or a size in GB; "
               "picks the fastest preset that fits (see `tables`)")


def main(argv: list[str] | None = None) -> None:
    ap = argparse.ArgumentParser(prog=PROG, description="Optimal Rubik's cube solver (half-turn metric) on an AMD GPU.")
    sub = ap.add_subparsers(dest="command", required=True)
    build = sub.add_parser("build-table", help="build a table now instead of on first use")
    build.add_argument("--out", type=Path, default=Path("tables"), help="output directory (default: tables)")
    build.add_argument("--memory", default="auto", help=MEMORY_HELP)
    build.add_argument("--table", help="a preset name (see `tables`)")
    tables = sub.add_parser("tables", help="list the table presets and the one `solve` would use")
    tables.add_argument("--tables", type=Path, default=Path("tables"), help="table directory (default: tables)")
    tables.add_argument("--memory", default="auto", help=MEMORY_HELP)
    solve = sub.add_parser("solve", help="solve cubes read from stdin (scrambles or facelet strings)",
                           description="Solve cubes optimally.", epilog=SOLVE_EPILOG,
                           formatter_class=argparse.RawDescriptionHelpFormatter)
    solve.add_argument("--tables", type=Path, default=Path("tables"),
                       help="table directory; missing tables are built there (default: tables)")
    solve.add_argument("--memory", default="auto", help=MEMORY_HELP)
    solve.add_argument("--table", help="use this preset instead of choosing by --memory")
    solve.add_argument("--no-build", action="store_true", help="fail instead of building a missing table")
    solve.add_argument("--format", choices=["auto", "moves", "facelets"], default="auto",
                       help="input format (default: auto-detect per line)")
    solve.add_argument("--batch", type=int, default=100, help="cubes searched together on the GPU (default: 100)")
    solve.add_argument("--mode", default="step", choices=["step", "batched"],
                       help="search: step (per-round launches, default) or batched (one persistent launch)")
    args = ap.parse_args(argv)

    from . import kinds

    if getattr(args, "table", None) and args.command == "solve" and args.table not in kinds.PRESET_BY_NAME:
        sys.exit(f"{PROG}: unknown table {args.table!r}; choose from {', '.join(kinds.PRESET_BY_NAME)}")

    if args.command == "tables":
        best, better = _pick(args.tables, args.memory, None)
        print("\n".join(kinds.describe(args.tables)))
        print(f"\nsolve would use: {best.name if best else 'nothing (not enough GPU memory)'}"
              + (f"; {better.name} fits too but can't be built here ({better.build_bytes / 1e9:.1f} GB)"
                 if better else ""))
        return

    if args.command == "build-table":
        from . import table

        if args.table and args.table not in kinds.PRESET_BY_NAME:
            sys.exit(f"{PROG}: unknown table {args.table!r}; choose from {', '.join(kinds.PRESET_BY_NAME)}")
        preset = _pick(args.out, args.memory, args.table, to_build=True)[0]
        if preset is None:
            sys.exit(_NO_TABLE)
        _log(f"building table {preset.name} in {args.out}/")
        print(table.build_preset(args.out, preset, log=lambda m: _log(f"  {m}")))
        return

    from .cubie import MOVES_HTM, CubieState, apply_moves, parse_moves
    from .facelets import FaceletError, is_facelet_string, parse_facelets
    from .solver import NoTableFits, TableNotFound, open_solver

    cubes: list[tuple[int, CubieState]] = []
    errors: list[tuple[int, str]] = []
    for number, raw in enumerate(sys.stdin, start=1):
        line = raw.strip()
        if not line or line.startswith("#"):
            continue
        try:
            if args.format == "facelets" or (args.format == "auto" and is_facelet_string(line)):
                state = parse_facelets(line)
            else:
                state = apply_moves(CubieState(), [MOVES_HTM[m] for m in parse_moves(line)])
        except (FaceletError, KeyError) as exc:
            reason = exc.args[0] if isinstance(exc, FaceletError) else f"unknown move {exc.args[0]!r}"
            errors.append((number, f"{number} error {reason}"))
            continue
        cubes.append((number, state))

    def flush_errors(before: float) -> None:
        while errors and errors[0][0] < before:
            print(errors.pop(0)[1], flush=True)

    if cubes:
        try:
            solver = open_solver(args.tables, args.memory, args.table, build=not args.no_build, log=_log)
        except (TableNotFound, NoTableFits, ValueError) as exc:
            sys.exit(f"{PROG}: {exc}")
        solver.mode = "batched" if args.mode == "batched" else "plain"
        for start in range(0, len(cubes), args.batch):
            chunk = cubes[start:start + args.batch]
            for (number, _), sol in zip(chunk, solver.solve([state for _, state in chunk])):
                flush_errors(number)
                print(f"{number} {len(sol.moves)} {sol.text()}".rstrip(), flush=True)
    flush_errors(float("inf"))


def _pick(tables: Path, memory: str, name: str | None, to_build: bool = False):
    from . import kinds

    try:
        return kinds.pick(tables, memory, name, to_build)
    except ValueError as exc:
        sys.exit(f"{PROG}: {exc}")


def _log(message: str) -> None:
    print(f"{PROG}: {message}", file=sys.stderr, flush=True)


if __name__ == "__main__":
    main()
