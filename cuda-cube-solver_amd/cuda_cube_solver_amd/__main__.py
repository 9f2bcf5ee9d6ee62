"""Command line: list the table presets, build a table, or solve cubes read
from stdin.

The commands, flags, output lines and error messages are the reference's
interface (src/cuda_cube_solver/__main__.py:34-146) so that scripts written
for it run unchanged; ``--mode`` selects this library's search forms.
"""
from __future__ import annotations

import argparse
import sys
from pathlib import Path

PROG = "cuda-cube-solver"

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

MEMORY_HELP = ("GPU memory for the table: 'auto' (default: the free device memory) or a size in GB; "
               "picks the fastest preset that fits (see `tables`)")
TABLES_DIR = dict(type=Path, default=Path("tables"))

# (command, help, [(flag, argparse keywords)], extra parser keywords)
COMMANDS = [
    ("build-table", "build a table now instead of on first use", [
        ("--out", dict(TABLES_DIR, help="output directory (default: tables)")),
        ("--memory", dict(default="auto", help=MEMORY_HELP)),
        ("--table", dict(help="a preset name (see `tables`)")),
    ], {}),
    ("tables", "list the table presets and the one `solve` would use", [
        ("--tables", dict(TABLES_DIR, help="table directory (default: tables)")),
        ("--memory", dict(default="auto", help=MEMORY_HELP)),
    ], {}),
    ("solve", "solve cubes read from stdin (scrambles or facelet strings)", [
        ("--tables", dict(TABLES_DIR, help="table directory; missing tables are built there (default: tables)")),
        ("--memory", dict(default="auto", help=MEMORY_HELP)),
        ("--table", dict(help="use this preset instead of choosing by --memory")),
        ("--no-build", dict(action="store_true", help="fail instead of building a missing table")),
        ("--format", dict(choices=["auto", "moves", "facelets"], default="auto",
                          help="input format (default: auto-detect per line)")),
        ("--batch", dict(type=int, default=100, help="cubes searched together on the GPU (default: 100)")),
        ("--mode", dict(default="step", choices=["step", "batched"],
                        help="search: step (per-round launches, default) or batched (one persistent launch)")),
    ], dict(description="Solve cubes optimally.", epilog=SOLVE_EPILOG,
            formatter_class=argparse.RawDescriptionHelpFormatter)),
]


def _parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(prog=PROG, description="Optimal Rubik's cube solver (half-turn metric) on an AMD GPU.")
    sub = ap.add_subparsers(dest="command", required=True)
    for name, help_text, flags, extra in COMMANDS:
        p = sub.add_parser(name, help=help_text, **extra)
        for flag, kw in flags:
            p.add_argument(flag, **kw)
    return ap


def _die(message: str) -> None:
    sys.exit(f"{PROG}: {message}")


def _log(message: str) -> None:
    print(f"{PROG}: {message}", file=sys.stderr, flush=True)


def _preset_pair(tables: Path, memory: str, name: str | None, to_build: bool = False):
    from . import kinds

    try:
        return kinds.pick(tables, memory, name, to_build)
    except ValueError as exc:
        _die(str(exc))


def _check_name(name: str | None) -> None:
    from . import kinds

    if name and name not in kinds.PRESET_BY_NAME:
        _die(f"unknown table {name!r}; choose from {', '.join(kinds.PRESET_BY_NAME)}")


def cmd_tables(args) -> None:
    from . import kinds

    best, better = _preset_pair(args.tables, args.memory, None)
    print("\n".join(kinds.describe(args.tables)))
    line = f"\nsolve would use: {best.name if best else 'nothing (not enough GPU memory)'}"
    if better:
        line += f"; {better.name} fits too but can't be built here ({better.build_bytes / 1e9:.1f} GB)"
    print(line)


def cmd_build_table(args) -> None:
    from . import table

    _check_name(args.table)
    preset = _preset_pair(args.out, args.memory, args.table, to_build=True)[0]
    if preset is None:
        _die("no table fits in that GPU memory (the smallest, 304, needs 1.8 GB)")
    _log(f"building table {preset.name} in {args.out}/")
    print(table.build_preset(args.out, preset, log=lambda m: _log(f"  {m}")))


def read_cubes(lines, fmt: str):
    """(line number, cube or None, error text or None) for every input line
    that is not blank or a comment."""
    from .cubie import MOVES_HTM, CubieState, apply_moves, parse_moves
    from .facelets import FaceletError, is_facelet_string, parse_facelets

    for number, raw in enumerate(lines, start=1):
        text = raw.strip()
        if not text or text.startswith("#"):
            continue
        as_facelets = fmt == "facelets" or (fmt == "auto" and is_facelet_string(text))
        try:
            if as_facelets:
                yield number, parse_facelets(text), None
            else:
                yield number, apply_moves(CubieState(), [MOVES_HTM[m] for m in parse_moves(text)]), None
        except FaceletError as exc:
            yield number, None, str(exc.args[0])
        except KeyError as exc:
            yield number, None, f"unknown move {exc.args[0]!r}"


class InOrder:
    """Prints one line per numbered input line, strictly in input order, as
    the lines' outputs become known."""

    def __init__(self, numbers: list[int]):
        self.order = numbers
        self.ready: dict[int, str] = {}
        self.next = 0

    def put(self, number: int, text: str) -> None:
        self.ready[number] = text

    def flush(self) -> None:
        while self.next < len(self.order) and self.order[self.next] in self.ready:
            print(self.ready.pop(self.order[self.next]), flush=True)
            self.next += 1


def cmd_solve(args) -> None:
    from .solver import NoTableFits, TableNotFound, open_solver

    _check_name(args.table)
    items = list(read_cubes(sys.stdin, args.format))
    out = InOrder([n for n, _, _ in items])
    for number, _, err in items:
        if err is not None:
            out.put(number, f"{number} error {err}")
    todo = [(n, cube) for n, cube, err in items if err is None]
    if todo:
        try:
            solver = open_solver(args.tables, args.memory, args.table, build=not args.no_build, log=_log)
        except (TableNotFound, NoTableFits, ValueError) as exc:
            _die(str(exc))
        solver.mode = "batched" if args.mode == "batched" else "plain"
        for start in range(0, len(todo), args.batch):
            part = todo[start:start + args.batch]
            for (number, _), sol in zip(part, solver.solve([cube for _, cube in part])):
                out.put(number, f"{number} {len(sol.moves)} {sol.text()}".rstrip())
            out.flush()
    out.flush()


HANDLERS = {"tables": cmd_tables, "build-table": cmd_build_table, "solve": cmd_solve}


def main(argv: list[str] | None = None) -> None:
    args = _parser().parse_args(argv)
    HANDLERS[args.command](args)


if __name__ == "__main__":
    main()
