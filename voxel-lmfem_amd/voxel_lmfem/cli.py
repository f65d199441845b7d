"""Command line: ``python -m voxel_lmfem solve ... [--backend cuda|gpu|hip]`` and
``python -m voxel_lmfem import-model ...``.

Mirrors the reference's `solve` subcommand (src/voxel_lmfem/cli.py:166-171,
213-223, 35-40, 234-235): NPZ + JSON input, refuses an existing output,
writes JSON with open('x') and allow_nan=False, prints
{output, backend, max_true_residual}, and exits with status 2 and
"error: ..." on input or solver errors.  `python -O` is refused.
`import-model` takes the reference's flags (cli.py:155-164, 203-209) and
prints {status, output, shape, domain_size, phases}.
"""
import argparse
import json
import sys
from pathlib import Path

import numpy as np

from . import __version__


def _plain(value):
    if isinstance(value, dict):
        return {k: _plain(v) for k, v in value.items()}
    if isinstance(value, (list, tuple)):
        return [_plain(v) for v in value]
    if isinstance(value, np.ndarray):
        return value.tolist()
    if isinstance(value, np.generic):
        return value.item()
    return value


def _save(path, value):
    path.parent.mkdir(parents=True, exist_ok=True)
    with path.open("x", encoding="utf-8") as stream:
        json.dump(_plain(value), stream, indent=2, allow_nan=False)


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    if sys.flags.optimize:
        raise SystemExit("Run without -O: numerical assertions are required.")
    parser = argparse.ArgumentParser(prog="voxel-lmfem",
                                     description="Periodic fully integrated HEX8 elastic homogenization (MI355X).")
    parser.add_argument("--version", action="version", version=__version__)
    sub = parser.add_subparsers(dest="command", required=True)
    importer = sub.add_parser("import-model", help="Import and validate an external labelled voxel volume")
    importer.add_argument("--input", type=Path, required=True)
    importer.add_argument("--materials", type=Path, required=True)
    importer.add_argument("--out", type=Path, required=True, help="New output directory")
    importer.add_argument("--axes", default="zyx", choices=("xyz", "xzy", "yxz", "yzx", "zxy", "zyx"),
                          help="Source array dimension order (default: zyx)")
    size = importer.add_mutually_exclusive_group()
    size.add_argument("--spacing", type=float, nargs=3, metavar=("DX", "DY", "DZ"))
    size.add_argument("--domain-size", type=float, nargs=3, metavar=("LX", "LY", "LZ"))
    importer.add_argument("--key", default="phase_id", help="Label-array key in an NPZ file")
    importer.add_argument("--orientation", type=Path, help="NPY rotation array or NPZ orientation/axis arrays")
    solve = sub.add_parser("solve", help="Homogenize an NPZ phase field and JSON material table")
    solve.add_argument("--input", type=Path, required=True)
    solve.add_argument("--materials", type=Path, required=True)
    solve.add_argument("--output", type=Path, required=True)
    solve.add_argument("--backend", choices=("cuda", "gpu", "hip", "cpu"), default="cuda")
    solve.add_argument("--precision", choices=("mixed", "fp64"), default="mixed")
    solve.add_argument("--warmup", action="store_true")
    solve.add_argument("--fields", action="store_true", help="add element-centre field statistics (loads 0, 5)")
    args = parser.parse_args(argv)
    try:
        if args.command == "import-model":
            from .importers import import_model
            report = import_model(args.input, args.materials, args.out, axes=args.axes, spacing=args.spacing,
                                  domain_size=args.domain_size, key=args.key, orientation_path=args.orientation)
            print(json.dumps(dict(status="imported", output=str(args.out.resolve()), shape=report["shape"],
                                  domain_size=report["domain_size"], phases=report["phases"]), indent=2))
        elif args.command == "solve":
            if args.output.exists():
                raise FileExistsError(args.output)
            from .inputs import load_input
            from .solver import homogenize
            phase, materials, domain, orientation = load_input(args.input, args.materials)
            result = homogenize(phase, materials, domain, orientation=orientation, backend=args.backend,
                                warmup=args.warmup, precision=args.precision, fields=args.fields)
            _save(args.output, result)
            print(json.dumps(dict(output=str(args.output), backend=result["backend"],
                                  max_true_residual=result["max_true_residual"])))
    except (ValueError, FileNotFoundError, FileExistsError, RuntimeError) as exc:
        parser.exit(2, f"error: {exc}\n")
    return 0
