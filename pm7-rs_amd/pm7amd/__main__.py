"""``python -m pm7amd energy FILE.xyz [--json]``: the molecular single point of pm7-rs's
``pm7_rs energy`` (src/bin/pm7_rs.rs:374-470), same JSON record, on the GPU engine."""
from __future__ import annotations

import argparse
import sys

from . import Pm7Error, Pm7Options, run_pm7
from .formats import energy_json, energy_text, read_xyz_file


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m pm7amd")
    sub = ap.add_subparsers(dest="command", required=True)
    e = sub.add_parser("energy", help="PM7 single point of an XYZ molecule")
    e.add_argument("xyz")
    e.add_argument("--json", action="store_true")
    e.add_argument("--charge", type=float, default=0.0)
    e.add_argument("--multiplicity", type=int, default=1)
    e.add_argument("--method", default="pm7")
    e.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    a = ap.parse_args(argv)
    try:
        mol = read_xyz_file(a.xyz, a.charge, a.multiplicity)
        res = run_pm7(mol, options=Pm7Options(method=a.method, charge=a.charge, multiplicity=a.multiplicity,
                                              precision=a.precision))
    except Pm7Error as err:
        print(f"error: {err}", file=sys.stderr)
        return 1
    print(energy_json(res, a.method) if a.json else energy_text(res, a.method))
    return 0


if __name__ == "__main__":
    sys.exit(main())
