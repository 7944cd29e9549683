#!/usr/bin/env python3
# SPDX-License-Identifier: GPL-3.0-or-later
"""Regenerate the bundled PM7 parameter tables from the reference's MOPAC-derived CSV data.

The numbers are MOPAC v23.2.5 PM7 parameters (Apache-2.0, (c) Virginia Tech), as tabulated by
pm7-rs in ``src/data/pm7_elements.csv``, ``pm7_pairs.csv`` and ``pm7_vpar.csv``.  Parameters are
data, not algorithm: the engine must see the same values the reference does, so this script only
re-shapes them into one tab-separated file per table with a fixed column order that the C++
loader (``csrc/params.cpp``) reads with ``strtod``.  Rows whose ``zeta_s`` and ``beta_s`` are both
zero are unparameterised in PM7 and are dropped, as the reference does (``src/params.rs:212-216``).

Usage (in a tree that has the reference checked out)::

    python pm7-rs_amd/data/extract_params.py /root/reference/src/data
"""

from __future__ import annotations

import csv
import sys
from pathlib import Path

ELEMENT_COLUMNS = [
    "z", "n", "ios", "iop", "iod", "core_charge", "eheat_kcal",
    "u_ss", "u_pp", "u_dd", "zeta_s", "zeta_p", "zeta_d", "beta_s", "beta_p", "beta_d",
    "g_ss", "g_sp", "g_pp", "g_p2", "h_sp", "zsn", "zpn", "zdn", "f0sd", "g2sd", "poc",
    "g1_1", "g2_1", "g3_1", "g1_2", "g2_2", "g3_2", "g1_3", "g2_3", "g3_3", "g1_4", "g2_4", "g3_4",
]

HEADER = (
    "# PM7 parameters, MOPAC v23.2.5 (Apache-2.0, (c) 2021 Virginia Tech),\n"
    "# re-shaped from pm7-rs src/data/{src} by extract_params.py. Do not hand edit.\n"
)


def rows(path: Path) -> list[dict[str, str]]:
    lines = [ln for ln in path.read_text().splitlines() if ln.strip() and not ln.startswith("#")]
    return list(csv.DictReader(lines))


def main(src_dir: str) -> None:
    src = Path(src_dir)
    out = Path(__file__).resolve().parent
    for method, prefix in (("pm7", "pm7"), ("pm7ts", "pm7ts")):
        elements = rows(src / f"{prefix}_elements.csv")
        lines = [HEADER.format(src=f"{prefix}_elements.csv"), "\t".join(ELEMENT_COLUMNS) + "\n"]
        for r in elements:
            if float(r["zeta_s"]) == 0.0 and float(r["beta_s"]) == 0.0:
                continue
            lines.append("\t".join(r[c] for c in ELEMENT_COLUMNS) + "\n")
        (out / f"{method}_elements.tsv").write_text("".join(lines))

        pairs = rows(src / f"{prefix}_pairs.csv")
        lines = [HEADER.format(src=f"{prefix}_pairs.csv"), "z_hi\tz_lo\talpb\txfac\n"]
        lines += [f"{r['z_hi']}\t{r['z_lo']}\t{r['alpb']}\t{r['xfac']}\n" for r in pairs]
        (out / f"{method}_pairs.tsv").write_text("".join(lines))

        vpar = rows(src / f"{prefix}_vpar.csv")
        lines = [HEADER.format(src=f"{prefix}_vpar.csv"), "index\tvalue\n"]
        lines += [f"{r['index']}\t{r['value']}\n" for r in vpar]
        (out / f"{method}_vpar.tsv").write_text("".join(lines))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/data")
