# SPDX-License-Identifier: GPL-3.0-or-later
"""XYZ file -> GPU engine -> ``energy --json`` record, end to end, on the reference's example files
(tests/golden/xyz, copied from the reference's examples/). The heats of formation are the ones the
reference publishes for exactly these files (README.md "MOPAC validation" table: pm7-rs column, 8
decimals), and each run is also held to the CPU oracle at the full bar."""

from __future__ import annotations

import json
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle_binding as ob
import pm7amd
from pm7amd import formats as F

pytestmark = pytest.mark.gpu

XYZ = Path(__file__).parent / "golden" / "xyz"
ROOT = Path(__file__).resolve().parent.parent

# (file, multiplicity, published heat of formation in kcal/mol)
PUBLISHED = [("water", 1, -57.78933784), ("methane", 1, -14.39979910), ("h2s", 1, -3.05174959),
             ("methyl", 2, 28.46324808)]


@pytest.fixture(scope="module")
def eng():
    e = pm7amd.Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("name, mult, hof", PUBLISHED)
def test_xyz_to_json_matches_published_and_oracle(eng, name, mult, hof):
    print(f"xyz json {name}", flush=True)
    mol = F.read_xyz_file(XYZ / f"{name}.xyz", multiplicity=mult)
    r = eng.run(mol, pm7amd.Pm7Options(multiplicity=mult), raise_on_error=True)
    rec = json.loads(F.energy_json(r, "pm7"))
    # published to 8 decimals; the engine agrees with the oracle far below that
    assert abs(rec["heat_of_formation_kcal"] - hof) <= 1e-6
    res, dens, spin, eps, q = ob.run(mol.z, mol.xyz, 0.0, mult)
    assert abs(r.total_ev - res.total_ev) <= 1e-6
    assert abs(r.heat_of_formation_kcal - res.hof_kcal) <= 1e-4
    assert len(rec["charges"]) == len(mol.z)
    assert np.max(np.abs(np.array(rec["charges"]) - q)) <= 1e-6
    if mult == 2:
        # one unpaired electron: <S^2> at or just above S(S+1) = 0.75
        assert 0.75 - 1e-9 <= rec["spin_squared"] < 0.8
        assert "homo_ev_beta" in rec and "lumo_ev_beta" in rec
    else:
        assert "spin_squared" not in rec


def test_energy_entry_point_prints_the_record():
    print("xyz json entry point", flush=True)
    env = dict(os.environ)
    env["PYTHONPATH"] = str(ROOT / "pm7-rs_amd") + os.pathsep + env.get("PYTHONPATH", "")
    out = subprocess.run([sys.executable, "-m", "pm7amd", "energy", str(XYZ / "ethanol.xyz"), "--json"],
                         capture_output=True, text=True, timeout=120, env=env)
    assert out.returncode == 0, out.stderr
    rec = json.loads(out.stdout)
    mol = F.read_xyz_file(XYZ / "ethanol.xyz")
    res, dens, spin, eps, q = ob.run(mol.z, mol.xyz, 0.0, 1)
    assert abs(rec["heat_of_formation_kcal"] - res.hof_kcal) <= 1e-4
    assert abs(rec["energy_ev"] - res.total_ev) <= 1e-6
