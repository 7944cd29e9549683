# SPDX-License-Identifier: GPL-3.0-or-later
"""Pin the CPU FP64 oracle to pm7-rs's own recorded outputs (tools/oracle/baselines/reference.json).

175 cases: closed shells, ions, UHF doublets/triplets/quartets/quintets, H-bonded dimers and wires
(EH+ correction), stacked benzenes (dispersion), a 92-atom alkane past the 7 A feather range, heavy
sp elements (numerical Slater overlaps) and 85 molecules holding a d-shell element (the MNDO/d
kernels: S/P/Cl/Si/Al, the heavier p block, d0/d6/d8/d10 transition-metal halides, oxo anions,
carbonyls, ferrocene and titanocene dichloride).
Tolerances are the north-star bars: 1e-6 eV total energy, 1e-4 kcal/mol heat of formation.
"""

import json
from pathlib import Path

import numpy as np
import pytest

import oracle_binding as ob

CASES = json.loads((Path(__file__).parent / "golden" / "reference_cases.json").read_text())["cases"]


# Orbital-energy tolerance per case (eV). Chloride is a closed-shell atom whose Fock and density
# stay exactly diagonal, so every DIIS error vector is zero up to rounding; the reference's
# relative-scaled DIIS solve (src/scf.rs:1714-1760) then mixes earlier Fock matrices with weights
# set by that rounding, which moves its reported HOMO/LUMO by 1-3e-5 eV while the energy, charges
# and dipole stay at the converged values. The rounding of a different implementation cannot be
# reproduced, so for that case the eigenvalues are held to 5e-5 eV.
ORBITAL_TOL = {"chloride": 5e-5}


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_reference_golden(name):
    c = CASES[name]
    z, xyz = ob.atoms_to_arrays(c["atoms"])
    res, dens, spin, eps, q = ob.run(z, xyz, c["charge"], c["multiplicity"])
    assert res.status == 0, res.message.decode()
    assert abs(res.total_ev - c["energy_ev"]) <= 1e-6
    assert abs(res.hof_kcal - c["heat_of_formation_kcal"]) <= 1e-4
    assert np.max(np.abs(q - np.array(c["charges"]))) <= 1e-6
    for key, got in (("homo_ev", res.homo_ev), ("lumo_ev", res.lumo_ev)):
        if c[key] is not None:
            assert abs(got - c[key]) <= ORBITAL_TOL.get(name, 1e-5), key
    if c["homo_ev_beta"] is not None:
        assert abs(res.homo_ev_beta - c["homo_ev_beta"]) <= 1e-5
    # The dipole follows the density to the SCF's own convergence (RMS dP < 1e-7); the most
    # sensitive case (CfF3, a large f-shell core on a planar ligand set) moves 1.3e-5 D.
    assert np.max(np.abs(np.array(res.dipole[:]) - np.array(c["dipole_debye"]))) <= 2e-5
    assert np.max(np.abs(np.array(res.dipole_sp[:]) - np.array(c["dipole_sp_hybrid_debye"]))) <= 2e-5
    # the p-d hybrid part is what remains of the total (src/dipole.rs:332-340)
    pd_ref = np.array(c["dipole_debye"]) - np.array(c["dipole_point_charge_debye"]) - np.array(c["dipole_sp_hybrid_debye"])
    assert np.max(np.abs(np.array(res.dipole_pd[:]) - pd_ref)) <= 2e-5
    # Invariants of any SCF solution: the Mulliken charges sum to the molecular charge
    # (trace P = electron count), and an RHF density is idempotent, P P = 2 P.
    assert abs(float(np.sum(q)) - c["charge"]) <= 1e-8
    if res.unrestricted == 0:
        assert np.allclose(dens @ dens, 2.0 * dens, atol=1e-6)
