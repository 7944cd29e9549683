# SPDX-License-Identifier: GPL-3.0-or-later
"""The molecule wire and disk formats (pm7amd.formats), checked on the host: the XYZ reader of
pm7-rs (src/system.rs:149-345) on the reference's own example files (tests/golden/xyz, copied from
the reference's examples/) and on the malformed inputs its parser names, the XYZ writer, and the
``energy --json`` record (src/bin/pm7_rs.rs:374-441) with <S^2> and the two-channel gap."""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

import pm7amd
from pm7amd import formats as F

XYZ = Path(__file__).parent / "golden" / "xyz"
A2B = 1.0 / 0.529177210903


def test_reference_examples_parse():
    w = F.read_xyz_file(XYZ / "water.xyz")
    assert w.z.tolist() == [8, 1, 1]
    assert np.allclose(w.xyz, np.array([[0, 0, 0], [0.9584, 0, 0], [-0.24, 0.9278, 0]]) * A2B, atol=0, rtol=0)
    e = F.read_xyz_file(XYZ / "ethanol.xyz")
    assert e.z.tolist() == [6, 6, 8, 1, 1, 1, 1, 1, 1]
    assert abs(e.xyz[8, 0] - 1.9994 * A2B) < 1e-15
    h2s = F.read_xyz_file(XYZ / "h2s.xyz")
    assert h2s.z.tolist() == [16, 1, 1] and abs(h2s.xyz[1, 2] - 0.9430 * A2B) < 1e-15
    assert F.read_xyz_file(XYZ / "methyl.xyz", multiplicity=2).multiplicity == 2


def test_symbols_case_numbers_and_bom():
    text = "\ufeff4\ncomment with extra words\ncL 0 0 0\n8 1 0 0\nh 0 1 0 extra columns 7\nNO 0 0 1\n"
    m = F.read_xyz(text)
    assert m.z.tolist() == [17, 8, 1, 102]  # "NO" normalises to nobelium, as the reference does
    assert F.symbol_to_z("0") is None and F.symbol_to_z("103") is None and F.symbol_to_z("102") == 102
    assert F.symbol_to_z("Xx") is None and F.symbol_to_z("") is None


@pytest.mark.parametrize("text, status, fragment", [
    ("", 1, "empty XYZ"),
    ("two\n\nH 0 0 0\n", 1, "invalid XYZ atom count"),
    ("2\n\nH 0 0 0\n", 1, "XYZ ended before atom 1"),
    ("1\n\nH 0 0\n", 1, "line 3 has fewer than 4 fields"),
    ("1\n\nQq 0 0 0\n", 1, "unknown element on line 3: Qq"),
    ("1\n\nH 0 zero 0\n", 1, "line 3: invalid floating point value: zero"),
    ('1\nLattice="5 0 0 0 5 0 0 0 5" pbc="T T T"\nH 0 0 0\n', 4, "periodic"),
    ("1\nLattice=5\nH 0 0 0\n", 4, "periodic"),
])
def test_malformed_inputs_are_refused(text, status, fragment):
    with pytest.raises(pm7amd.Pm7Error) as exc:
        F.read_xyz(text)
    assert exc.value.status == status and fragment in str(exc.value)


def test_lattice_key_needs_a_word_boundary():
    # "MyLattice=" is not the extended-XYZ key (src/system.rs:275-279): still a molecule.
    m = F.read_xyz('1\nMyLattice="1 2 3" pbc="T T T"\nHe 0 0 0\n')
    assert m.z.tolist() == [2]


def test_writer_round_trips():
    m = F.read_xyz_file(XYZ / "ethanol.xyz")
    text = F.to_xyz_string(m, 'a "quoted"\ntitle')
    assert text.splitlines()[1] == "a  quoted  title"
    back = F.read_xyz(text)
    assert back.z.tolist() == m.z.tolist()
    assert np.max(np.abs(back.xyz - m.xyz)) < 1e-11


def _result(unrestricted=False, **kw):
    n = 4
    base = dict(status=0, density=np.eye(n), spin_density=None, mo_energies=np.zeros(n), mo_coeff=np.eye(n),
                n_occ=2, mo_energies_beta=None, mo_coeff_beta=None, n_occ_beta=None, electronic_ev=-10.0,
                core_ev=5.0, total_ev=-5.0, heat_of_formation_kcal=-1.25, correction_kcal=0.0,
                charges=np.array([0.25, -0.25]), dipole_debye=np.array([1.0, -2.0, 0.5]), dipole_magnitude=2.29,
                dipole_point_charge_debye=np.array([0.1, 0.2, 0.3]), dipole_sp_hybrid_debye=np.array([0.0, 0.0, -0.0]),
                dipole_pd_hybrid_debye=np.zeros(3), homo_ev=-11.0, lumo_ev=1.5, homo_ev_beta=None,
                lumo_ev_beta=None, iterations=7, converged=True, unrestricted=unrestricted, density_error=1e-9)
    base.update(kw)
    return pm7amd.Pm7Result(**base)


def test_energy_json_record_restricted():
    r = _result()
    s = F.energy_json(r, "pm7")
    d = json.loads(s)
    assert list(d) == ["method", "energy_ev", "heat_of_formation_kcal", "charges", "dipole_debye",
                       "dipole_point_charge_debye", "dipole_sp_hybrid_debye", "dipole_pd_hybrid_debye",
                       "orbital_source", "homo_ev", "lumo_ev", "gap_ev"]
    assert '"energy_ev":-5.000000000000,' in s and '"charges":[0.250000000000,-0.250000000000]' in s
    assert d["gap_ev"] == 12.5 and d["orbital_source"] == "molecular"
    assert "spin_squared" not in d and r.spin_squared() is None


def _orthonormal(n, seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(n, n)))
    return q


def test_spin_squared_matches_alpha_beta_overlap():
    n, na, nb = 6, 4, 2
    ca, cb = _orthonormal(n, 1)[:, :na], _orthonormal(n, 2)[:, :nb]
    pa, pb = ca @ ca.T, cb @ cb.T
    r = _result(True, density=pa + pb, spin_density=pa - pb, n_occ=na, n_occ_beta=nb,
                homo_ev=-9.0, lumo_ev=2.0, homo_ev_beta=-8.5, lumo_ev_beta=0.5)
    sz = 0.5 * (na - nb)
    want = sz * (sz + 1.0) + nb - np.trace(pa @ pb)
    assert abs(r.spin_squared() - want) < 1e-12
    # beta orbitals inside the alpha space (restricted open shell): <S^2> is exactly S(S+1)
    pb2 = ca[:, :nb] @ ca[:, :nb].T
    r2 = _result(True, density=pa + pb2, spin_density=pa - pb2, n_occ=na, n_occ_beta=nb)
    assert abs(r2.spin_squared() - sz * (sz + 1.0)) < 1e-12
    # two-channel frontier: higher HOMO (beta -8.5), lower LUMO (beta 0.5)
    assert abs(r.gap_ev - 9.0) < 1e-15
    d = json.loads(F.energy_json(r))
    assert list(d)[-6:] == ["homo_ev", "lumo_ev", "homo_ev_beta", "lumo_ev_beta", "gap_ev", "spin_squared"]
    assert "(SZ) = 1.000000" in F.energy_text(r)
