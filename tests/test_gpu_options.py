# SPDX-License-Identifier: GPL-3.0-or-later
"""Every option of the drop-in boundary on the GPU against the CPU oracle run with the same option
(the oracle's option paths are pinned to pm7-rs's own numbers in test_oracle_options.py):
methods (PM7-TS, PM7-, PM7-HH), level shift, accelerators, plain iteration, forced UHF, force_dpath,
smooth dispersion, dipole origins of ions, seeded densities and spin densities; plus the refusals
(PM7-Sparkle, unknown methods, a method that does not match the loaded tables, more electrons of one
spin than orbitals) with their error payloads."""

from __future__ import annotations

import ctypes as C
import json
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle_binding as ob
import pm7amd
from pm7amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = json.loads((Path(__file__).parent / "golden" / "reference_cases.json").read_text())["cases"]
A = ob.ANGSTROM_TO_BOHR

ORACLE_KEYS = {"method": {"pm7": 0, "pm7-ts": 1, "pm7-minus": 3, "pm7-hh": 4},
               "accelerator": {"none": 0, "cdiis": 1, "adiis-cdiis": 2},
               "reference": {"auto": 0, "rhf": 1, "uhf": 2},
               "dipole_origin": {"centre-of-mass": 0, "coordinates": 1, "centre-of-charge": 2}}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = pm7amd.Engine(0)
    yield e
    e.close()


def both(eng, atoms, charge=0.0, mult=1, dens_bar=1e-7, **opts):
    """Engine and oracle with the same options; returns (engine result, oracle result tuple)."""
    mol = pm7amd.Molecule.from_symbols_angstrom(atoms, charge, mult)
    r = eng.run(mol, pm7amd.Pm7Options(charge=charge, multiplicity=mult, **opts), raise_on_error=False)
    okw = {}
    for k, v in opts.items():
        if k in ("initial_density", "initial_spin_density"):
            okw[k] = v
        elif k in ORACLE_KEYS:
            okw[k] = ORACLE_KEYS[k][v]
        elif isinstance(v, bool):
            okw[k] = int(v)
        else:
            okw[k] = v
    z, xyz = ob.atoms_to_arrays(atoms)
    ref = ob.run(z, xyz, charge, mult, **okw)
    res, dens, spin, eps, q = ref
    assert r.status == 0 and res.status == 0, (r.message, res.message)
    assert abs(r.total_ev - res.total_ev) <= 1e-6
    assert abs(r.heat_of_formation_kcal - res.hof_kcal) <= 1e-4
    assert np.max(np.abs(r.density - dens)) <= dens_bar
    assert np.max(np.abs(r.charges - q)) <= 1e-6
    assert np.max(np.abs(r.dipole_debye - np.array(res.dipole[:]))) <= 1e-5
    assert np.max(np.abs(r.dipole_point_charge_debye - np.array(res.dipole_point[:]))) <= 1e-5
    if r.unrestricted:
        assert np.max(np.abs(r.spin_density - spin)) <= dens_bar
    return r, ref


FORMALDEHYDE = [["C", 0.0, 0.0, 0.0], ["O", 0.03, 0.0, 1.25], ["H", 0.95, 0.02, -0.55], ["H", -0.94, -0.03, -0.52]]
WATER = [["O", 0.0, 0.0, 0.0], ["H", 0.9584, 0.0, 0.0], ["H", -0.24, 0.9278, 0.0]]
CH4 = [["C", 0, 0, 0], ["H", 0.629, 0.629, 0.629], ["H", -0.629, -0.629, 0.629], ["H", -0.629, 0.629, -0.629],
       ["H", 0.629, -0.629, -0.629]]
DIMER = CH4 + [["C", 0, 0, 3.6], ["H", 0.629, 0.629, 4.229], ["H", -0.629, -0.629, 4.229],
               ["H", -0.629, 0.629, 2.971], ["H", 0.629, -0.629, 2.971]]


@pytest.mark.parametrize("method,atoms,mopac", [("pm7-ts", CH4, -6.85849), ("pm7-minus", WATER, -57.78417),
                                                ("pm7-hh", DIMER, -7.03415)])
def test_methods_match_oracle_and_mopac(eng, method, atoms, mopac):
    say(f"method {method}")
    r, _ = both(eng, atoms, method=method)
    assert abs(r.heat_of_formation_kcal - mopac) < 1e-3
    eng.load("pm7")


@pytest.mark.parametrize("name", ["glycine", "methyl_radical", "thiophene"])
def test_level_shift_matches_oracle(eng, name):
    say(f"level shift {name}")
    c = CASES[name]
    r, _ = both(eng, c["atoms"], c["charge"], c["multiplicity"], level_shift=2.0)
    assert r.converged


@pytest.mark.parametrize("accel,use_diis", [("cdiis", True), ("none", True), ("adiis-cdiis", False)])
def test_accelerators_match_oracle(eng, accel, use_diis):
    say(f"accelerator {accel} use_diis={use_diis}")
    for atoms in (WATER, FORMALDEHYDE):
        both(eng, atoms, accelerator=accel, use_diis=use_diis)


def test_all_accelerators_reach_one_energy(eng):
    # src/scf.rs:2201-2227: every accelerator reaches the same energy to 1e-6 eV
    mol = pm7amd.Molecule.from_symbols_angstrom(CASES["glycine"]["atoms"])
    e = [eng.run(mol, pm7amd.Pm7Options(accelerator=a, max_scf=500)).total_ev for a in ("adiis-cdiis", "cdiis", "none")]
    assert max(e) - min(e) <= 1e-6


def test_forced_uhf_singlet_equals_rhf(eng):
    say("forced UHF singlet")
    for name in ("water", "benzene", "formaldehyde"):
        c = CASES[name]
        mol = pm7amd.Molecule.from_symbols_angstrom(c["atoms"])
        rhf = eng.run(mol, pm7amd.Pm7Options())
        uhf = eng.run(mol, pm7amd.Pm7Options(reference="uhf"))
        assert uhf.unrestricted and abs(uhf.total_ev - rhf.total_ev) <= 1e-4  # src/scf.rs:2090-2112
        assert np.max(np.abs(uhf.spin_density)) <= 1e-8
        both(eng, c["atoms"], reference="uhf")


@pytest.mark.parametrize("atoms", [WATER, FORMALDEHYDE], ids=["water", "formaldehyde"])
def test_force_dpath_matches_sp_path_and_oracle(eng, atoms):
    say("force_dpath")
    mol = pm7amd.Molecule.from_symbols_angstrom(atoms)
    sp = eng.run(mol, pm7amd.Pm7Options())
    dp, _ = both(eng, atoms, force_dpath=True)
    assert abs(sp.total_ev - dp.total_ev) <= 1e-4  # src/scf.rs:1965-2031


def test_smooth_dispersion_matches_oracle(eng):
    say("smooth dispersion")
    s = 1.391 / 1.09
    methane = [["C", 0, 0, 0], ["H", 0.63, 0.63, 0.63], ["H", -0.63, -0.63, 0.63], ["H", 0.63, -0.63, -0.63],
               ["H", -0.63 * s, 0.63 * s, -0.63 * s]]
    smooth, _ = both(eng, methane, smooth_dispersion=True)
    step, _ = both(eng, methane)
    assert smooth.correction_kcal != step.correction_kcal
    both(eng, CASES["benzene_dimer_stacked"]["atoms"], smooth_dispersion=True)


@pytest.mark.parametrize("origin", ["centre-of-mass", "coordinates", "centre-of-charge"])
def test_ion_dipoles_match_oracle(eng, origin):
    say(f"ion dipoles {origin}")
    for name in ("ammonium", "hydroxide", "carbonate", "nitrate"):
        c = CASES[name]
        both(eng, c["atoms"], c["charge"], c["multiplicity"], dipole_origin=origin)


def test_seeded_density_matches_oracle(eng):
    say("initial density")
    c = CASES["glycine"]
    z, xyz = ob.atoms_to_arrays(c["atoms"])
    _, d0, *_ = ob.run(z, xyz)
    # a perturbed converged density: the run must start from it, not from the atomic guess
    rng = np.random.default_rng(5)
    noise = rng.normal(0.0, 1e-3, d0.shape)
    seed = d0 + 0.5 * (noise + noise.T)
    r, (res, *_) = both(eng, c["atoms"], initial_density=seed)
    cold = eng.run(pm7amd.Molecule.from_symbols_angstrom(c["atoms"]), pm7amd.Pm7Options())
    assert r.iterations < cold.iterations and r.iterations == res.iterations
    # open shell: alpha/beta split of the seeded total plus half the spin seed each
    m = CASES["methyl_radical"]
    zr, xr = ob.atoms_to_arrays(m["atoms"])
    _, dr, sr, *_ = ob.run(zr, xr, m["charge"], m["multiplicity"])
    both(eng, m["atoms"], m["charge"], m["multiplicity"], initial_density=dr, initial_spin_density=sr)


def test_spin_seed_breaks_closed_shell_symmetry(eng):
    say("spin seed")
    h2 = [["H", 0, 0, 0], ["H", 2.5, 0, 0]]
    mol = pm7amd.Molecule.from_symbols_angstrom(h2)
    rhf = eng.run(mol, pm7amd.Pm7Options())
    broken, _ = both(eng, h2, reference="uhf", initial_density=rhf.density, initial_spin_density=np.diag([0.5, -0.5]))
    assert broken.total_ev < rhf.total_ev - 0.1


def test_seeded_density_of_the_wrong_shape_is_refused(eng):
    mol = pm7amd.Molecule.from_symbols_angstrom(WATER)
    r = eng.run(mol, pm7amd.Pm7Options(initial_density=np.eye(5)), raise_on_error=False)
    assert r.status == 1 and "initial density" in r.message


def test_large_path_seeded_and_mixed_check(eng):
    """The grid-wide path (> 196 basis functions) takes a seeded density too, and a mixed run there
    reports the size of its final FP64 step."""
    say("large path seeded")
    atoms = W.water_cluster(40, 4)
    z, xyz = ob.atoms_to_arrays(atoms)
    res, dens, *_ = ob.run(z, xyz)
    mol = pm7amd.Molecule.from_symbols_angstrom(atoms)
    r = eng.run(mol, pm7amd.Pm7Options(initial_density=dens))
    assert r.iterations <= 3 and abs(r.total_ev - res.total_ev) <= 1e-6
    mx = eng.run(mol, pm7amd.Pm7Options(precision="mixed"))
    assert 0.0 <= mx.err_vs_fp64_ev <= 1e-7 and 0.0 <= mx.err_vs_fp64_density <= 1e-6
    assert abs(mx.total_ev - res.total_ev) <= 1e-6
    _, fixed, *_ = ob.run(z, xyz, e_tol=1e-12, p_tol=1e-10, max_scf=500)
    assert np.max(np.abs(mx.density - fixed)) <= 1e-7


def test_refusals_carry_their_payloads(eng):
    say("refusals")
    water = pm7amd.Molecule.from_symbols_angstrom(WATER)
    europium = pm7amd.Molecule.from_symbols_angstrom([["Eu", 0, 0, 0], ["F", 2.0, 0, 0], ["F", -2.0, 0, 0],
                                                      ["F", 0, 2.0, 0]])
    r = eng.run(europium, pm7amd.Pm7Options(), raise_on_error=False)
    assert r.status == 2 and r.missing_element == 63
    # more electrons of one spin than orbitals: H2 with charge -4 (6 electrons, 2 orbitals)
    h2 = pm7amd.Molecule.from_symbols_angstrom([["H", 0, 0, 0], ["H", 0.74, 0, 0]], -4.0)
    r = eng.run(h2, pm7amd.Pm7Options(charge=-4.0), raise_on_error=False)
    assert r.status == 1
    # the request server refuses it the same way
    with pm7amd.Server(eng, lanes=1) as srv:
        rid = srv.submit(h2, pm7amd.Pm7Options(charge=-4.0))
        _, rr, _ = srv.poll(rid, 10000)
        assert rr.status == 1
    # PM7-Sparkle is not offered; unknown codes are invalid; the tables must match the options
    lib = pm7amd.library()
    assert lib.pm7_params_load(eng._h, 2) == 4
    assert lib.pm7_params_load(eng._h, 99) == 1
    assert lib.pm7_params_load(eng._h, 0) == 0
    o = pm7amd.Pm7Options().to_c()
    o.method = 4  # PM7-HH options against loaded PM7 tables
    cm, keep = pm7amd.Engine._mol(water)
    res = (pm7amd._Result * 1)()
    assert lib.pm7_run(eng._h, C.byref(cm), C.byref(o), res) == 1
    lib.pm7_result_free(res)
    for code in (2, 99):
        o.method = code
        res = (pm7amd._Result * 1)()
        assert lib.pm7_run(eng._h, C.byref(cm), C.byref(o), res) == (4 if code == 2 else 1)
        lib.pm7_result_free(res)
    big = pm7amd.Molecule.from_symbols_angstrom(W.alkane_chain(32, 2))
    r = eng.run(big, pm7amd.Pm7Options(max_memory_mb=1), raise_on_error=False)
    assert r.status == 7 and r.budget_mb == pytest.approx(1.0) and r.needed_mb > r.budget_mb
