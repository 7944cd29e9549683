# SPDX-License-Identifier: GPL-3.0-or-later
"""Pin the oracle's option paths (the methods, force_dpath, smooth_dispersion, dipole_origin and
the seeded densities) against what pm7-rs's own tests hold for them:

* PM7-TS / PM7- / PM7-HH heats of formation against MOPAC (tests/molecules.rs:322-344, 1e-3);
* the dipole breakdown of water, H2S, SO2 and PF3 against MOPAC (tests/dipole.rs:46-130, 2e-3 D),
  a charged species recentred, a neutral one bit-identical across origins (tests/dipole.rs:160-210);
* force_dpath reproducing the sp energy of water and formaldehyde (src/scf.rs:1965-2031, 1e-4 eV);
* the smooth dispersion: no step across the C-H counting threshold, the discrete value away from it
  (tests/dispersion_smooth.rs:90-155);
* the PM7-HH polynomial's branches (src/hh_rep.rs:97-108);
* a seeded density: the converged density restarts at its own fixed point, and a spin seed breaks
  the spin symmetry a closed-shell UHF run otherwise keeps (src/scf.rs:161-170, 1336-1378).
"""

import numpy as np
import pytest

import oracle_binding as ob

A = ob.ANGSTROM_TO_BOHR
TIGHT = dict(e_tol=1e-12, p_tol=1e-10, max_scf=500)


def arrays(atoms):
    return ob.atoms_to_arrays(atoms)


CH4 = [["C", 0, 0, 0], ["H", 0.629, 0.629, 0.629], ["H", -0.629, -0.629, 0.629], ["H", -0.629, 0.629, -0.629],
       ["H", 0.629, -0.629, -0.629]]
H2O_M = [["O", 0.0, 0.0, 0.0], ["H", 0.9584, 0.0, 0.0], ["H", -0.24, 0.9278, 0.0]]
DIMER = CH4 + [["C", 0, 0, 3.6], ["H", 0.629, 0.629, 4.229], ["H", -0.629, -0.629, 4.229],
               ["H", -0.629, 0.629, 2.971], ["H", 0.629, -0.629, 2.971]]


@pytest.mark.parametrize("method,atoms,mopac", [(1, CH4, -6.85849), (3, H2O_M, -57.78417), (4, DIMER, -7.03415)])
def test_methods_match_mopac(method, atoms, mopac):
    z, xyz = arrays(atoms)
    res, *_ = ob.run(z, xyz, method=method)
    assert res.status == 0
    assert abs(res.hof_kcal - mopac) < 1e-3


def test_hh_repulsion_is_the_difference_between_pm7_hh_and_pm7():
    z, xyz = arrays(DIMER)
    pm7, *_ = ob.run(z, xyz, method=0)
    hh, *_ = ob.run(z, xyz, method=4)
    assert abs((hh.hof_kcal - pm7.hof_kcal) - ob.corrections2(z, xyz)[1]) < 1e-9
    assert abs(hh.electronic_ev - pm7.electronic_ev) < 1e-12  # a post-SCF term only


def test_hh_polynomial_branches():
    # plateau below 1 A, exponential tail beyond 1.5 A (src/hh_rep.rs:97-108)
    z = np.array([1, 1], dtype=np.int32)
    for r in (0.8, 1.0):
        xyz = np.array([[0, 0, 0], [r * A, 0, 0]], dtype=float)
        assert abs(ob.corrections2(z, xyz)[1] - 25.46293603147693) < 1e-9
    xyz = np.array([[0, 0, 0], [2.0 * A, 0, 0]], dtype=float)
    assert abs(ob.corrections2(z, xyz)[1] - 118.7326 * np.exp(-1.53965 * 2.0 ** 1.72905)) < 1e-9
    # the polynomial piece is positive and joins the plateau within the fit's accuracy
    xyz = np.array([[0, 0, 0], [1.25 * A, 0, 0]], dtype=float)
    assert ob.corrections2(z, xyz)[1] > 0.0
    z3 = np.array([1, 1, 8], dtype=np.int32)  # only H-H pairs count
    xyz3 = np.array([[0, 0, 0], [2.0 * A, 0, 0], [1.0 * A, 0.5 * A, 0]], dtype=float)
    assert abs(ob.corrections2(z3, xyz3)[1] - 118.7326 * np.exp(-1.53965 * 2.0 ** 1.72905)) < 1e-9


DIPOLE_CASES = [
    ("water", [["O", 0, 0, 0], ["H", 0.96, 0, 0], ["H", -0.24, 0.93, 0]],
     [1.115, 1.441, 0.0], [0.197, 0.254, 0.0], [1.312, 1.695, 0.0]),
    ("hydrogen_sulfide", [["S", 0, 0, 0], ["H", 1.34, 0, 0], ["H", -0.35, 1.29, 0]],
     [0.824, 1.069, 0.0], [0.285, 0.374, 0.0], [1.109, 1.443, 0.0]),
    ("sulfur_dioxide", [["S", 0, 0, 0], ["O", 1.24, 0, 0.72], ["O", -1.24, 0, 0.72]],
     [0.0, 0.0, -4.764], [0.0, 0.0, 1.472], [0.0, 0.0, -3.292]),
    ("phosphorus_trifluoride", [["P", 0, 0, 0], ["F", 1.24, 0, 0.51], ["F", -0.62, 1.07, 0.51],
                                ["F", -0.62, -1.07, 0.51]],
     [-0.004, 0.0, -1.677], [-0.001, 0.0, 0.883], [-0.005, 0.0, -0.795]),
]


@pytest.mark.parametrize("name,atoms,pc,hyb,total", DIPOLE_CASES, ids=[c[0] for c in DIPOLE_CASES])
def test_dipole_breakdown_matches_mopac(name, atoms, pc, hyb, total):
    z, xyz = arrays(atoms)
    res, *_ = ob.run(z, xyz, e_tol=1e-12, p_tol=1e-9, max_scf=800)
    assert res.status == 0
    got_h = np.array(res.dipole_sp[:]) + np.array(res.dipole_pd[:])
    assert np.max(np.abs(np.array(res.dipole_point[:]) - pc)) < 2e-3
    assert np.max(np.abs(got_h - hyb)) < 2e-3
    assert np.max(np.abs(np.array(res.dipole[:]) - total)) < 2e-3


def test_dipole_origin_of_an_ion_and_of_a_neutral_molecule():
    z, xyz = arrays([["O", 0, 0, 0], ["H", 0.96, 0, 0]])
    com, *_, q = ob.run(z, xyz, charge=-1.0, **TIGHT)
    raw, *_ = ob.run(z, xyz, charge=-1.0, dipole_origin=1, **TIGHT)
    coc, *_ = ob.run(z, xyz, charge=-1.0, dipole_origin=2, **TIGHT)
    assert abs(np.linalg.norm(raw.dipole[:]) - np.linalg.norm(com.dipole[:])) > 0.1
    # closed form: mu_pc(c) = sum q_A (R_A - c) = mu_pc(0) - (sum q) c, in Debye
    debye = 2.541746473  # e * Bohr in Debye
    assert np.max(np.abs(np.array(raw.dipole_point[:]) - debye * (q[:, None] * xyz).sum(0))) < 1e-9
    masses = np.array([15.9994, 1.0079])
    c_com = (masses[:, None] * xyz).sum(0) / masses.sum()
    zc = np.array([6.0, 1.0])  # core charges of O and H
    c_coc = (zc[:, None] * xyz).sum(0) / zc.sum()
    for res, c in ((com, c_com), (coc, c_coc)):
        expect = np.array(raw.dipole_point[:]) - (np.sum(q) * c) * debye
        assert np.max(np.abs(np.array(res.dipole_point[:]) - expect)) < 1e-9
        # the hybrid parts do not move
        assert np.max(np.abs(np.array(res.dipole_sp[:]) - np.array(raw.dipole_sp[:]))) < 1e-12
    zw, xw = arrays(DIPOLE_CASES[0][1])
    ref, *_ = ob.run(zw, xw, **TIGHT)
    for origin in (0, 1, 2):
        other, *_ = ob.run(zw, xw, dipole_origin=origin, **TIGHT)
        assert list(other.dipole[:]) == list(ref.dipole[:])


@pytest.mark.parametrize("atoms", [
    [["O", 0.0, 0.0, 0.0], ["H", 0.9584, 0.0, 0.0], ["H", -0.24, 0.9278, 0.0]],
    [["C", 0.0, 0.0, 0.0], ["O", 0.03, 0.0, 1.25], ["H", 0.95, 0.02, -0.55], ["H", -0.94, -0.03, -0.52]],
], ids=["water", "formaldehyde"])
def test_force_dpath_reproduces_the_sp_energy(atoms):
    z, xyz = arrays(atoms)
    sp, *_ = ob.run(z, xyz)
    dp, *_ = ob.run(z, xyz, force_dpath=1)
    assert sp.status == 0 and dp.status == 0
    assert abs(sp.total_ev - dp.total_ev) < 1e-4


def methane(d):
    s = d / 1.09
    return np.array([[0, 0, 0], [0.63, 0.63, 0.63], [-0.63, -0.63, 0.63], [0.63, -0.63, -0.63],
                     [-0.63 * s, 0.63 * s, -0.63 * s]]) * A


def test_smooth_dispersion_has_no_step_and_matches_away_from_it():
    z = np.array([6, 1, 1, 1, 1], dtype=np.int32)
    centre = 1.3910

    def second(h, smooth):
        f = (lambda x: ob.corrections2(z, methane(x))[0]) if smooth else (lambda x: ob.corrections(z, methane(x))[0])
        return abs(f(centre + h) - 2.0 * f(centre) + f(centre - h))

    coarse, fine = second(0.0040, False), second(0.0020, False)
    assert coarse > 1e-4 and abs(fine / coarse - 1.0) < 0.05  # the discrete path steps
    prev = second(0.0040, True)
    for h in (0.0020, 0.0010, 0.0005):
        v = second(h, True)
        assert 3.6 <= prev / v <= 4.4  # curvature, not a step
        prev = v
    # relaxed methane: both paths agree (kcal/mol; 1e-6 eV = 2.3e-5 kcal/mol)
    x = methane(1.09)
    assert abs(ob.corrections2(z, x)[0] - ob.corrections(z, x)[0]) < 2.3e-5


def test_seeded_density_restarts_at_its_fixed_point():
    z, xyz = arrays(H2O_M)
    ref, dens, *_ = ob.run(z, xyz, **TIGHT)
    again, dens2, *_ = ob.run(z, xyz, initial_density=dens, **TIGHT)
    assert again.iterations <= 3
    assert abs(again.total_ev - ref.total_ev) < 1e-9
    assert np.max(np.abs(dens2 - dens)) < 1e-8
    # UHF splits a seeded total density n_alpha/n : n_beta/n and adds half the spin seed to each
    # spin (P_alpha = P n_a/n + S/2, not (P + S)/2), so the converged pair is not a fixed point of
    # the seed; the run still reaches the same state.
    zr, xr = arrays([["C", 0, 0, 0], ["H", 1.08, 0, 0], ["H", -0.54, 0.935, 0], ["H", -0.54, -0.935, 0]])
    r0, d0, s0, *_ = ob.run(zr, xr, multiplicity=2, **TIGHT)
    r1, d1, s1, *_ = ob.run(zr, xr, multiplicity=2, initial_density=d0, initial_spin_density=s0, **TIGHT)
    assert r1.status == 0 and abs(r1.total_ev - r0.total_ev) < 1e-8
    assert np.max(np.abs(s1 - s0)) < 1e-6


def test_spin_seed_breaks_the_symmetry_of_a_closed_shell_uhf_run():
    # stretched H2: closed-shell UHF keeps P_alpha = P_beta unless seeded with a spin density
    z = np.array([1, 1], dtype=np.int32)
    xyz = np.array([[0, 0, 0], [2.5 * A, 0, 0]], dtype=float)
    rhf, dens, *_ = ob.run(z, xyz, **TIGHT)
    uhf, *_ = ob.run(z, xyz, reference=2, **TIGHT)
    assert abs(uhf.total_ev - rhf.total_ev) < 1e-6  # forced UHF singlet = RHF (src/scf.rs:2090-2112)
    seed = np.diag([0.5, -0.5])
    broken, d2, spin, *_ = ob.run(z, xyz, reference=2, initial_density=dens, initial_spin_density=seed, **TIGHT)
    assert broken.status == 0
    assert broken.total_ev < rhf.total_ev - 0.1
    assert np.max(np.abs(spin)) > 0.5
