# SPDX-License-Identifier: GPL-3.0-or-later
"""GPU parity: the HIP engine against the CPU FP64 oracle and pm7-rs's golden vectors.

Bars (north star): |dE| <= 1e-6 eV, |dHf| <= 1e-4 kcal/mol, max |dP| <= 1e-7; stage-level
pieces (H_core, one Fock build, core-core) to 1e-9 or better since they involve no iteration.
"""

import json
import sys
from pathlib import Path

import os

import numpy as np
import pytest

import oracle_binding as ob
import pm7amd
from pm7amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = json.loads((Path(__file__).parent / "golden" / "reference_cases.json").read_text())["cases"]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = pm7amd.Engine(0)
    yield e
    e.close()


def mol_of(atoms, charge=0.0, mult=1):
    return pm7amd.Molecule.from_symbols_angstrom(atoms, charge, mult)


def oracle_of(atoms, charge=0.0, mult=1):
    z, xyz = ob.atoms_to_arrays(atoms)
    return ob.run(z, xyz, charge, mult)


@pytest.mark.parametrize("name", ["water", "formaldehyde", "ammonium", "hf_chain_8", "alkane_c20", "methyl_radical",
                                  "germane", "dimethylzinc", "calcium_difluoride", "tetramethyltin"])
def test_core_hamiltonian_matches_oracle(eng, name):
    say(f"hcore {name}")
    c = CASES[name]
    h = eng.core_hamiltonian(mol_of(c["atoms"], c["charge"], c["multiplicity"]))
    z, xyz = ob.atoms_to_arrays(c["atoms"])
    ho = ob.hcore(z, xyz)
    assert h.shape == ho.shape
    assert np.max(np.abs(h - ho)) <= 1e-9


# Molecules with a d-bearing atom: every pair through the MNDO/d kernels (spd rotation, multipole
# integrals, d-balance for transition metals, spd overlaps).
D_PATH = ["hydrogen_sulfide", "sulfur_hexafluoride", "thiophene", "titanium_tetrachloride", "permanganate",
          "chromium_hexacarbonyl", "ferrocene", "tetrachloroplatinate", "titanocene_dichloride", "iodomethane"]


@pytest.mark.parametrize("name", D_PATH)
def test_core_hamiltonian_d_path_matches_oracle(eng, name):
    say(f"hcore d {name}")
    c = CASES[name]
    h = eng.core_hamiltonian(mol_of(c["atoms"], c["charge"], c["multiplicity"]))
    z, xyz = ob.atoms_to_arrays(c["atoms"])
    ho = ob.hcore(z, xyz)
    assert h.shape == ho.shape
    assert np.max(np.abs(h - ho)) <= 1e-9


@pytest.mark.parametrize("name", ["hydrogen_sulfide", "sulfur_hexafluoride", "ferrocene", "permanganate"])
def test_fock_build_d_path_matches_oracle(eng, name):
    say(f"fock d {name}")
    c = CASES[name]
    z, xyz = ob.atoms_to_arrays(c["atoms"])
    n = ob.nao(z)
    rng = np.random.default_rng(11)
    a = rng.normal(size=(n, n)) * 0.3
    pt = a + a.T
    b = rng.normal(size=(n, n)) * 0.2
    ps = b + b.T
    f = eng.fock_spin(mol_of(c["atoms"], c["charge"], c["multiplicity"]), pt, ps)
    fo = ob.fock(z, xyz, pt, ps)
    assert np.max(np.abs(f - fo)) <= 1e-9


@pytest.mark.parametrize("name", ["thiophene", "ferrocene", "iron_pentacarbonyl", "chlorine_dioxide",
                                  "tetrachloroplatinate", "perrhenate"])
def test_engine_d_path_matches_oracle_density(eng, name):
    say(f"scf d {name}")
    c = CASES[name]
    r = eng.run(mol_of(c["atoms"], c["charge"], c["multiplicity"]),
                pm7amd.Pm7Options(charge=c["charge"], multiplicity=c["multiplicity"]))
    res, dens, spin, eps, q = oracle_of(c["atoms"], c["charge"], c["multiplicity"])
    assert abs(r.total_ev - res.total_ev) <= 1e-6
    assert abs(r.heat_of_formation_kcal - res.hof_kcal) <= 1e-4
    assert np.max(np.abs(r.density - dens)) <= 1e-7
    assert np.max(np.abs(np.array(r.dipole_pd_hybrid_debye) - np.array(res.dipole_pd[:]))) <= 1e-5
    if r.unrestricted:
        assert np.max(np.abs(r.spin_density - spin)) <= 1e-7


@pytest.mark.parametrize("name", ["water", "formic_acid_dimer", "alkane_c20", "water_wire_8"])
def test_fock_build_matches_oracle(eng, name):
    say(f"fock {name}")
    c = CASES[name]
    z, xyz = ob.atoms_to_arrays(c["atoms"])
    n = ob.nao(z)
    rng = np.random.default_rng(7)
    a = rng.normal(size=(n, n)) * 0.3
    pt = a + a.T
    b = rng.normal(size=(n, n)) * 0.2
    ps = b + b.T
    f = eng.fock_spin(mol_of(c["atoms"]), pt, ps)
    fo = ob.fock(z, xyz, pt, ps)
    assert np.max(np.abs(f - fo)) <= 1e-9


@pytest.mark.parametrize("name", ["water", "ammonium", "alkane_c30", "boron_trifluoride"])
def test_core_core_matches_oracle(eng, name):
    say(f"core-core {name}")
    c = CASES[name]
    z, xyz = ob.atoms_to_arrays(c["atoms"])
    e = eng.core_core_energy(mol_of(c["atoms"]))
    assert abs(e - ob.core_core(z, xyz)) <= 1e-9


def check_golden(k, r, bad):
    c = CASES[k]
    de = r.total_ev - c["energy_ev"]
    dh = r.heat_of_formation_kcal - c["heat_of_formation_kcal"]
    dq = float(np.max(np.abs(r.charges - np.array(c["charges"])))) if r.status == 0 else float("inf")
    if r.status != 0 or not r.converged or abs(de) > 1e-6 or abs(dh) > 1e-4 or dq > 1e-6:
        bad.append((k, r.status, de, dh, dq, r.fp32_iterations, r.message))


def test_golden_cases_in_one_batch(eng):
    """Every golden case, FP64, in one call, at the full bar."""
    names = sorted(CASES)
    say(f"golden batch of {len(names)}")
    mols = [mol_of(CASES[k]["atoms"], CASES[k]["charge"], CASES[k]["multiplicity"]) for k in names]
    opts = [pm7amd.Pm7Options(charge=CASES[k]["charge"], multiplicity=CASES[k]["multiplicity"]) for k in names]
    res = eng.run_batch(mols, opts)
    bad = []
    for k, r in zip(names, res):
        check_golden(k, r, bad)
        assert r.err_vs_fp64_ev == 0.0 and r.err_vs_fp64_density == 0.0
    for b in bad:
        say(f"  {b}")
    assert not bad, bad


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_degenerate_open_shell_reaches_the_reference_state(eng, precision):
    """O2(2+) triplet: the atomic guess puts both spin frontiers inside exactly degenerate pi sets,
    one alpha pi* and one beta pi electron. The eigensolver orders a degenerate set by slot,
    ascending for alpha and descending for beta on the first (cold) solve, which puts the two
    electrons on different axes: the state pm7-rs and MOPAC report (813.98294 kcal/mol), 0.99 eV
    below the one where both take the same axis."""
    say(f"O2 2+ triplet {precision}")
    c = CASES["oxygen_dication_triplet"]
    r = eng.run(mol_of(c["atoms"], c["charge"], c["multiplicity"]),
                pm7amd.Pm7Options(charge=c["charge"], multiplicity=c["multiplicity"], precision=precision))
    bad = []
    check_golden("oxygen_dication_triplet", r, bad)
    assert not bad, bad
    assert abs(r.homo_ev - c["homo_ev"]) <= 1e-5 and abs(r.lumo_ev - c["lumo_ev"]) <= 1e-5
    assert abs(r.homo_ev_beta - c["homo_ev_beta"]) <= 1e-5 and abs(r.lumo_ev_beta - c["lumo_ev_beta"]) <= 1e-5
    # The reference's state and this one differ by a quarter turn about the molecular axis (which of
    # y and z carries the alpha pi* electron), so the densities are compared through invariants of
    # that rotation: their eigenvalues.
    z, xyz = ob.atoms_to_arrays(c["atoms"])
    res, dens, spin, eps, q = ob.run(z, xyz, c["charge"], c["multiplicity"],
                                     **({} if precision == "fp64" else dict(e_tol=1e-12, p_tol=1e-10, max_scf=500)))
    for a, b in ((r.density, dens), (r.spin_density, spin)):
        assert np.max(np.abs(np.linalg.eigvalsh(a) - np.linalg.eigvalsh(b))) <= 1e-7


@pytest.mark.parametrize("name", ["water", "benzene", "glycine", "alkane_c30", "water_wire_12", "nitrogen_dioxide",
                                  "allyl_radical", "carbonate", "hydrogen_telluride", "tetramethyltin"])
def test_engine_matches_oracle_density(eng, name):
    say(f"scf {name}")
    c = CASES[name]
    r = eng.run(mol_of(c["atoms"], c["charge"], c["multiplicity"]),
                pm7amd.Pm7Options(charge=c["charge"], multiplicity=c["multiplicity"]))
    res, dens, spin, eps, q = oracle_of(c["atoms"], c["charge"], c["multiplicity"])
    assert abs(r.total_ev - res.total_ev) <= 1e-6
    assert abs(r.heat_of_formation_kcal - res.hof_kcal) <= 1e-4
    assert np.max(np.abs(r.density - dens)) <= 1e-7
    assert np.max(np.abs(r.mo_energies - eps)) <= 1e-5
    if r.unrestricted:
        assert np.max(np.abs(r.spin_density - spin)) <= 1e-7


def tight(**kw):
    """FP64 converged far past the default stopping test: the fixed point itself."""
    return pm7amd.Pm7Options(e_tol=1e-12, p_tol=1e-10, max_scf=500, **kw)


# Cases whose mixed run settles a state related to the reference's by a symmetry operation of the
# nuclear frame (same energy, same heat of formation): O2 2+ (which of y and z carries the odd
# electrons), the Si atom triplet (which two p orbitals hold the pair) and CfF3 (which of the three
# equivalent F atoms carries the larger charge in the symmetry-broken state). Densities are compared
# through their eigenvalues and charges as sorted lists.
SYMMETRY_EQUIVALENT = {"oxygen_dication_triplet", "silicon_atom_triplet", "californium_trifluoride"}


def test_mixed_precision_golden_batch_within_bars(eng):
    """PRECISION_MIXED (FP32 G(P) and P until the density change is below 1e-3, then FP64, with
    occupied-virtual rotations near the fixed point, a tail converged 100x past p_tol and a final
    full FP64 Roothaan step) on every golden case: the energy and heat-of-formation bars against
    pm7-rs, and the density within 1e-7 of the FP64 fixed point (an FP64 run converged to
    p_tol 1e-10). A default FP64 run stops where its own path crosses p_tol, up to ~4e-6 from that
    point on the long alkanes, and the reference's charges carry that stopping error, so the mixed
    charges are held to 1e-5 against them. Each result's stated error against FP64
    (err_vs_fp64_*) is small and bounds the actual one."""
    names = sorted(CASES)
    say(f"mixed golden batch of {len(names)}")
    mols = [mol_of(CASES[k]["atoms"], CASES[k]["charge"], CASES[k]["multiplicity"]) for k in names]
    mx = eng.run_batch(mols, [pm7amd.Pm7Options(charge=CASES[k]["charge"], multiplicity=CASES[k]["multiplicity"],
                                                 precision="mixed") for k in names])
    fix = eng.run_batch(mols, [tight(charge=CASES[k]["charge"], multiplicity=CASES[k]["multiplicity"]) for k in names])
    bad = []
    worst = 0.0
    for k, r, f in zip(names, mx, fix):
        c = CASES[k]
        de = r.total_ev - c["energy_ev"]
        dh = r.heat_of_formation_kcal - c["heat_of_formation_kcal"]
        ok = r.status == 0 and r.converged
        sym = k in SYMMETRY_EQUIVALENT
        qr, qc = (np.sort(r.charges), np.sort(c["charges"])) if sym else (r.charges, np.array(c["charges"]))
        dq = float(np.max(np.abs(qr - qc))) if ok else float("inf")
        if not ok or abs(de) > 1e-6 or abs(dh) > 1e-4 or dq > 1e-5:
            bad.append((k, r.status, de, dh, dq, r.fp32_iterations, r.message))
            continue
        if sym:
            dp = float(np.max(np.abs(np.linalg.eigvalsh(r.density) - np.linalg.eigvalsh(f.density))))
        else:
            dp = float(np.max(np.abs(r.density - f.density)))
        worst = max(worst, dp)
        if dp > 1e-7 or not (0.0 <= r.err_vs_fp64_ev <= 1e-6) or not (0.0 <= r.err_vs_fp64_density <= 1e-6):
            bad.append((k, "density vs fp64 fixed point", dp, r.err_vs_fp64_ev, r.err_vs_fp64_density))
    say(f"mixed golden: worst |dP| vs FP64 fixed point {worst:.2e}")
    for b in bad:
        say(f"  {b}")
    assert not bad, bad
    assert sum(r.fp32_iterations for r in mx) > 0


def test_benchmark_batch_fp64_matches_oracle_everywhere(eng):
    """The whole 256-molecule organic benchmark batch (seed 3), FP64, each molecule against the
    oracle at the full bar: the eigensolver's stopping test must leave every density within 1e-7,
    not only the golden cases'."""
    batch = W.organic_batch(256, 3)
    say("benchmark batch of 256 vs oracle")
    res = eng.run_batch([mol_of(a) for a in batch], pm7amd.Pm7Options())
    worst = (0.0, 0.0, 0.0)
    for a, r in zip(batch, res):
        ref, dens, *_ = oracle_of(a)
        assert r.status == 0 and r.converged
        d = (abs(r.total_ev - ref.total_ev), abs(r.heat_of_formation_kcal - ref.hof_kcal),
             float(np.max(np.abs(r.density - dens))))
        worst = tuple(max(x, y) for x, y in zip(worst, d))
    say(f"worst |dE| {worst[0]:.2e} eV, |dHf| {worst[1]:.2e} kcal/mol, |dP| {worst[2]:.2e}")
    assert worst[0] <= 1e-6 and worst[1] <= 1e-4 and worst[2] <= 1e-7


def test_mixed_precision_organic_batch_vs_fp64(eng):
    """Mixed against FP64 on 48 benchmark molecules: energies and heats of formation against the
    default FP64 run, densities against the FP64 fixed point (the tightly converged oracle)."""
    batch = W.organic_batch(48, 3)
    say("mixed organic batch of 48")
    mols = [mol_of(a) for a in batch]
    r64 = eng.run_batch(mols, pm7amd.Pm7Options())
    rmx = eng.run_batch(mols, pm7amd.Pm7Options(precision="mixed"))
    worst = 0.0
    for atoms, a, b in zip(batch, r64, rmx):
        assert a.status == 0 and b.status == 0 and b.converged
        assert abs(a.total_ev - b.total_ev) <= 1e-6
        assert abs(a.heat_of_formation_kcal - b.heat_of_formation_kcal) <= 1e-4
        z, xyz = ob.atoms_to_arrays(atoms)
        _, dens, *_ = ob.run(z, xyz, e_tol=1e-12, p_tol=1e-10, max_scf=500)
        worst = max(worst, float(np.max(np.abs(b.density - dens))))
        assert b.fp32_iterations > 0
        assert b.err_vs_fp64_density <= 1e-6
    say(f"mixed organic: worst |dP| vs the FP64 fixed point {worst:.2e}")
    assert worst <= 1e-7


@pytest.mark.parametrize("n_water,precision", [(40, "fp64"), (40, "mixed"), (70, "fp64")])
def test_large_path_water_cluster_matches_oracle(eng, n_water, precision):
    """Basis sizes above 196 run the grid-wide path (rocSOLVER dsyevd, rocBLAS products, grid Fock
    kernels): same bars against the oracle as the megakernel, FP64 and mixed."""
    atoms = W.water_cluster(n_water, 4)
    say(f"large path: {n_water} waters ({6 * n_water} AOs), {precision}")
    r = eng.run(mol_of(atoms), pm7amd.Pm7Options(precision=precision))
    res, dens, *_ = oracle_of(atoms)
    assert r.converged and r.density.shape == (6 * n_water, 6 * n_water)
    assert abs(r.total_ev - res.total_ev) <= 1e-6
    assert abs(r.heat_of_formation_kcal - res.hof_kcal) <= 1e-4
    if precision == "mixed":  # another path to the fixed point: against the tightly converged oracle
        z, xyz = ob.atoms_to_arrays(atoms)
        _, dens, *_ = ob.run(z, xyz, e_tol=1e-12, p_tol=1e-10, max_scf=500)
    assert np.max(np.abs(r.density - dens)) <= 1e-7
    if precision == "mixed":
        assert r.fp32_iterations > 0
    # orbitals: orthonormal columns that reproduce the density
    c = r.mo_coeff
    assert np.max(np.abs(c.T @ c - np.eye(c.shape[0]))) <= 1e-9
    occ = c[:, : r.n_occ]
    assert np.max(np.abs(2.0 * occ @ occ.T - r.density)) <= 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_large_path_integral_streams_equal_pair_kernels(eng, precision):
    """The per-atom integral streams of the closed-shell s/p large path (one workgroup per atom) and
    the element and pair kernels (PM7_AMD_LARGE_STREAMS=0) are two orders of the same sums: same
    SCF to well inside the parity bars, and the streamed run is reproducible bit for bit."""
    atoms = W.water_cluster(40, 4)
    say(f"large path streams vs pair kernels, {precision}")
    opts = pm7amd.Pm7Options(precision=precision)
    a = eng.run(mol_of(atoms), opts)
    b = eng.run(mol_of(atoms), opts)
    old = os.environ.get("PM7_AMD_LARGE_STREAMS")
    os.environ["PM7_AMD_LARGE_STREAMS"] = "0"
    try:
        c = eng.run(mol_of(atoms), opts)
    finally:
        if old is None:
            os.environ.pop("PM7_AMD_LARGE_STREAMS", None)
        else:
            os.environ["PM7_AMD_LARGE_STREAMS"] = old
    assert a.converged and c.converged
    assert a.total_ev == b.total_ev and np.array_equal(a.density, b.density)
    assert abs(a.total_ev - c.total_ev) <= 1e-8
    assert np.max(np.abs(a.density - c.density)) <= 1e-8


def test_large_group_matches_single_runs(eng):
    """Large-class molecules (197-512 basis functions) of one batch run as a lockstep group with
    batched eigensolves; each result equals the molecule's own run to the parity bars, and one of
    them matches the oracle."""
    rng = W.Rng(21)
    atoms = [W.hetero_chain(rng, nc, het) for nc, het in ((34, "Cl"), (40, "S"), (45, "P"), (52, "Cl"))]
    say("large group of %d" % len(atoms))
    mols = [mol_of(a) for a in atoms]
    group = eng.run_batch(mols, [pm7amd.Pm7Options()] * len(mols))
    for m, g in zip(mols, group):
        alone = eng.run(m, pm7amd.Pm7Options())
        assert g.status == 0 and g.converged
        assert abs(g.total_ev - alone.total_ev) <= 1e-6
        assert abs(g.heat_of_formation_kcal - alone.heat_of_formation_kcal) <= 1e-4
        assert np.max(np.abs(g.density - alone.density)) <= 1e-7
    res, dens, spin, eps, q = oracle_of(atoms[0])
    assert abs(group[0].total_ev - res.total_ev) <= 1e-6
    assert abs(group[0].heat_of_formation_kcal - res.hof_kcal) <= 1e-4
    assert np.max(np.abs(group[0].density - dens)) <= 1e-7


def test_call_after_a_grouped_large_batch_on_a_fresh_context():
    """A context whose first call runs only grouped large-class SCFs (two molecules of 197-512 basis
    functions) and whose next call is an ordinary batch: the timing of the first call must leave no
    error behind for the second (the large class's timing events were once recorded only on the
    one-at-a-time branch)."""
    rng = W.Rng(21)
    atoms = [W.hetero_chain(rng, nc, het) for nc, het in ((34, "Cl"), (40, "S"))]
    say("fresh context: grouped large batch, then a small batch")
    with pm7amd.Engine(0) as e:
        res = e.run_batch([mol_of(a) for a in atoms], pm7amd.Pm7Options(), raise_on_error=True)
        assert all(r.converged for r in res)
        again = e.run_batch([mol_of(W.water()), mol_of(CASES["benzene"]["atoms"])], pm7amd.Pm7Options(),
                            raise_on_error=True)
        assert abs(again[0].total_ev - oracle_of(W.water())[0].total_ev) <= 1e-6


def test_config1_water(eng):
    say("config 1 water")
    atoms = W.water()
    r = eng.run(mol_of(atoms))
    res, dens, *_ = oracle_of(atoms)
    assert r.converged and res.status == 0
    assert abs(r.total_ev - res.total_ev) <= 1e-6
    assert abs(r.heat_of_formation_kcal - res.hof_kcal) <= 1e-4
    assert np.max(np.abs(r.density - dens)) <= 1e-7


def test_config2_alkane_c32(eng):
    say("config 2 C32H66")
    atoms = W.alkane_chain(32, 2)
    r = eng.run(mol_of(atoms))
    assert r.density.shape == (194, 194)
    res, dens, *_ = oracle_of(atoms)
    assert abs(r.total_ev - res.total_ev) <= 1e-6
    assert abs(r.heat_of_formation_kcal - res.hof_kcal) <= 1e-4
    assert np.max(np.abs(r.density - dens)) <= 1e-7


def test_batch_result_is_the_same_alone_and_in_the_mix(eng):
    mols_atoms = W.organic_batch(24, 11)
    say("batch vs alone, 24 molecules")
    mols = [mol_of(a) for a in mols_atoms]
    together = eng.run_batch(mols)
    for i in (0, 5, 17, 23):
        alone = eng.run(mols[i])
        assert alone.total_ev == together[i].total_ev
        assert np.array_equal(alone.density, together[i].density)
    for i, a in enumerate(mols_atoms[:6]):
        res, dens, *_ = oracle_of(a)
        assert abs(together[i].total_ev - res.total_ev) <= 1e-6
        assert np.max(np.abs(together[i].density - dens)) <= 1e-7


def test_error_paths_are_reported_per_molecule(eng):
    say("error paths")
    water = mol_of(W.water())
    coincident = mol_of([["O", 0, 0, 0], ["H", 0, 0, 0], ["H", 0.96, 0, 0]])
    sulfur = mol_of(CASES["hydrogen_sulfide"]["atoms"])
    odd = mol_of([["O", 0, 0, 0], ["H", 0.97, 0, 0]])
    europium = mol_of([["Eu", 0, 0, 0], ["F", 2.0, 0, 0]])  # no PM7 orbital parameters for Eu
    res = eng.run_batch([water, coincident, sulfur, odd, water, europium],
                        [pm7amd.Pm7Options(), pm7amd.Pm7Options(), pm7amd.Pm7Options(), pm7amd.Pm7Options(),
                         pm7amd.Pm7Options(max_scf=1), pm7amd.Pm7Options()])
    assert [r.status for r in res] == [0, 1, 0, 1, 6, 2]
    assert res[5].missing_element == 63
    assert res[2].total_ev == pytest.approx(CASES["hydrogen_sulfide"]["energy_ev"], abs=1e-6)
    assert np.isfinite(res[4].total_ev) and res[4].iterations == 1
    assert res[0].total_ev == pytest.approx(oracle_of(W.water())[0].total_ev, abs=1e-6)


def test_insufficient_memory_is_refused_before_allocation(eng):
    say("memory budget")
    big = mol_of(W.alkane_chain(32, 2))
    r = eng.run(big, pm7amd.Pm7Options(max_memory_mb=1), raise_on_error=False)
    assert r.status == 7
