# SPDX-License-Identifier: GPL-3.0-or-later
"""Two-electron response Fock of many densities against one resident integral set
(pm7_response_fock_batch; pm7-rs cphf_ov's kernel, src/hessian.rs:1074-1081: build_fock_x(X) - H_core,
src/fock.rs:665-684). Held to the CPU oracle's Fock build minus its core Hamiltonian, to the
linearity every response kernel has (G(aX + bY) = aG(X) + bG(Y), at any size), to the converged
SCF (H + G(P) commutes with P), and to batch/alone identity. The open-shell kernel
(pm7_response_fock_spin_batch; analytic_hessian_uhf, src/hessian.rs:1850-1880) is held to the same
oracle and reduces to the closed-shell one for equal alpha and beta halves."""

from __future__ import annotations

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


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = pm7amd.Engine(0)
    yield e
    e.close()


def response_densities(r, k, seed):
    """k CPHF-shaped densities C_v U C_o^T + transpose (pm7-rs ao_response_density,
    src/hessian.rs:1026-1041, weight 2) from the converged orbitals."""
    rng = np.random.default_rng(seed)
    c = r.mo_coeff
    co, cv = c[:, : r.n_occ], c[:, r.n_occ:]
    out = []
    for _ in range(k):
        u = 2.0 * rng.normal(size=(cv.shape[1], co.shape[1])) * 0.1
        b = cv @ u @ co.T
        out.append(b + b.T)
    return np.array(out)


@pytest.mark.parametrize("name", ["water", "formic_acid_dimer", "alkane_c20", "thiophene", "germane", "ferrocene"])
def test_response_fock_matches_oracle(eng, name):
    say(f"response fock {name}")
    c = CASES[name]
    mol = pm7amd.Molecule.from_symbols_angstrom(c["atoms"], c["charge"], c["multiplicity"])
    r = eng.run(mol, pm7amd.Pm7Options(charge=c["charge"], multiplicity=c["multiplicity"]))
    z, xyz = mol.z, mol.xyz
    n = ob.nao(z)
    rng = np.random.default_rng(11)
    a = rng.normal(size=(3, n, n)) * 0.2
    xs = np.concatenate([a + a.transpose(0, 2, 1), response_densities(r, 3, 5)]) if not r.unrestricted \
        else a + a.transpose(0, 2, 1)
    g = eng.response_fock(mol, xs)
    h = ob.hcore(z, xyz)
    for x, gk in zip(xs, g):
        ref = ob.fock(z, xyz, x, 0.5 * x) - h
        assert np.max(np.abs(gk - ref)) <= 1e-9, np.max(np.abs(gk - ref))
        assert np.max(np.abs(gk - gk.T)) <= 1e-12


def test_response_fock_large_path_molecule(eng):
    """240 AOs: the molecule's integrals come from the large-system preparation."""
    atoms = W.water_cluster(40, 4)
    say("response fock 40 waters")
    mol = pm7amd.Molecule.from_symbols_angstrom(atoms)
    n = 6 * 40
    rng = np.random.default_rng(3)
    a = rng.normal(size=(2, n, n)) * 0.1
    xs = a + a.transpose(0, 2, 1)
    g = eng.response_fock(mol, xs)
    h = ob.hcore(mol.z, mol.xyz)
    for x, gk in zip(xs, g):
        assert np.max(np.abs(gk - (ob.fock(mol.z, mol.xyz, x, 0.5 * x) - h))) <= 1e-9


def test_response_fock_linear_batched_and_consistent(eng):
    """On the largest molecule of the organic benchmark batch: 3N response densities (one per
    Cartesian degree of freedom, as the Hessian's CPHF makes), linearity, batch/alone identity,
    agreement with the Fock entry point, and the converged SCF's commutator."""
    batch = W.organic_batch(256, seed=3)
    atoms = max(batch, key=len)
    mol = pm7amd.Molecule.from_symbols_angstrom(atoms)
    say(f"response fock batch: {len(atoms)} atoms, {3 * len(atoms)} densities")
    r = eng.run(mol, pm7amd.Pm7Options(e_tol=1e-12, p_tol=1e-10), raise_on_error=True)
    xs = response_densities(r, 3 * len(atoms), 9)
    g = eng.response_fock(mol, xs)
    scale = np.max(np.abs(g))
    # linearity: G(aX + bY) = aG(X) + bG(Y)
    lin = eng.response_fock(mol, np.array([0.75 * xs[0] - 1.5 * xs[1]]))[0]
    assert np.max(np.abs(lin - (0.75 * g[0] - 1.5 * g[1]))) <= 1e-12 * max(scale, 1.0) * 10
    # each density's G is the same alone as in the batch, bit for bit
    for k in (0, 7, len(xs) - 1):
        assert np.array_equal(eng.response_fock(mol, xs[k]), g[k])
    # F(X, X/2) - H from the Fock entry point
    h = eng.core_hamiltonian(mol)
    f = eng.fock_spin(mol, xs[3], 0.5 * xs[3])
    assert np.max(np.abs((f - h) - g[3])) <= 1e-12 * max(scale, 1.0)
    # at the converged density, H + G(P) is the SCF's Fock matrix: it commutes with P
    gp = eng.response_fock(mol, r.density)
    fp = h + gp
    comm = fp @ r.density - r.density @ fp
    assert np.max(np.abs(comm)) <= 1e-4


def test_response_fock_edge_cases(eng):
    say("response fock edge cases")
    mol = pm7amd.Molecule.from_symbols_angstrom(W.water())
    assert eng.response_fock(mol, np.zeros((0, 6, 6))).shape == (0, 6, 6)
    assert np.array_equal(eng.response_fock(mol, np.zeros((2, 6, 6))), np.zeros((2, 6, 6)))
    with pytest.raises(ValueError):
        eng.response_fock(mol, np.zeros((1, 5, 5)))


@pytest.mark.parametrize("name", ["methyl_radical", "chlorine_dioxide", "water", "thiophene"])
def test_response_fock_spin_matches_oracle(eng, name):
    say(f"response fock spin {name}")
    c = CASES[name]
    mol = pm7amd.Molecule.from_symbols_angstrom(c["atoms"], c["charge"], c["multiplicity"])
    n = ob.nao(mol.z)
    rng = np.random.default_rng(21)
    a = rng.normal(size=(3, n, n)) * 0.2
    b = rng.normal(size=(3, n, n)) * 0.2
    xa, xb = a + a.transpose(0, 2, 1), b + b.transpose(0, 2, 1)
    ga, gb = eng.response_fock_spin(mol, xa, xb)
    h = ob.hcore(mol.z, mol.xyz)
    for k in range(3):
        assert np.max(np.abs(ga[k] - (ob.fock(mol.z, mol.xyz, xa[k] + xb[k], xa[k]) - h))) <= 1e-9
        assert np.max(np.abs(gb[k] - (ob.fock(mol.z, mol.xyz, xa[k] + xb[k], xb[k]) - h))) <= 1e-9
    # equal halves: the open-shell kernel is the closed-shell one
    g = eng.response_fock(mol, xa)
    ha, hb = eng.response_fock_spin(mol, 0.5 * xa, 0.5 * xa)
    assert np.max(np.abs(ha - g)) <= 1e-12 * max(1.0, np.max(np.abs(g))) and np.array_equal(ha, hb)
    # one pair alone equals the same pair in the batch
    a1, b1 = eng.response_fock_spin(mol, xa[2], xb[2])
    assert np.array_equal(a1, ga[2]) and np.array_equal(b1, gb[2])
