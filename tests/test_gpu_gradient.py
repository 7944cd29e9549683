# SPDX-License-Identifier: GPL-3.0-or-later
"""Nuclear gradient at the converged density (pm7_gradient_batch; pm7-rs closed_form_gradient,
src/gradient.rs:500-555) on the GPU, held to the derivative it claims to be: the fourth-order
central difference of the CPU oracle's full SCF total energy (tight SCF), so nothing of the device
code enters the reference value. Larger molecules are held to the invariances of any molecular
energy (zero net force and zero net torque) and to batch/alone identity."""

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


def oracle_fd_gradient(z, xyz, charge, mult, h=1.0e-3):
    def energy(x):
        res, *_ = ob.run(z, x, charge, mult, e_tol=1e-12, p_tol=1e-11)
        return res.total_ev

    g = np.zeros_like(xyz)
    for a in range(len(z)):
        for k in range(3):
            ev = []
            for st in (h, -h, 2 * h, -2 * h):
                x = xyz.copy()
                x[a, k] += st
                ev.append(energy(x))
            g[a, k] = (8.0 * (ev[0] - ev[1]) - (ev[2] - ev[3])) / (12.0 * h)
    return g


# water: closed shell; methyl radical: UHF; water dimer: H-bond and dispersion corrections active;
# methylamine, hydrogen fluoride chain: other sp elements; thiophene, H2S, ClO2 (UHF), TiCl4,
# ferrocene: every pair through the MNDO/d kernels (k_pair_grad_d); germane, tetramethyltin:
# valence n >= 4, overlaps by the numerical quadrature at the displaced separations
@pytest.mark.parametrize("name", ["water", "methyl_radical", "water_dimer", "formaldehyde", "methylamine",
                                  "hf_chain_8", "thiophene", "hydrogen_sulfide", "chlorine_dioxide",
                                  "titanium_tetrachloride", "ferrocene", "germane", "tetramethyltin"])
def test_gradient_matches_oracle_scf_finite_difference(eng, name):
    say(f"gradient {name}")
    c = CASES[name]
    mol = pm7amd.Molecule.from_symbols_angstrom(c["atoms"], c["charge"], c["multiplicity"])
    r, g = eng.gradient(mol, pm7amd.Pm7Options(charge=c["charge"], multiplicity=c["multiplicity"],
                                               e_tol=1e-12, p_tol=1e-11))
    ref = oracle_fd_gradient(mol.z, mol.xyz, c["charge"], c["multiplicity"])
    assert np.max(np.abs(g - ref)) <= 1e-5, (g, ref)
    # energies are those of the plain single point
    alone = eng.run(mol, pm7amd.Pm7Options(charge=c["charge"], multiplicity=c["multiplicity"],
                                            e_tol=1e-12, p_tol=1e-11))
    assert r.total_ev == alone.total_ev


def test_gradient_invariances_and_batch_identity(eng):
    """The device part of the gradient (PM7-: no geometric corrections) on the 24 largest molecules
    of the organic benchmark batch. The corrections' host gradient is a difference quotient of the
    correction energy, whose H-bond term selects among exactly equivalent contacts in these
    generated geometries; there it averages the branches, so invariance is checked without it."""
    say("gradient organic batch invariances")
    batch = W.organic_batch(256, seed=3)
    order = sorted(range(len(batch)), key=lambda i: -len(batch[i]))[:24]
    mols = [pm7amd.Molecule.from_symbols_angstrom(batch[i]) for i in order]
    tight = pm7amd.Pm7Options(method="pm7-minus", e_tol=1e-12, p_tol=1e-10)
    res, grads = eng.gradient_batch(mols, tight, raise_on_error=True)
    for m, r, g in zip(mols, res, grads):
        assert np.all(np.isfinite(g))
        assert np.max(np.abs(g.sum(axis=0))) <= 1e-7                      # no net force
        torque = np.cross(m.xyz - m.xyz.mean(axis=0), g).sum(axis=0)
        assert np.max(np.abs(torque)) <= 1e-6                             # no net torque
    r0, g0 = eng.gradient(mols[3], tight)
    assert np.array_equal(g0, grads[3]) and r0.total_ev == res[3].total_ev


def test_mixed_precision_gradient_agrees_with_fp64(eng):
    say("gradient mixed vs fp64")
    batch = W.organic_batch(32, seed=3)
    mols = [pm7amd.Molecule.from_symbols_angstrom(a) for a in batch[:8]]
    _, g64 = eng.gradient_batch(mols, pm7amd.Pm7Options(), raise_on_error=True)
    _, gmx = eng.gradient_batch(mols, pm7amd.Pm7Options(precision="mixed"), raise_on_error=True)
    for a, b in zip(g64, gmx):
        assert np.max(np.abs(a - b)) <= 1e-4


def test_mixed_kinds_in_one_batch_match_alone(eng):
    """sp, MNDO/d and numerical-overlap molecules in one call: each gradient lands at its own offset
    and equals the molecule's gradient alone, bit for bit."""
    say("gradient mixed kinds batch")
    names = ["water", "thiophene", "germane", "methyl_radical", "titanium_tetrachloride"]
    mols = [pm7amd.Molecule.from_symbols_angstrom(CASES[n]["atoms"], CASES[n]["charge"], CASES[n]["multiplicity"])
            for n in names]
    opts = [pm7amd.Pm7Options(charge=CASES[n]["charge"], multiplicity=CASES[n]["multiplicity"]) for n in names]
    res, grads = eng.gradient_batch(mols, opts, raise_on_error=True)
    for m, o, r, g in zip(mols, opts, res, grads):
        r1, g1 = eng.gradient(m, o)
        assert g.shape == (len(m.z), 3) and np.array_equal(g, g1) and r.total_ev == r1.total_ev
