# SPDX-License-Identifier: GPL-3.0-or-later
"""Divide-and-conquer SCF on the GPU (pm7_run_dandc) against the CPU oracle's D&C (pinned to the
reference's scaling record and to closed forms in test_oracle_dandc.py), through the C ABI.

Bars: both run the same deterministic iteration (linear mixing, bisection Fermi level), so the
step counts agree within one and the converged quantities to what the eigensolvers' rounding
leaves: electronic energy 1e-6 eV, Fermi level 1e-6 eV, every stored density element and charge
1e-6. The engine's own exact SCF checks the one-subsystem closed form at 1e-6 eV. Refusals carry
their numbers (src/dandc/mod.rs:171-213, tests/dandc.rs:530-575)."""

from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle_binding as ob
import pm7amd
from pm7amd import workloads as W
from pm7amd.dandc import DandcOptions, run_dandc

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = pm7amd.Engine(0)
    yield e
    e.close()


def compare(eng, atoms, mult=1, charge=0.0, de_bar=1e-6, p_bar=1e-6, fermi=True, **dkw):
    say(f"dandc case: {len(atoms)} atoms, multiplicity {mult}, {dkw}")
    mol = pm7amd.Molecule.from_symbols_angstrom(atoms, charge, mult)
    r = run_dandc(mol, options=pm7amd.Pm7Options(charge=charge, multiplicity=mult), dandc=DandcOptions(**dkw),
                  engine=eng)
    z, xyz = ob.atoms_to_arrays(atoms)
    okw = dict(dkw)
    if "buffer" in okw:
        okw["buffer"] = okw.pop("buffer")
    ref, dense, spin, q = ob.run_dandc(z, xyz, charge=charge, multiplicity=mult, **okw)
    say(f"  engine E={r.electronic_ev:.9f} it={r.iterations} conv={r.converged} fermi={r.fermi_ev:.9f} "
        f"device {r.device_ms:.2f} ms; oracle E={ref.electronic_ev:.9f} it={ref.iterations} fermi={ref.fermi_ev:.9f}")
    assert ref.status == 0
    assert r.converged == bool(ref.converged)
    assert abs(r.iterations - ref.iterations) <= 1
    assert r.subsystems == ref.subsystems and r.largest_subsystem == ref.largest_subsystem
    assert r.density.stored_elements() == ref.stored_elements
    assert abs(r.electronic_ev - ref.electronic_ev) < de_bar
    if fermi:
        assert abs(r.fermi_ev - ref.fermi_ev) < de_bar
    mask = dense != 0.0
    got = r.density.to_dense()
    assert np.max(np.abs(got[mask] - dense[mask])) < p_bar
    assert np.max(np.abs(r.charges - q)) < p_bar
    if r.unrestricted:
        assert np.max(np.abs(r.spin_density.to_dense() - spin)) < p_bar
    return r, ref


def test_alkane_c10_against_the_oracle(eng):
    compare(eng, W.flat_alkane(10), buffer=10.0, core_size=6, p_tol=1e-8, max_scf=400)


def test_alkane_c40_scaling_record_against_the_oracle(eng):
    r, ref = compare(eng, W.flat_alkane(40), buffer=15.0, core_size=8, p_tol=1e-7, max_scf=400)
    assert r.subsystems == 16 and r.iterations in (53, 54, 55)


def test_open_shell_against_the_oracle(eng):
    atoms = [["C", 0.0, 0.0, 0.0], ["C", 1.50, 0.0, 0.0], ["H", -0.5, 0.9, 0.0], ["H", -0.5, -0.9, 0.0],
             ["H", 2.05, 0.92, 0.0]]
    # One Fermi level for both spins: the electron count is flat at n_elec (to rounding) across the
    # window between the alpha SOMO and the beta LUMO, so where bisection ends inside it is decided
    # by the last bits of the count, in the reference too; the density and energy do not depend on it.
    r, _ = compare(eng, atoms, mult=2, fermi=False, buffer=25.0, core_size=3, p_tol=1e-8, max_scf=400)
    assert r.unrestricted
    assert abs(r.spin_density.trace() - 1.0) < 0.15


def test_sulfur_chain_mndo_d_subsystems_against_the_oracle(eng):
    # thioether chain: subsystems holding S take the MNDO/d integrals, the others the sp ones
    rng = W.Rng(11)
    atoms = W.hetero_chain(rng, 14, "S")
    compare(eng, atoms, buffer=12.0, core_size=8, p_tol=1e-8, max_scf=400)


def test_charged_chain_against_the_oracle(eng):
    atoms = W.flat_alkane(12)[:-1]  # remove a cap H: odd electron count as a cation
    compare(eng, atoms, charge=1.0, mult=1, buffer=12.0, core_size=6, p_tol=1e-8, max_scf=400)


def test_one_subsystem_equals_the_engines_exact_scf(eng):
    mol = pm7amd.Molecule.from_symbols_angstrom(W.water())
    ex = eng.run(mol, pm7amd.Pm7Options(e_tol=1e-12, p_tol=1e-11, max_scf=500))
    r = run_dandc(mol, dandc=DandcOptions(core_size=12, p_tol=1e-11, max_scf=2000), engine=eng)
    assert r.converged and r.subsystems == 1
    assert abs(r.electronic_ev - ex.electronic_ev) < 1e-6
    assert np.max(np.abs(r.density.to_dense() - ex.density)) < 1e-7


def test_a_run_is_bitwise_repeatable(eng):
    mol = pm7amd.Molecule.from_symbols_angstrom(W.flat_alkane(16))
    d = DandcOptions(buffer=12.0, core_size=6, p_tol=1e-8, max_scf=400)
    a = run_dandc(mol, dandc=d, engine=eng)
    b = run_dandc(mol, dandc=d, engine=eng)
    assert a.iterations == b.iterations and a.electronic_ev == b.electronic_ev
    assert np.array_equal(a.density.data, b.density.data)


def test_c160_against_the_engines_exact_scf(eng):
    # docs/divide_and_conquer.md:92: 482 atoms, 64 subsystems, dE 0.0089 eV, 9.9 % stored, 51 steps
    mol = pm7amd.Molecule.from_symbols_angstrom(W.flat_alkane(160))
    ex = eng.run(mol, pm7amd.Pm7Options(e_tol=1e-10, p_tol=1e-8, max_scf=400))
    r = run_dandc(mol, dandc=DandcOptions(buffer=15.0, core_size=8, p_tol=1e-7, max_scf=400), engine=eng)
    say(f"C160: dE {abs(r.electronic_ev - ex.electronic_ev):.5f} eV, {r.iterations} steps, {r.subsystems} subsystems, "
        f"{100 * r.density.stored_elements() / ex.density.shape[0] ** 2:.1f} % stored, device {r.device_ms:.1f} ms")
    assert r.converged and r.subsystems == 64
    assert abs(abs(r.electronic_ev - ex.electronic_ev) - 0.0089) < 3e-4
    assert abs(r.density.stored_elements() / ex.density.shape[0] ** 2 - 0.099) < 1e-3
    assert abs(r.iterations - 51) <= 1


def test_a_subsystem_too_large_is_refused_with_the_numbers(eng):
    mol = pm7amd.Molecule.from_symbols_angstrom(W.flat_alkane(40))
    with pytest.raises(pm7amd.Pm7Error) as e:
        run_dandc(mol, dandc=DandcOptions(buffer=60.0, core_size=8), engine=eng)
    assert e.value.kind == "UnsupportedFeature"
    text = str(e.value)
    for wanted in ("subsystem", "buffer", "exact SCF", "atoms", "orbitals"):
        assert wanted in text
    assert e.value.largest_subsystem_nao > 196


def test_a_budget_too_small_is_refused_with_the_numbers(eng):
    mol = pm7amd.Molecule.from_symbols_angstrom(W.flat_alkane(40))
    with pytest.raises(pm7amd.Pm7Error) as e:
        run_dandc(mol, options=pm7amd.Pm7Options(max_memory_mb=1), dandc=DandcOptions(buffer=15.0, core_size=8),
                  engine=eng)
    assert e.value.kind == "InsufficientMemory"
    assert "subsystem" in str(e.value) and "exact SCF" in str(e.value)
    assert e.value.needed_mb > e.value.budget_mb > 0


@pytest.mark.parametrize("bad", [dict(buffer=-1.0), dict(core_size=0), dict(fermi_width_ev=0.0), dict(mixing=1.5)])
def test_invalid_options_are_refused(eng, bad):
    mol = pm7amd.Molecule.from_symbols_angstrom(W.water())
    with pytest.raises(pm7amd.Pm7Error) as e:
        run_dandc(mol, dandc=DandcOptions(**bad), engine=eng)
    assert e.value.kind == "InvalidInput"


@pytest.mark.parametrize("mult", [1, 2])
def test_graph_replayed_steps_equal_eager_steps(mult):
    """Steps after the first replay two captured HIP graphs (one per eigenvector-buffer parity);
    the same run with every launch issued eagerly (PM7_AMD_DC_GRAPH=0) gives the same bits. Run
    twice on one context, so graphs captured in the first call are rebuilt, not reused stale."""
    import os
    atoms = W.flat_alkane(40)
    if mult == 2:
        atoms = W.remove_one_h(atoms)
    mol = pm7amd.Molecule.from_symbols_angstrom(atoms, 0.0, mult)
    opts = pm7amd.Pm7Options(multiplicity=mult)
    dopt = DandcOptions(buffer=15.0, core_size=8)
    say(f"dandc graph vs eager, multiplicity {mult}")
    old = os.environ.get("PM7_AMD_DC_GRAPH")
    os.environ["PM7_AMD_DC_GRAPH"] = "0"
    try:
        with pm7amd.Engine(0) as e0:
            eager = run_dandc(mol, options=opts, dandc=dopt, engine=e0)
    finally:
        if old is None:
            os.environ.pop("PM7_AMD_DC_GRAPH", None)
        else:
            os.environ["PM7_AMD_DC_GRAPH"] = old
    with pm7amd.Engine(0) as e1:
        runs = [run_dandc(mol, options=opts, dandc=dopt, engine=e1) for _ in range(2)]
    for r in runs:
        assert r.converged and r.iterations == eager.iterations and r.iterations > 3
        assert r.electronic_ev == eager.electronic_ev and r.fermi_ev == eager.fermi_ev
        assert np.array_equal(r.density.to_dense(), eager.density.to_dense())
