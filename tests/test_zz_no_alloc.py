# SPDX-License-Identifier: GPL-3.0-or-later
"""pm7_fock_spin_noalloc (the allocation-free variant of pm7_fock_spin, whose device layout lives
in a static per-device scratch) against pm7_fock_spin through the C ABI: bit-identical Fock
matrices on two calls in a row, the second on a smaller molecule (the scratch keeps the first
call's contents), and a molecule too large for the scratch refused with INSUFFICIENT_MEMORY."""

import ctypes as C
import json
import sys
from pathlib import Path

import numpy as np
import pytest

import pm7amd
from pm7amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = json.loads((Path(__file__).parent / "golden" / "reference_cases.json").read_text())["cases"]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def call(lib, h, fn, mol, pt, ps):
    cm, keep = pm7amd.Engine._mol(mol)
    n = pt.shape[0]
    f = np.zeros((n, n))
    dp = C.POINTER(C.c_double)
    st = fn(h, C.byref(cm), np.ascontiguousarray(pt).ctypes.data_as(dp),
            np.ascontiguousarray(ps).ctypes.data_as(dp), f.ctypes.data_as(dp))
    return st, f


def densities(eng, mol):
    r = eng.run(mol, pm7amd.Pm7Options(charge=mol.charge, multiplicity=mol.multiplicity))
    ps = 0.5 * r.density if r.spin_density is None else 0.5 * (r.density + r.spin_density)
    return r.density, ps


def test_noalloc_fock_equals_fock_spin_on_two_calls():
    lib = pm7amd.library()
    with pm7amd.Engine(0) as eng:
        cases = [pm7amd.Molecule.from_symbols_angstrom(W.alkane_chain(32, 2)),
                 pm7amd.Molecule.from_symbols_angstrom(CASES["thiophene"]["atoms"]),
                 pm7amd.Molecule.from_symbols_angstrom(CASES["methyl_radical"]["atoms"], 0.0, 2)]
        for i, mol in enumerate(cases):  # large, then smaller ones (MNDO/d, open shell)
            say(f"noalloc call {i}: {len(mol.z)} atoms")
            pt, ps = densities(eng, mol)
            st0, f0 = call(lib, eng._h, lib.pm7_fock_spin, mol, pt, ps)
            st1, f1 = call(lib, eng._h, lib.pm7_fock_spin_noalloc, mol, pt, ps)
            assert st0 == 0 and st1 == 0, eng.last_error()
            assert np.array_equal(f0, f1)
            assert np.max(np.abs(f1 - f1.T)) == 0.0


def test_noalloc_refuses_a_molecule_larger_than_its_scratch():
    lib = pm7amd.library()
    with pm7amd.Engine(0) as eng:
        mol = pm7amd.Molecule.from_symbols_angstrom(W.water_cluster(80, 4))  # 480 AOs: ~60 MB of slab
        n = 480
        say("noalloc too large")
        st, _ = call(lib, eng._h, lib.pm7_fock_spin_noalloc, mol, np.eye(n), 0.5 * np.eye(n))
        assert st == 7, eng.last_error()
        assert "scratch" in eng.last_error()
        small = pm7amd.Molecule.from_symbols_angstrom(W.water())
        pt, ps = densities(eng, small)
        st0, f0 = call(lib, eng._h, lib.pm7_fock_spin, small, pt, ps)
        st1, f1 = call(lib, eng._h, lib.pm7_fock_spin_noalloc, small, pt, ps)
        assert st0 == 0 and st1 == 0 and np.array_equal(f0, f1)
