# SPDX-License-Identifier: GPL-3.0-or-later
"""Host-side pieces of the Python mirror that need no device: molecule marshalling and the
zero-copy result conversion (arrays stay owned by the library until the last view is gone)."""

from __future__ import annotations

import ctypes as C
import gc

import numpy as np

import pm7amd


def test_molecule_arrays_are_bohr_and_int32():
    m = pm7amd.Molecule.from_symbols_angstrom([("O", 0, 0, 0), ("H", 0.96, 0, 0)], charge=0.0)
    z, xyz = m.arrays()
    assert z.dtype == np.int32 and z.tolist() == [8, 1]
    assert xyz.shape == (2, 3) and abs(xyz[1, 0] - 0.96 / 0.529177210903) < 1e-12
    assert [a.z for a in m.atoms] == [8, 1]


def test_molecule_from_xyz_matches_symbols():
    text = "3\nwater\nO 0 0 0\nH 0.96 0 0\nH -0.24 0.93 0\n"
    a = pm7amd.Molecule.from_xyz_str(text)
    b = pm7amd.Molecule.from_symbols_angstrom([("O", 0, 0, 0), ("H", 0.96, 0, 0), ("H", -0.24, 0.93, 0)])
    assert np.array_equal(a.z, b.z) and np.array_equal(a.xyz, b.xyz)


def test_result_views_free_once_after_last_reference():
    n, na = 5, 3
    r = pm7amd._Result()
    r.nao, r.n_atoms, r.total_ev, r.status = n, na, -1.5, 0
    keep = {}
    for name, k in (("density", n * n), ("mo_energies", n), ("mo_coeff", n * n), ("charges", na)):
        buf = (C.c_double * k)(*[float(i) for i in range(k)])
        keep[name] = buf
        setattr(r, name, C.cast(buf, C.POINTER(C.c_double)))
    calls = []

    def free(ptr):
        calls.append(ptr)

    res = pm7amd._convert(r, free)
    assert res.total_ev == -1.5 and res.density.shape == (n, n)
    assert res.density[1, 2] == 7.0 and res.charges.tolist() == [0.0, 1.0, 2.0]
    assert res.spin_density is None and res.mo_coeff_beta is None
    dens = res.density
    del res
    gc.collect()
    assert calls == []          # a view is still alive
    del dens
    gc.collect()
    assert len(calls) == 1
