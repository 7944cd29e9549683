# SPDX-License-Identifier: GPL-3.0-or-later
"""Pin the oracle's divide-and-conquer SCF (src/dandc/mod.rs, partition.rs, sparse.rs) and the
engine's host partition (pm7_dandc_partition), CPU only.

What fixes the oracle:

* the reference's published scaling record (docs/divide_and_conquer.md:85-92, linear alkanes,
  buffer 15 Bohr, core_size 8, D&C p_tol 1e-7, exact SCF e_tol 1e-10 / p_tol 1e-8): C40 122 atoms
  dE 0.0019 eV, 36.4 % stored, 16 subsystems, 54 iterations; C80 dE 0.0043 eV, 19.2 %, 32, 52.
  Those four numbers come out of the whole loop (partition, environment field, Fermi level,
  weights, mixing), so a wrong sign, weight or index moves at least one of them;
* closed forms: one subsystem holding every atom as core, or buffers reaching every atom, make
  divide and conquer the exact SCF (no environment, identical subsystem spectra, weights that sum
  to one), so the energy must equal the oracle's exact SCF;
* the invariants of tests/dandc.rs:129-238 and src/dandc/partition.rs:224-330 (buffer convergence,
  sparsity, bounded subsystem size, linear work, exact buffer reach, core ownership, the partition
  of unity of the weights) and the open-shell spin count (tests/dandc.rs:409-436).
"""

import numpy as np
import pytest

import oracle_binding as ob
from pm7amd import Molecule
from pm7amd import workloads as W
from pm7amd.dandc import partition

A = ob.ANGSTROM_TO_BOHR
EXACT = dict(e_tol=1e-10, p_tol=1e-8, max_scf=400)


def alkane(n):
    return ob.atoms_to_arrays(W.flat_alkane(n))


@pytest.mark.parametrize("n,de,stored,subs,iters", [(40, 0.0019, 0.364, 16, 54), (80, 0.0043, 0.192, 32, 52)])
def test_scaling_record_of_the_reference(n, de, stored, subs, iters):
    z, xyz = alkane(n)
    ex, *_ = ob.run(z, xyz, **EXACT)
    r, *_ = ob.run_dandc(z, xyz, buffer=15.0, core_size=8, p_tol=1e-7, max_scf=400)
    assert ex.status == 0 and r.status == 0 and r.converged
    assert abs(abs(r.electronic_ev - ex.electronic_ev) - de) < 5e-5
    assert abs(r.stored_elements / ex.nao ** 2 - stored) < 5e-4
    assert r.subsystems == subs
    assert r.iterations == iters


def test_one_subsystem_is_the_exact_scf():
    # core_size >= n: a single core, no buffer, no environment; Fermi smearing of 0.1 eV changes
    # nothing across a 10 eV gap, and linear mixing has the SCF's fixed point.
    z, xyz = ob.atoms_to_arrays(W.water())
    ex, dens, *_ = ob.run(z, xyz, e_tol=1e-12, p_tol=1e-11, max_scf=500)
    r, dd, _, q = ob.run_dandc(z, xyz, core_size=12, p_tol=1e-11, max_scf=2000)
    assert r.converged and r.subsystems == 1
    assert abs(r.electronic_ev - ex.electronic_ev) < 1e-8
    assert np.max(np.abs(dd - dens)) < 1e-9


def test_buffers_that_reach_every_atom_are_the_exact_scf():
    # Every subsystem holds the whole molecule: each gives the exact Fock matrix and spectrum, the
    # core-weighted norms of a state sum to one over the subsystems, the normalised weights average
    # identical densities.
    z, xyz = alkane(6)
    ex, dens, *_ = ob.run(z, xyz, e_tol=1e-12, p_tol=1e-11, max_scf=500)
    r, dd, _, _ = ob.run_dandc(z, xyz, buffer=200.0, core_size=5, p_tol=1e-10, max_scf=2000)
    assert r.converged and r.subsystems == 4 and r.largest_subsystem == len(z)
    assert abs(r.electronic_ev - ex.electronic_ev) < 1e-7
    assert np.max(np.abs(dd - dens)) < 1e-8


def test_the_energy_converges_to_the_exact_scf_as_the_buffer_grows():
    # tests/dandc.rs:129-170
    z, xyz = alkane(10)
    ex, *_ = ob.run(z, xyz, e_tol=1e-11, p_tol=1e-9)
    errors = []
    for buffer in (6.0, 10.0, 14.0, 20.0, 30.0):
        r, *_ = ob.run_dandc(z, xyz, buffer=buffer, core_size=6, p_tol=1e-8, max_scf=400)
        assert r.converged, buffer
        errors.append(abs(r.electronic_ev - ex.electronic_ev))
    assert errors[-1] < errors[0]
    assert errors[-1] < 0.05


def test_the_density_is_stored_sparsely():
    # tests/dandc.rs:173-208
    fractions = []
    for n in (10, 20, 40):
        z, xyz = alkane(n)
        r, *_ = ob.run_dandc(z, xyz, buffer=10.0, core_size=6, max_scf=1)
        fractions.append(r.stored_elements / ob.nao(z) ** 2)
    assert fractions[0] > fractions[1] > fractions[2]
    assert fractions[-1] < 0.35


def test_an_open_shell_system_runs_unrestricted():
    # tests/dandc.rs:409-436: the ethyl radical's unpaired electron survives partitioning.
    atoms = [["C", 0.0, 0.0, 0.0], ["C", 1.50, 0.0, 0.0], ["H", -0.5, 0.9, 0.0], ["H", -0.5, -0.9, 0.0],
             ["H", 2.05, 0.92, 0.0]]
    z, xyz = ob.atoms_to_arrays(atoms)
    r, dd, spin, _ = ob.run_dandc(z, xyz, multiplicity=2, buffer=25.0, core_size=3, p_tol=1e-8, max_scf=400)
    assert r.unrestricted and r.status == 0
    assert abs(np.trace(spin) - 1.0) < 0.15


# ---- partition (src/dandc/partition.rs) ------------------------------------------------------------

def chain(n):
    return np.array([[2.0 * i, 0.0, 0.0] for i in range(n)])


def test_every_atom_is_the_core_of_exactly_one_subsystem():
    owned = np.zeros(40, int)
    for members, nc in ob.dandc_partition(chain(40), 5, 6.0):
        for a in members[:nc]:
            owned[a] += 1
    assert np.all(owned == 1)


def test_the_buffer_reaches_exactly_as_far_as_it_is_told():
    pos = chain(40)
    for members, nc in ob.dandc_partition(pos, 4, 6.0):
        core = pos[members[:nc]]
        for b in members[nc:]:
            assert np.min(np.linalg.norm(core - pos[b], axis=1)) <= 6.0 + 1e-12
        for a in range(40):
            if np.min(np.linalg.norm(core - pos[a], axis=1)) <= 6.0 - 1e-12:
                assert a in members


def test_subsystem_size_stays_flat_as_the_system_grows():
    small = ob.dandc_partition(chain(40), 5, 6.0)
    large = ob.dandc_partition(chain(160), 5, 6.0)
    assert len(large) == 4 * len(small)
    assert max(len(m) for m, _ in large) <= max(len(m) for m, _ in small) + 1


def test_weights_are_a_partition_of_unity_away_from_buffer_edges():
    # src/dandc/partition.rs:275-330: a pair whose atoms both sit inside each other's buffers
    # collects exactly 1 (1 core-core, or 1/2 from each side).
    pos = chain(30)
    parts = ob.dandc_partition(pos, 5, 6.0)
    total = {}
    for members, nc in parts:
        for i, a in enumerate(members):
            for j, b in enumerate(members):
                w = 1.0 if (i < nc and j < nc) else (0.5 if (i < nc or j < nc) else 0.0)
                if w:
                    total[(a, b)] = total.get((a, b), 0.0) + w
    for (a, b), w in total.items():
        if abs(pos[a, 0] - pos[b, 0]) <= 2.0:
            assert abs(w - 1.0) < 1e-12, (a, b, w)


def log_log_slope(points):
    x = np.log([p[0] for p in points])
    y = np.log([p[1] for p in points])
    return float(np.polyfit(x, y, 1)[0])


def test_the_diagonalization_work_grows_linearly():
    # tests/dandc.rs:353-406 (partition arithmetic only)
    cubic, quad = [], []
    for n in (20, 40, 80, 160, 320):
        z, xyz = alkane(n)
        parts = ob.dandc_partition(xyz, 8, 12.0)
        cubic.append((len(z), sum(len(m) ** 3 for m, _ in parts)))
        quad.append((len(z), sum(len(m) ** 2 for m, _ in parts)))
    assert log_log_slope(cubic) <= 1.2
    assert log_log_slope(quad) <= 1.2


@pytest.mark.parametrize("n,core,buffer", [(40, 8, 15.0), (80, 8, 15.0), (10, 6, 10.0), (33, 3, 7.5)])
def test_the_engine_partition_equals_the_oracle_partition(n, core, buffer):
    # The engine's host partition (grid-accelerated) against the oracle's brute force: the same cores
    # in the same order, the same buffer sets.
    z, xyz = alkane(n)
    ref = ob.dandc_partition(xyz, core, buffer)
    got = partition(Molecule.from_arrays(z, xyz), core, buffer)
    assert len(got.subsystems) == len(ref)
    for s, (members, nc) in zip(got.subsystems, ref):
        assert s.n_core == nc
        assert s.parent[:nc] == [int(a) for a in members[:nc]]
        assert sorted(s.parent[nc:]) == sorted(int(a) for a in members[nc:])


def test_the_engine_partition_of_a_3d_cluster_equals_the_oracle():
    atoms = W.water_cluster(60, seed=7)
    z, xyz = ob.atoms_to_arrays(atoms)
    for core, buffer in ((12, 15.0), (5, 9.0)):
        ref = ob.dandc_partition(xyz, core, buffer)
        got = partition(Molecule.from_arrays(z, xyz), core, buffer)
        assert [s.n_core for s in got.subsystems] == [nc for _, nc in ref]
        for s, (members, nc) in zip(got.subsystems, ref):
            assert s.parent[:nc] == [int(a) for a in members[:nc]]
            assert sorted(s.parent[nc:]) == sorted(int(a) for a in members[nc:])
