# SPDX-License-Identifier: GPL-3.0-or-later
"""Asynchronous batches (pm7_batch_submit / pm7_batch_wait) through the C ABI: every batch's results
are bit for bit what pm7_run_batch returns for the same molecules, with several batches in flight,
waits in any order, more submits than lanes (a submit completes the oldest batch itself), batches
of mixed precision and of invalid molecules, and an unknown ticket refused."""

from __future__ import annotations

import sys

import numpy as np
import pytest

import pm7amd
from pm7amd import workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = pm7amd.Engine(0)
    yield e
    e.close()


def same(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert x.status == y.status
        if x.status not in (0, 6):
            continue
        assert x.total_ev == y.total_ev and x.heat_of_formation_kcal == y.heat_of_formation_kcal
        assert x.iterations == y.iterations
        assert np.array_equal(x.density, y.density) and np.array_equal(x.mo_energies, y.mo_energies)
        assert np.array_equal(x.charges, y.charges)


def test_batches_in_flight_equal_the_synchronous_call(eng):
    say("async: 5 batches of 64, waits out of order")
    mols = [pm7amd.Molecule.from_symbols_angstrom(a) for a in W.organic_batch(64, 3)]
    groups = [mols[k:] + mols[:k] for k in (0, 5, 17, 33, 60)]
    ref = [eng.run_batch(g) for g in groups]
    tickets = [eng.submit_batch(g) for g in groups]  # more than the lanes: submits retire the oldest
    for k in (3, 0, 4, 1, 2):
        same(eng.wait_batch(tickets[k]), ref[k])


def test_mixed_precision_and_open_shells_in_flight(eng):
    say("async: mixed precision, UHF")
    atoms = W.organic_batch(24, 7)
    mols = [pm7amd.Molecule.from_symbols_angstrom(a) for a in atoms]
    rad = [pm7amd.Molecule.from_symbols_angstrom(W.remove_one_h(a), 0.0, 2) for a in atoms
           if any(x[0] == "H" for x in a)][:8]
    o_mx = pm7amd.Pm7Options(precision="mixed")
    o_uhf = pm7amd.Pm7Options(multiplicity=2)
    ref_mx = eng.run_batch(mols, o_mx)
    ref_uhf = eng.run_batch(rad, o_uhf)
    t1 = eng.submit_batch(mols, o_mx)
    t2 = eng.submit_batch(rad, o_uhf)
    same(eng.wait_batch(t2), ref_uhf)
    same(eng.wait_batch(t1), ref_mx)


def test_invalid_molecules_and_an_unknown_ticket(eng):
    say("async: invalid molecules, unknown ticket")
    good = pm7amd.Molecule.from_symbols_angstrom(W.water())
    bad = pm7amd.Molecule.from_arrays([8, 1], [[0, 0, 0], [0, 0, 0]])  # coincident atoms
    ref = eng.run_batch([good, bad])
    t = eng.submit_batch([good, bad])
    got = eng.wait_batch(t)
    same(got, ref)
    assert got[1].status == 1
    t_empty = eng.submit_batch([bad])  # nothing valid: the wait returns at once
    assert eng.wait_batch(t_empty)[0].status == 1
    assert eng._lib.pm7_batch_wait(eng._h, 10 ** 9) == 1  # PM7_ERR_INVALID_INPUT


def test_timings_place_each_batch_on_the_device_clock(eng):
    mols = [pm7amd.Molecule.from_symbols_angstrom(a) for a in W.organic_batch(32, 9)]
    t1 = eng.submit_batch(mols)
    t2 = eng.submit_batch(mols)
    eng.wait_batch(t1)
    a = eng.last_timings_ms()
    eng.wait_batch(t2)
    b = eng.last_timings_ms()
    assert 0.0 <= a[12] < a[13] and 0.0 <= b[12] < b[13]
    assert b[13] >= a[12]
