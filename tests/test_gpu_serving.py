# SPDX-License-Identifier: GPL-3.0-or-later
"""GPU: the continuous-batching request server. Requests arrive while others are on the device;
every result must be the one the same molecule gives alone, and admission must respect the
device-memory budget. A molecule that runs the same path alone and in the mix (a megakernel:
no atomics, fixed reduction orders) gives the same bits; one that runs grid-wide alone (a lone
molecule above 128 basis functions) and in a megakernel or a lockstep group in the mix is held
to the parity bars (|dE| 1e-6 eV, |dHf| 1e-4 kcal/mol, max |dP| 1e-7)."""

import sys
import threading
import time

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


def _stream(n, seed):
    # the full request mix: closed-shell H/C/N/O, S/P/Cl molecules (MNDO/d pairs), UHF doublets
    reqs = W.serving_stream(n, seed=seed, rate_per_s=2000.0, frac_sp=0.7, frac_d=0.2)
    return [(t, pm7amd.Molecule.from_symbols_angstrom(a, c, m), pm7amd.Pm7Options(charge=c, multiplicity=m))
            for t, a, c, m, _ in reqs]


GRID_MIN_NAO = 128  # above this a molecule alone runs grid-wide (the engine's routing)


def _same(a, b):
    assert a.status == b.status
    if a.density.shape[0] <= GRID_MIN_NAO:
        assert a.total_ev == b.total_ev and a.heat_of_formation_kcal == b.heat_of_formation_kcal
        assert a.iterations == b.iterations
        assert np.array_equal(a.density, b.density)
        if a.unrestricted:
            assert np.array_equal(a.spin_density, b.spin_density)
    else:
        assert abs(a.total_ev - b.total_ev) <= 1e-6
        assert abs(a.heat_of_formation_kcal - b.heat_of_formation_kcal) <= 1e-4
        assert np.max(np.abs(a.density - b.density)) <= 1e-7
        if a.unrestricted:
            assert np.max(np.abs(a.spin_density - b.spin_density)) <= 1e-7


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_each_result_is_the_same_alone_and_in_the_mix(eng, precision):
    reqs = _stream(40, 17)
    say(f"server mix of {len(reqs)} ({precision})")
    ids = {}
    with pm7amd.Server(eng, lanes=4, max_batch=8) as srv:
        t0 = time.perf_counter()

        def feed():
            for k, (t, mol, opt) in enumerate(reqs):
                dt = t - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                opt.precision = precision
                ids[srv.submit(mol, opt)] = k

        th = threading.Thread(target=feed)
        th.start()
        got = {}
        while len(got) < len(reqs):
            out = srv.poll(-1, timeout_ms=20000)
            assert out is not None, "server stalled"
            rid, res, lat = out
            assert lat >= 0.0
            got[rid] = res
        th.join()
    assert len(got) == len(reqs)
    for rid, res in got.items():
        _, mol, opt = reqs[ids[rid]]
        alone = eng.run(mol, opt, raise_on_error=False)
        _same(res, alone)


def test_admission_respects_the_memory_budget(eng):
    say("server budget")
    big = [pm7amd.Molecule.from_symbols_angstrom(a) for a in W.organic_batch(12, 23, h_min=18, h_max=25)]
    small_budget_mb = 8
    with pm7amd.Server(eng, lanes=4, max_batch=64, max_memory_mb=small_budget_mb) as srv:
        ids = [srv.submit(m) for m in big]
        # a request larger than the whole budget is refused at once
        huge = pm7amd.Molecule.from_symbols_angstrom(W.alkane_chain(32, seed=2))
        hid = srv.submit(huge)
        rid, res, _ = srv.poll(hid, timeout_ms=20000)
        assert res.status == 7 and "budget" in res.message
        results = {}
        for i in ids:
            out = srv.poll(i, timeout_ms=60000)
            assert out is not None
            results[i] = out[1]
    for i, m in zip(ids, big):
        _same(results[i], eng.run(m, raise_on_error=False))


def test_poll_reports_not_ready_and_unknown_ids(eng):
    with pm7amd.Server(eng, lanes=1, max_batch=4) as srv:
        assert srv.poll(-1, timeout_ms=0) is None
        with pytest.raises(pm7amd.Pm7Error):
            srv.poll(12345, timeout_ms=0)
