# SPDX-License-Identifier: GPL-3.0-or-later
"""GPU tests of pm7_autotune and of every setting it can choose.

The tuner picks, per size class, the block Jacobi solver or the scalar rounds, the sweep cap of the
reduced-precision phase (3, 5 or 8) and the number of batches in flight. Each of those settings is
held to the parity bars here (the scalar rounds against the CPU oracle, the sweep caps against the
FP64 results); the tuner's own contract is checked on a small batch: a valid choice, kept in the
context, written to and read back from its cache file, results after tuning bit for bit those of a
context that takes the same choice from the cache.
"""

import os
import sys

import numpy as np
import pytest

import oracle_binding as ob
import pm7amd
from pm7amd import workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def batch(n, seed=3):
    atoms = W.organic_batch(n, seed)
    return atoms, [pm7amd.Molecule.from_symbols_angstrom(a) for a in atoms]


def engine_with(env):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update({k: str(v) for k, v in env.items()})
    try:
        return pm7amd.Engine(0)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def test_scalar_rounds_match_the_oracle():
    """PM7_AMD_EIG_BLOCK=0 (the scalar Jacobi rounds in every class) against the oracle: molecules
    from all three megakernel classes of the organic batch."""
    atoms, mols = batch(256)
    nao = [ob.nao(ob.atoms_to_arrays(a)[0]) for a in atoms]
    pick = sorted(range(len(atoms)), key=lambda i: nao[i])
    sample = [pick[0], pick[60], pick[128], pick[200], pick[250], pick[-1]]
    say(f"scalar rounds vs oracle, nao {[nao[i] for i in sample]}")
    with engine_with({"PM7_AMD_EIG_BLOCK": 0}) as eng:
        res = eng.run_batch([mols[i] for i in sample])
    for i, r in zip(sample, res):
        z, xyz = ob.atoms_to_arrays(atoms[i])
        ref, dens, *_ = ob.run(z, xyz)
        assert r.status == 0 and r.converged
        assert abs(r.total_ev - ref.total_ev) <= 1e-6
        assert abs(r.heat_of_formation_kcal - ref.hof_kcal) <= 1e-4
        assert np.max(np.abs(r.density - dens)) <= 1e-7


@pytest.mark.parametrize("cap", [5, 8])
def test_reduced_precision_sweep_caps_stay_within_bars(cap):
    """Mixed runs with the sweep caps the tuner may choose, against FP64 on the same molecules."""
    _, mols = batch(48, seed=7)
    say(f"mixed sweep cap {cap}")
    with engine_with({"PM7_AMD_MIXED_SWEEPS": cap}) as eng:
        fp = eng.run_batch(mols, pm7amd.Pm7Options())
        mx = eng.run_batch(mols, pm7amd.Pm7Options(precision="mixed"))
    for f, m in zip(fp, mx):
        assert m.status == 0 and m.converged
        assert abs(m.total_ev - f.total_ev) <= 1e-6
        assert abs(m.heat_of_formation_kcal - f.heat_of_formation_kcal) <= 1e-4


def test_autotune_choice_is_valid_kept_and_cached(tmp_path):
    _, mols = batch(64, seed=5)
    cache = tmp_path / "tune.txt"
    opts = pm7amd.Pm7Options(precision="mixed")
    say("autotune on 64 molecules")
    with pm7amd.Engine(0) as eng:
        t = eng.autotune(mols, opts, steps=2, latency_budget=2.0, cache_path=str(cache))
        say(f"  choice {t}")
        assert 0 <= t["eig_block_mask"] <= 15
        assert t["mixed_sweeps"] in (3, 5, 8)
        assert 1 <= t["batches_in_flight"] <= 4
        assert t["candidates"] >= 6 and t["solves_per_s"] > 0 and t["step_ms"] > 0
        assert t["default_solves_per_s"] > 0
        after = eng.run_batch(mols, opts)
    assert cache.exists() and len(cache.read_text().splitlines()) == 1
    with pm7amd.Engine(0) as eng2:
        t2 = eng2.autotune(mols, opts, steps=2, cache_path=str(cache))
        assert t2["candidates"] == 0
        for k in ("eig_block_mask", "mixed_sweeps", "batches_in_flight"):
            assert t2[k] == t[k], k
        again = eng2.run_batch(mols, opts)
    for a, b in zip(after, again):
        assert a.status == b.status == 0
        assert a.total_ev == b.total_ev
        assert np.array_equal(a.density, b.density)


def test_autotune_refuses_while_batches_are_in_flight():
    _, mols = batch(8, seed=9)
    with pm7amd.Engine(0) as eng:
        ticket = eng.submit_batch(mols)
        with pytest.raises(pm7amd.Pm7Error):
            eng.autotune(mols, steps=1)
        eng.wait_batch(ticket)
