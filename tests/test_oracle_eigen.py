# SPDX-License-Identifier: GPL-3.0-or-later
"""Pins for the oracle's symmetric eigensolver (the reference's `symmetric_eigen`,
src/linalg.rs:231-278): the reference's own known-answer cases (src/linalg.rs:314-356), numpy's
LAPACK `eigh` on random and degenerate matrices, and the Jacobi variant (the device's ordering) run
in a child process because the variant is chosen once per process."""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle_binding as ob

HERE = Path(__file__).resolve().parent


def _check_decomposition(a, w, v, tol):
    n = a.shape[0]
    assert np.all(np.diff(w) >= -1e-12), "eigenvalues must ascend"
    assert np.max(np.abs(v @ np.diag(w) @ v.T - a)) < tol
    assert np.max(np.abs(v.T @ v - np.eye(n))) < tol


def test_reference_known_2x2():
    # src/linalg.rs:315-332: [[2,1],[1,2]] -> 1, 3 and V diag V^T = A.
    a = np.array([[2.0, 1.0], [1.0, 2.0]])
    w, v = ob.eigh(a)
    assert abs(w[0] - 1.0) < 1e-10 and abs(w[1] - 3.0) < 1e-10
    _check_decomposition(a, w, v, 1e-10)


def test_reference_known_3x3_asymmetric_vectors():
    # src/linalg.rs:334-356: distinct eigenvalues, non-symmetric eigenvector matrix.
    a = np.array([[4.0, 1.0, 2.0], [1.0, 3.0, 0.0], [2.0, 0.0, 1.0]])
    w, v = ob.eigh(a)
    _check_decomposition(a, w, v, 1e-9)
    assert not np.allclose(v, v.T)


@pytest.mark.parametrize("n,seed", [(6, 1), (17, 2), (64, 3), (115, 4)])
def test_matches_lapack_eigh(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, n))
    a = 0.5 * (x + x.T) * 10.0
    w, v = ob.eigh(a)
    w_ref, v_ref = np.linalg.eigh(a)
    assert np.max(np.abs(w - w_ref)) < 1e-11 * max(1.0, np.max(np.abs(w_ref)))
    _check_decomposition(a, w, v, 1e-10)
    # the eigenvector gauge is free; the projectors on each eigenvector are not
    for k in range(n):
        assert np.max(np.abs(np.outer(v[:, k], v[:, k]) - np.outer(v_ref[:, k], v_ref[:, k]))) < 1e-8


def test_degenerate_subspace_projector():
    # An exactly threefold eigenvalue: only the subspace projector is fixed.
    rng = np.random.default_rng(5)
    q, _ = np.linalg.qr(rng.standard_normal((8, 8)))
    d = np.array([-3.0, -1.0, -1.0, -1.0, 0.5, 2.0, 2.0, 4.0])
    a = q @ np.diag(d) @ q.T
    a = 0.5 * (a + a.T)
    w, v = ob.eigh(a)
    assert np.max(np.abs(w - d)) < 1e-12
    p = v[:, 1:4] @ v[:, 1:4].T
    p_ref = q[:, 1:4] @ q[:, 1:4].T
    assert np.max(np.abs(p - p_ref)) < 1e-12


def test_jacobi_variant_agrees_with_default():
    code = (
        "import sys, numpy as np; sys.path.insert(0, %r); import oracle_binding as ob\n"
        "rng = np.random.default_rng(7)\n"
        "for n in (2, 7, 30, 64):\n"
        "    x = rng.standard_normal((n, n)); a = 0.5 * (x + x.T)\n"
        "    w, v = ob.eigh(a); wr = np.linalg.eigh(a)[0]\n"
        "    assert np.max(np.abs(w - wr)) < 1e-12, n\n"
        "    assert np.max(np.abs(v @ np.diag(w) @ v.T - a)) < 1e-12, n\n"
        "    assert np.max(np.abs(v.T @ v - np.eye(n))) < 1e-12, n\n"
        "print('ok')\n" % str(HERE)
    )
    env = dict(os.environ, PM7_ORACLE_EIGEN="jacobi")
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "ok"
