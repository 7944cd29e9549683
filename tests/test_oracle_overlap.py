# SPDX-License-Identifier: GPL-3.0-or-later
"""Pins for the oracle's Slater overlaps: the numerical quadrature used for valence shells with
n >= 4 (src/overlap_numeric.rs) against textbook closed forms and against the analytic kernel
for n <= 3 (the reference's own checks, src/overlap_numeric.rs:180-205, src/overlap.rs:367-398)."""

import math

import numpy as np
import pytest

import oracle_binding as ob


@pytest.mark.parametrize("r", [1.0, 2.0, 3.5])
def test_1s1s_equal_exponents_closed_form(r):
    # S = e^{-R} (1 + R + R^2/3) for zeta = 1 (Mulliken, Rieke, Orloff & Orloff 1949).
    s = ob.overlap_locals(1, 1.0, 1.0, 1, 1.0, 1.0, r, numeric=True)
    assert abs(s[0] - math.exp(-r) * (1.0 + r + r * r / 3.0)) < 1e-10


@pytest.mark.parametrize("r", [1.5, 3.0])
def test_2s2s_equal_exponents_closed_form(r):
    z = 1.3
    p = z * r
    exact = math.exp(-p) * (1.0 + p + 4.0 / 9.0 * p * p + p ** 3 / 9.0 + p ** 4 / 45.0)
    s = ob.overlap_locals(2, z, z, 2, z, z, r, numeric=True)
    assert abs(s[0] - exact) < 1e-9


@pytest.mark.parametrize("na,nb", [(1, 1), (2, 1), (2, 2), (3, 1), (3, 2), (3, 3)])
@pytest.mark.parametrize("r", [1.2, 2.5, 4.0])
def test_quadrature_reproduces_the_closed_forms(na, nb, r):
    za = (1.8, 1.5) if na > 1 else (1.2, 1.2)
    zb = (1.6, 1.3) if nb > 1 else (1.1, 1.1)
    a = ob.overlap_locals(na, *za, nb, *zb, r)
    q = ob.overlap_locals(na, *za, nb, *zb, r, numeric=True)
    sel = [0, 1] if nb == 1 else [0, 1, 2, 3, 4]
    if na == 1:
        sel = [0]
    # the quadrature's own error is ~1e-7 on these shells (the reference accepts 5e-5)
    assert np.max(np.abs(a[sel] - q[sel])) < 1e-6


def test_n_ge_4_goes_through_the_quadrature():
    s = ob.overlap_locals(4, 2.0, 1.8, 1, 1.2, 1.2, 2.8)
    q = ob.overlap_locals(4, 2.0, 1.8, 1, 1.2, 1.2, 2.8, numeric=True)
    assert np.array_equal(s, q)
    assert 0.0 < s[0] < 1.0 and np.all(np.isfinite(s))
