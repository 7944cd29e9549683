#!/usr/bin/env python3
# SPDX-License-Identifier: GPL-3.0-or-later
"""Write ``reference_cases.json``: golden vectors for the PM7 single point.

The expected values are pm7-rs 0.2.4's own outputs recorded in the reference's
``tools/oracle/baselines/reference.json`` (CLI ``energy --json``, e_tol 1e-8 eV, p_tol 1e-7): the
total energy, the heat of formation, HOMO/LUMO, Mulliken charges and the dipole parts. That file
holds no geometries; the reference builds each case from small geometric recipes described in
``tools/oracle/baseline.py`` and prints every coordinate with ``%.6f`` Angstrom. The builders
below are written here from those descriptions (trans zig-zag alkane, tetrahedral caps, regular
rings, ...), and each emitted coordinate is rounded to 6 decimals exactly as the reference does,
so the geometry that reaches the SCF is the same point.

Run once in a tree that has the reference at /root/reference; the output is committed.
"""

from __future__ import annotations

import json
import math
import sys
from pathlib import Path

REF_JSON = Path("/root/reference/tools/oracle/baselines/reference.json")
OUT = Path(__file__).resolve().parent / "reference_cases.json"

Atom = tuple[str, float, float, float]

TET_DIRS = ((1, 1, 1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1))
OCT_DIRS = ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1))
COS_T = 1.0 / 3.0
SIN_T = math.sqrt(1.0 - COS_T * COS_T)


def fmt(atoms: list[Atom]) -> list[Atom]:
    """Round like the reference's ``%.6f`` text round trip."""
    return [(e, float(f"{x:.6f}"), float(f"{y:.6f}"), float(f"{z:.6f}")) for e, x, y, z in atoms]


def unit(v):
    n = math.sqrt(sum(c * c for c in v))
    return tuple(c / n for c in v)


def cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def any_perp(u):
    seed = (0.0, 0.0, 1.0) if abs(u[2]) < 0.9 else (1.0, 0.0, 0.0)
    return unit(cross(u, seed))


def cap(centre, neighbour, el, r, count=3, twist=0.0) -> list[Atom]:
    """`count` substituents tetrahedrally about the axis pointing away from `neighbour`."""
    u = unit(tuple(centre[i] - neighbour[i] for i in range(3)))
    v = any_perp(u)
    w = cross(u, v)
    out = []
    for k in range(count):
        phi = math.radians(twist + 360.0 * k / count)
        d = [COS_T * u[i] + SIN_T * (math.cos(phi) * v[i] + math.sin(phi) * w[i]) for i in range(3)]
        out.append((el, centre[0] + r * d[0], centre[1] + r * d[1], centre[2] + r * d[2]))
    return out


def bridge_cap(centre, first, second, el, r, angle=109.47) -> list[Atom]:
    """Methylene pair on a centre with two heavy neighbours, in the bisecting plane."""
    a = unit(tuple(centre[i] - first[i] for i in range(3)))
    b = unit(tuple(centre[i] - second[i] for i in range(3)))
    bis = unit(tuple(a[i] + b[i] for i in range(3)))
    nrm = unit(cross(a, b))
    half = math.radians(angle) / 2.0
    out = []
    for sgn in (1.0, -1.0):
        d = [math.cos(half) * bis[i] + sgn * math.sin(half) * nrm[i] for i in range(3)]
        out.append((el, centre[0] + r * d[0], centre[1] + r * d[1], centre[2] + r * d[2]))
    return out


def diatomic(a, b, r):
    return [(a, 0.0, 0.0, 0.0), (b, r, 0.0, 0.0)]


def linear(els, bonds):
    out, x = [(els[0], 0.0, 0.0, 0.0)], 0.0
    for el, step in zip(els[1:], bonds):
        x += step
        out.append((el, x, 0.0, 0.0))
    return out


def bent(c, lig, r, angle):
    h = math.radians(angle) / 2.0
    return [(c, 0.0, 0.0, 0.0), (lig, r * math.sin(h), r * math.cos(h), 0.0),
            (lig, -r * math.sin(h), r * math.cos(h), 0.0)]


def trigonal_planar(c, lig, r):
    out = [(c, 0.0, 0.0, 0.0)]
    for k in range(3):
        t = math.radians(120.0 * k)
        out.append((lig, r * math.cos(t), r * math.sin(t), 0.0))
    return out


def pyramidal(c, lig, r, angle):
    h = math.radians(angle) / 2.0
    polar = math.asin(min(1.0, math.sin(h) / math.sin(math.radians(60.0))))
    out = [(c, 0.0, 0.0, 0.0)]
    for k in range(3):
        t = math.radians(120.0 * k)
        out.append((lig, r * math.sin(polar) * math.cos(t), r * math.sin(polar) * math.sin(t),
                    r * math.cos(polar)))
    return out


def substituted_methane(ligs, centre="C"):
    out = [(centre, 0.0, 0.0, 0.0)]
    for (el, r), (x, y, z) in zip(ligs, TET_DIRS):
        s = r / math.sqrt(3.0)
        out.append((el, s * x, s * y, s * z))
    return out


def tetrahedral(c, lig, r):
    return substituted_methane([(lig, r)] * 4, c)


def octahedral(c, lig, r):
    return [(c, 0.0, 0.0, 0.0)] + [(lig, r * x, r * y, r * z) for x, y, z in OCT_DIRS]


def dumbbell(a, b, bond, cap_a, cap_b, twist=60.0):
    p1, p2 = (0.0, 0.0, 0.0), (bond, 0.0, 0.0)
    out = [(a, *p1), (b, *p2)]
    out += cap(p1, p2, cap_a[0], cap_a[1], cap_a[2])
    out += cap(p2, p1, cap_b[0], cap_b[1], cap_b[2], twist=twist)
    return out


def square_planar(c, lig, r):
    return [(c, 0.0, 0.0, 0.0)] + [(lig, r * x, r * y, 0.0) for x, y in ((1, 0), (-1, 0), (0, 1), (0, -1))]


def carbonyl(metal, m_c, c_o, shape):
    """Homoleptic M(CO)n with every CO radial: metal first, then C, O per direction."""
    dirs = {
        "tetrahedral": [tuple(c / math.sqrt(3.0) for c in v) for v in TET_DIRS],
        "octahedral": [tuple(float(c) for c in v) for v in OCT_DIRS],
        "bipyramidal": [(1.0, 0.0, 0.0), (-0.5, 0.866025403784, 0.0), (-0.5, -0.866025403784, 0.0),
                        (0.0, 0.0, 1.0), (0.0, 0.0, -1.0)],
    }[shape]
    out = [(metal, 0.0, 0.0, 0.0)]
    for u in dirs:
        out.append(("C", m_c * u[0], m_c * u[1], m_c * u[2]))
        reach = m_c + c_o
        out.append(("O", reach * u[0], reach * u[1], reach * u[2]))
    return out


def metallocene(metal, height, ring_radius, c_h, tilt=0.0):
    """Two C5H5 rings on a metal: eclipsed-staggered sandwich, or bent by moving the ring
    centroids to y = +-tilt (both at z = height)."""
    out = [(metal, 0.0, 0.0, 0.0)]
    for sign in (1.0, -1.0):
        cz = sign * height if tilt == 0.0 else height
        cy = 0.0 if tilt == 0.0 else sign * tilt
        for k in range(5):
            t = 2.0 * math.pi * k / 5.0 + (0.0 if sign > 0 else math.pi / 5.0)
            out.append(("C", ring_radius * math.cos(t), ring_radius * math.sin(t) + cy, cz))
            reach = ring_radius + c_h
            out.append(("H", reach * math.cos(t), reach * math.sin(t) + cy, cz))
    return out


def ring_atoms(els, radius, hs, h_radius, height=0.0):
    n = len(els)
    out = [(el, radius * math.cos(2 * math.pi * k / n), radius * math.sin(2 * math.pi * k / n), height)
           for k, el in enumerate(els)]
    for k, carries in enumerate(hs):
        if carries:
            t = 2 * math.pi * k / n
            out.append(("H", h_radius * math.cos(t), h_radius * math.sin(t), height))
    return out


def substituted_benzene(tail):
    return ring_atoms(["C"] * 6, 1.397, [False] + [True] * 5, 2.483) + tail


def alkane(nc):
    cc, ang = 1.526, math.radians(112.0)
    dx, dz = cc * math.sin(ang / 2.0), cc * math.cos(ang / 2.0)
    out, cs = [], []
    for i in range(nc):
        x, z = i * dx, (i % 2) * dz
        cs.append((x, z))
        out.append(("C", x, 0.0, z))
    for i, (x, z) in enumerate(cs):
        flip = 1.0 if i % 2 == 0 else -1.0
        out.append(("H", x, 0.89, z + 0.45 * flip))
        out.append(("H", x, -0.89, z + 0.45 * flip))
    out.append(("H", cs[0][0] - 1.09, 0.0, cs[0][1]))
    out.append(("H", cs[-1][0] + 1.09, 0.0, cs[-1][1]))
    return out


def hf_chain(n):
    out = []
    for i in range(n):
        out += [("F", i * 2.80, 0.0, 0.0), ("H", i * 2.80 + 0.93, 0.0, 0.0)]
    return out


def water_wire(n):
    out = []
    for i in range(n):
        b, lift = i * 2.85, 0.15 * (i % 2)
        out += [("O", b, lift, 0.0), ("H", b + 0.96, lift + 0.02, 0.0), ("H", b - 0.24, lift + 0.90, 0.30)]
    return out


def cubane():
    half, reach = 0.785, 1.415
    corners = [(x, y, z) for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)]
    return [("C", half * x, half * y, half * z) for x, y, z in corners] + \
           [("H", reach * x, reach * y, reach * z) for x, y, z in corners]


def naphthalene_like():
    a = 1.397
    dx, dy = a * 1.5, a * math.sqrt(3.0) / 2.0
    core = [(a, 0.0), (a / 2, dy), (-a / 2, dy), (-a, 0.0), (-a / 2, -dy), (a / 2, -dy)]
    fused = [(x + dx, y + dy) for x, y in core]
    ring = list(core)
    for p in fused:
        if all(math.hypot(p[0] - x, p[1] - y) > 1e-6 for x, y in ring):
            ring.append(p)
    out = [("C", x, y, 0.0) for x, y in ring]
    cx = sum(x for x, _ in ring) / len(ring)
    cy = sum(y for _, y in ring) / len(ring)
    for x, y in ring:
        nbr = sum(1 for u, v in ring if 0.1 < math.hypot(x - u, y - v) < 1.5 * a)
        if nbr > 2:
            continue
        ux, uy = unit((x - cx, y - cy, 0.0))[:2]
        out.append(("H", x + 1.09 * ux, y + 1.09 * uy, 0.0))
    return out


def glycine():
    n, ca, cc = (0.0, 0.0, 0.0), (1.47, 0.0, 0.0), (2.04, 1.41, 0.0)
    out = [("N", *n), ("C", *ca), ("C", *cc)]
    out += cap(n, ca, "H", 1.01, count=2)
    out += bridge_cap(ca, n, cc, "H", 1.09)
    out += [("O", 3.25, 1.55, 0.0), ("O", 1.20, 2.44, 0.0), ("H", 1.70, 3.28, 0.0)]
    return out


def formic_acid_dimer():
    mono = [("C", 0.0, 0.0, 0.0), ("O", 0.0, 1.20, 0.0), ("O", 1.20, -0.65, 0.0),
            ("H", 1.20, -1.61, 0.0), ("H", -0.94, -0.55, 0.0)]
    cx, cy = 1.20, 0.28
    return mono + [(e, 2 * cx - x, 2 * cy - y, z) for e, x, y, z in mono]


def xyz(lines):
    out = []
    for ln in lines:
        e, x, y, z = ln.split()
        out.append((e, float(x), float(y), float(z)))
    return out


def cases() -> dict[str, tuple[list[Atom], int, int]]:
    tet_h = [("H", 3.70 + 1.09 * x / math.sqrt(3.0), 1.09 * y / math.sqrt(3.0), 1.09 * z / math.sqrt(3.0))
             for x, y, z in TET_DIRS]
    c = {
        "water": (xyz(["O 0.0 0.0 0.0", "H 0.96 0.0 0.0", "H -0.24 0.93 0.0"]), 0, 1),
        "ammonia": (xyz(["N 0.0 0.0 0.0", "H 0.94 0.0 0.38", "H -0.47 0.81 0.38", "H -0.47 -0.81 0.38"]), 0, 1),
        "formaldehyde": (xyz(["C 0.0 0.0 0.0", "O 0.0 0.0 1.21", "H 0.94 0.0 -0.54", "H -0.94 0.0 -0.54"]), 0, 1),
        "methane": (tetrahedral("C", "H", 1.09), 0, 1),
        "hydrogen": (diatomic("H", "H", 0.74), 0, 1),
        "nitrogen": (diatomic("N", "N", 1.10), 0, 1),
        "carbon_monoxide": (diatomic("C", "O", 1.13), 0, 1),
        "hydrogen_fluoride": (diatomic("H", "F", 0.92), 0, 1),
        "carbon_dioxide": (linear(["O", "C", "O"], [1.16, 1.16]), 0, 1),
        "hydrogen_cyanide": (linear(["H", "C", "N"], [1.07, 1.16]), 0, 1),
        "acetylene": (linear(["H", "C", "C", "H"], [1.06, 1.20, 1.06]), 0, 1),
        "ethene": ([("C", 0.0, 0.67, 0.0), ("C", 0.0, -0.67, 0.0), ("H", 0.92, 1.23, 0.0),
                    ("H", -0.92, 1.23, 0.0), ("H", 0.92, -1.23, 0.0), ("H", -0.92, -1.23, 0.0)], 0, 1),
        "ethane": (dumbbell("C", "C", 1.53, ("H", 1.09, 3), ("H", 1.09, 3)), 0, 1),
        "methanol": (dumbbell("C", "O", 1.42, ("H", 1.09, 3), ("H", 0.96, 1)), 0, 1),
        "methylamine": (dumbbell("C", "N", 1.47, ("H", 1.09, 3), ("H", 1.01, 2)), 0, 1),
        "hydrogen_peroxide": (dumbbell("O", "O", 1.45, ("H", 0.97, 1), ("H", 0.97, 1), 90.0), 0, 1),
        "hydrazine": (dumbbell("N", "N", 1.45, ("H", 1.02, 2), ("H", 1.02, 2), 90.0), 0, 1),
        "nitrous_oxide": (linear(["N", "N", "O"], [1.13, 1.19]), 0, 1),
        "ozone": (bent("O", "O", 1.28, 117.0), 0, 1),
        "allene": (linear(["C", "C", "C"], [1.31, 1.31]) + [("H", -0.55, 0.93, 0.0), ("H", -0.55, -0.93, 0.0),
                                                           ("H", 3.17, 0.0, 0.93), ("H", 3.17, 0.0, -0.93)], 0, 1),
        "borane": (trigonal_planar("B", "H", 1.19), 0, 1),
        "boron_trifluoride": (trigonal_planar("B", "F", 1.31), 0, 1),
        "lithium_hydride": (diatomic("Li", "H", 1.60), 0, 1),
        "lithium_fluoride": (diatomic("Li", "F", 1.56), 0, 1),
        "beryllium_dihydride": (linear(["H", "Be", "H"], [1.33, 1.33]), 0, 1),
        "benzene": (ring_atoms(["C"] * 6, 1.397, [True] * 6, 2.483), 0, 1),
        "butadiene": ([("C", 0.0, 0.0, 0.0), ("C", 1.34, 0.0, 0.0), ("C", 2.10, 1.23, 0.0), ("C", 3.44, 1.23, 0.0),
                       ("H", -0.55, 0.93, 0.0), ("H", -0.55, -0.93, 0.0), ("H", 1.89, -0.93, 0.0),
                       ("H", 1.55, 2.16, 0.0), ("H", 3.99, 0.30, 0.0), ("H", 3.99, 2.16, 0.0)], 0, 1),
        "pyridine": (ring_atoms(["N", "C", "C", "C", "C", "C"], 1.395, [False] + [True] * 5, 2.48), 0, 1),
        "furan": (ring_atoms(["O", "C", "C", "C", "C"], 1.12, [False] + [True] * 4, 2.21), 0, 1),
        "pyrrole": (ring_atoms(["N", "C", "C", "C", "C"], 1.17, [True] * 5, 2.26), 0, 1),
        "imidazole": (ring_atoms(["N", "C", "N", "C", "C"], 1.17, [True, True, False, True, True], 2.26), 0, 1),
        "naphthalene_like": (naphthalene_like(), 0, 1),
        "phenol": (substituted_benzene([("O", 2.76, 0.0, 0.0), ("H", 3.10, 0.90, 0.0)]), 0, 1),
        "aniline": (substituted_benzene([("N", 2.80, 0.0, 0.0), ("H", 3.34, 0.82, 0.0), ("H", 3.34, -0.82, 0.0)]), 0, 1),
        "toluene": (substituted_benzene([("C", 2.90, 0.0, 0.0)] + cap((2.90, 0.0, 0.0), (1.397, 0.0, 0.0), "H", 1.09, 3)), 0, 1),
        "cyclopropane": ([("C", 0.870, 0.0, 0.0), ("C", -0.435, 0.753, 0.0), ("C", -0.435, -0.753, 0.0),
                          ("H", 1.520, 0.0, 0.885), ("H", 1.520, 0.0, -0.885), ("H", -0.760, 1.317, 0.885),
                          ("H", -0.760, 1.317, -0.885), ("H", -0.760, -1.317, 0.885), ("H", -0.760, -1.317, -0.885)], 0, 1),
        "cubane": (cubane(), 0, 1),
        "hydroxide": (xyz(["O 0.0 0.0 0.0", "H 0.96 0.0 0.0"]), -1, 1),
        "ammonium": (xyz(["N 0.0 0.0 0.0", "H 0.60 0.60 0.60", "H -0.60 -0.60 0.60", "H -0.60 0.60 -0.60",
                          "H 0.60 -0.60 -0.60"]), 1, 1),
        "hydronium": (pyramidal("O", "H", 0.98, 113.0), 1, 1),
        "fluoride": (xyz(["F 0.0 0.0 0.0"]), -1, 1),
        "cyanide": (diatomic("C", "N", 1.17), -1, 1),
        "nitrate": (trigonal_planar("N", "O", 1.26), -1, 1),
        "carbonate": (trigonal_planar("C", "O", 1.29), -2, 1),
        "acetate": ([("C", 0.0, 0.0, 0.0), ("C", 1.52, 0.0, 0.0), ("O", 2.13, 1.08, 0.0), ("O", 2.13, -1.08, 0.0)]
                    + cap((0.0, 0.0, 0.0), (1.52, 0.0, 0.0), "H", 1.09, 3), -1, 1),
        "methyl_cation": (trigonal_planar("C", "H", 1.09), 1, 1),
        "methyl_anion": (pyramidal("C", "H", 1.10, 108.0), -1, 1),
        "methylammonium": (dumbbell("N", "C", 1.50, ("H", 1.02, 3), ("H", 1.09, 3)), 1, 1),
        "methyl_radical": (xyz(["C 0.0 0.0 0.0", "H 1.08 0.0 0.0", "H -0.54 0.94 0.0", "H -0.54 -0.94 0.0"]), 0, 2),
        "hydroxyl_radical": (diatomic("O", "H", 0.97), 0, 2),
        "nitric_oxide": (diatomic("N", "O", 1.15), 0, 2),
        "nitrogen_dioxide": (bent("N", "O", 1.19, 134.0), 0, 2),
        "ethyl_radical": ([("C", 0.0, 0.0, 0.0), ("C", 1.49, 0.0, 0.0)] + cap((1.49, 0.0, 0.0), (0.0, 0.0, 0.0), "H", 1.09, 3)
                          + [("H", -0.55, 0.93, 0.0), ("H", -0.55, -0.93, 0.0)], 0, 2),
        "oxygen_triplet": (diatomic("O", "O", 1.21), 0, 3),
        "methylene_triplet": (bent("C", "H", 1.08, 134.0), 0, 3),
        "water_dimer": ([("O", 0.0, 0.0, 0.0), ("H", 0.96, 0.0, 0.0), ("H", -0.24, 0.93, 0.0),
                         ("O", 2.90, 0.10, 0.0), ("H", 3.30, -0.30, 0.75), ("H", 3.30, -0.30, -0.75)], 0, 1),
        "formic_acid_dimer": (formic_acid_dimer(), 0, 1),
        "ammonia_water": ([("N", 0.0, 0.0, 0.0), ("O", 2.95, 0.0, 0.0)] + cap((0.0, 0.0, 0.0), (2.95, 0.0, 0.0), "H", 1.02, 3)
                          + [("H", 2.00, 0.05, 0.0), ("H", 3.25, 0.80, 0.45)], 0, 1),
        "benzene_dimer_stacked": (ring_atoms(["C"] * 6, 1.397, [True] * 6, 2.483)
                                  + ring_atoms(["C"] * 6, 1.397, [True] * 6, 2.483, height=3.60), 0, 1),
        "methane_dimer": (tetrahedral("C", "H", 1.09) + [("C", 3.70, 0.0, 0.0)] + tet_h, 0, 1),
        "glycine": (glycine(), 0, 1),
        "alkane_c20": (alkane(20), 0, 1),
        "alkane_c30": (alkane(30), 0, 1),
        "hf_chain_8": (hf_chain(8), 0, 1),
        "water_wire_8": (water_wire(8), 0, 1),
        "water_wire_12": (water_wire(12), 0, 1),
        "water_wire_8_anion": (water_wire(8)[:-1], -1, 1),
        "helium": (xyz(["He 0.0 0.0 0.0"]), 0, 1),
        "neon": (xyz(["Ne 0.0 0.0 0.0"]), 0, 1),
        "carbon_atom_triplet": (xyz(["C 0.0 0.0 0.0"]), 0, 3),
        "nitrogen_atom_quartet": (xyz(["N 0.0 0.0 0.0"]), 0, 4),
        "oxygen_atom_triplet": (xyz(["O 0.0 0.0 0.0"]), 0, 3),
        "fluorine_atom_doublet": (xyz(["F 0.0 0.0 0.0"]), 0, 2),
        "difluoroamino_radical": (bent("N", "F", 1.36, 103.0), 0, 2),
        "cyanide_radical": (diatomic("C", "N", 1.17), 0, 2),
        "allyl_radical": ([("C", 0.0, 0.0, 0.0), ("C", 1.39, 0.0, 0.0), ("C", 2.09, 1.20, 0.0), ("H", -0.55, 0.93, 0.0),
                           ("H", -0.55, -0.93, 0.0), ("H", 1.94, -0.93, 0.0), ("H", 1.54, 2.13, 0.0),
                           ("H", 3.18, 1.20, 0.0)], 0, 2),
        "oxygen_dication_triplet": (diatomic("O", "O", 1.12), 2, 3),
        "carbon_dimer_quintet": (diatomic("C", "C", 1.50), 0, 5),
        # sp-only elements whose valence shell has n >= 4 (numerical Slater overlaps)
        "argon": ([("Ar", 0.0, 0.0, 0.0)], 0, 1),
        "krypton": ([("Kr", 0.0, 0.0, 0.0)], 0, 1),
        "xenon": ([("Xe", 0.0, 0.0, 0.0)], 0, 1),
        "calcium_difluoride": (bent("Ca", "F", 2.10, 140.0), 0, 1),
        "germane": (tetrahedral("Ge", "H", 1.53), 0, 1),
        "hydrogen_selenide": (bent("Se", "H", 1.46, 91.0), 0, 1),
        "stannane": (tetrahedral("Sn", "H", 1.71), 0, 1),
        "hydrogen_telluride": (bent("Te", "H", 1.66, 90.3), 0, 1),
        "thorium_tetrafluoride": (tetrahedral("Th", "F", 2.14), 0, 1),
        "californium_trifluoride": (trigonal_planar("Cf", "F", 2.16), 0, 1),
        "nobelium_trifluoride": (trigonal_planar("No", "F", 2.18), 0, 1),
        "dimethylzinc": ([("Zn", 0.0, 0.0, 0.0), ("C", 1.93, 0.0, 0.0), ("C", -1.93, 0.0, 0.0)]
                         + cap((1.93, 0.0, 0.0), (0.0, 0.0, 0.0), "H", 1.09, 3)
                         + cap((-1.93, 0.0, 0.0), (0.0, 0.0, 0.0), "H", 1.09, 3, twist=60.0), 0, 1),
        # MNDO/d: every molecule below holds an element with a d shell (Al..Cl, heavier p block,
        # transition metals), so all of its pairs take the MNDO/d kernels.
        "hydrogen_sulfide": (xyz(["S 0.0 0.0 0.0", "H 1.34 0.0 0.0", "H -0.35 1.29 0.0"]), 0, 1),
        "sulfur_dioxide": (xyz(["S 0.0 0.0 0.0", "O 1.24 0.0 0.72", "O -1.24 0.0 0.72"]), 0, 1),
        "phosphorus_trifluoride": (xyz(["P 0.0 0.0 0.0", "F 1.24 0.0 0.51", "F -0.62 1.07 0.51",
                                        "F -0.62 -1.07 0.51"]), 0, 1),
        "silane": (tetrahedral("Si", "H", 1.48), 0, 1),
        "phosphine": (pyramidal("P", "H", 1.42, 93.5), 0, 1),
        "aluminium_trichloride": (trigonal_planar("Al", "Cl", 2.06), 0, 1),
        "silicon_tetrafluoride": (tetrahedral("Si", "F", 1.55), 0, 1),
        "carbon_tetrachloride": (tetrahedral("C", "Cl", 1.77), 0, 1),
        "sulfur_hexafluoride": (octahedral("S", "F", 1.56), 0, 1),
        "carbon_disulfide": (linear(["S", "C", "S"], [1.55, 1.55]), 0, 1),
        "carbonyl_sulfide": (linear(["O", "C", "S"], [1.16, 1.56]), 0, 1),
        "thioformaldehyde": (xyz(["C 0.0 0.0 0.0", "S 0.0 0.0 1.61", "H 0.93 0.0 -0.56", "H -0.93 0.0 -0.56"]), 0, 1),
        "methanethiol": (dumbbell("C", "S", 1.82, ("H", 1.09, 3), ("H", 1.34, 1)), 0, 1),
        "silanol": (dumbbell("Si", "O", 1.63, ("H", 1.48, 3), ("H", 0.96, 1)), 0, 1),
        "hydrogen_chloride": (diatomic("H", "Cl", 1.27), 0, 1),
        "hydrogen_bromide": (diatomic("H", "Br", 1.41), 0, 1),
        "hydrogen_iodide": (diatomic("H", "I", 1.61), 0, 1),
        "chlorine": (diatomic("Cl", "Cl", 1.99), 0, 1),
        "bromine": (diatomic("Br", "Br", 2.28), 0, 1),
        "iodine": (diatomic("I", "I", 2.67), 0, 1),
        "chloromethane": (substituted_methane([("H", 1.09), ("H", 1.09), ("H", 1.09), ("Cl", 1.78)]), 0, 1),
        "bromomethane": (substituted_methane([("H", 1.09), ("H", 1.09), ("H", 1.09), ("Br", 1.93)]), 0, 1),
        "iodomethane": (substituted_methane([("H", 1.09), ("H", 1.09), ("H", 1.09), ("I", 2.14)]), 0, 1),
        "chloroform": (substituted_methane([("H", 1.09), ("Cl", 1.77), ("Cl", 1.77), ("Cl", 1.77)]), 0, 1),
        "chlorobenzene": (substituted_benzene([("Cl", 3.17, 0.0, 0.0)]), 0, 1),
        "sodium_chloride": (diatomic("Na", "Cl", 2.36), 0, 1),
        "potassium_bromide": (diatomic("K", "Br", 2.82), 0, 1),
        "magnesium_dichloride": (linear(["Cl", "Mg", "Cl"], [2.18, 2.18]), 0, 1),
        "thiophene": (ring_atoms(["S", "C", "C", "C", "C"], 1.25, [False] + [True] * 4, 2.34), 0, 1),
        "chloride": (xyz(["Cl 0.0 0.0 0.0"]), -1, 1),
        "sulfate": (tetrahedral("S", "O", 1.49), -2, 1),
        "chlorine_dioxide": (bent("Cl", "O", 1.47, 117.0), 0, 2),
        "rubidium_chloride": (diatomic("Rb", "Cl", 2.79), 0, 1),
        "caesium_chloride": (diatomic("Cs", "Cl", 2.91), 0, 1),
        "strontium_dichloride": (linear(["Cl", "Sr", "Cl"], [2.63, 2.63]), 0, 1),
        "barium_dichloride": (linear(["Cl", "Ba", "Cl"], [2.78, 2.78]), 0, 1),
        "gallium_trichloride": (trigonal_planar("Ga", "Cl", 2.09), 0, 1),
        "arsine": (pyramidal("As", "H", 1.51, 92.1), 0, 1),
        "indium_trichloride": (trigonal_planar("In", "Cl", 2.29), 0, 1),
        "stibine": (pyramidal("Sb", "H", 1.70, 91.7), 0, 1),
        "thallium_chloride": (diatomic("Tl", "Cl", 2.48), 0, 1),
        "lead_dichloride": (bent("Pb", "Cl", 2.44, 96.0), 0, 1),
        "bismuth_trichloride": (pyramidal("Bi", "Cl", 2.42, 100.0), 0, 1),
        "scandium_trifluoride": (trigonal_planar("Sc", "F", 1.85), 0, 1),
        "titanium_tetrachloride": (tetrahedral("Ti", "Cl", 2.17), 0, 1),
        "vanadium_oxytrichloride": ([("V", 0.0, 0.0, 0.0), ("O", 0.0, 0.0, 1.57)]
                                    + [("Cl", 2.14 * x / math.sqrt(3.0), 2.14 * y / math.sqrt(3.0),
                                        -2.14 * abs(z) / math.sqrt(3.0)) for x, y, z in TET_DIRS[:3]], 0, 1),
        "chromyl_chloride": ([("Cr", 0.0, 0.0, 0.0), ("O", 1.58, 0.0, 0.79), ("O", -1.58, 0.0, 0.79),
                              ("Cl", 0.0, 1.76, -0.88), ("Cl", 0.0, -1.76, -0.88)], 0, 1),
        "permanganate": (tetrahedral("Mn", "O", 1.63), -1, 1),
        "yttrium_trichloride": (trigonal_planar("Y", "Cl", 2.47), 0, 1),
        "zirconium_tetrachloride": (tetrahedral("Zr", "Cl", 2.32), 0, 1),
        "niobium_pentachloride": ([("Nb", 0.0, 0.0, 0.0)]
                                  + [("Cl", 2.30 * x, 2.30 * y, 0.0) for x, y in
                                     ((1.0, 0.0), (-0.5, 0.866025), (-0.5, -0.866025))]
                                  + [("Cl", 0.0, 0.0, 2.34), ("Cl", 0.0, 0.0, -2.34)], 0, 1),
        "molybdenum_hexafluoride": (octahedral("Mo", "F", 1.82), 0, 1),
        "pertechnetate": (tetrahedral("Tc", "O", 1.71), -1, 1),
        "lanthanum_trichloride": (trigonal_planar("La", "Cl", 2.59), 0, 1),
        "lutetium_trifluoride": (trigonal_planar("Lu", "F", 2.02), 0, 1),
        "hafnium_tetrachloride": (tetrahedral("Hf", "Cl", 2.32), 0, 1),
        "tantalum_pentachloride": ([("Ta", 0.0, 0.0, 0.0)]
                                   + [("Cl", 2.28 * x, 2.28 * y, 0.0) for x, y in
                                      ((1.0, 0.0), (-0.5, 0.866025), (-0.5, -0.866025))]
                                   + [("Cl", 0.0, 0.0, 2.32), ("Cl", 0.0, 0.0, -2.32)], 0, 1),
        "tungsten_hexafluoride": (octahedral("W", "F", 1.83), 0, 1),
        "perrhenate": (tetrahedral("Re", "O", 1.72), -1, 1),
        "osmium_tetroxide": (tetrahedral("Os", "O", 1.71), 0, 1),
        "copper_chloride": (diatomic("Cu", "Cl", 2.05), 0, 1),
        "zinc_dichloride": (linear(["Cl", "Zn", "Cl"], [2.07, 2.07]), 0, 1),
        "silver_chloride": (diatomic("Ag", "Cl", 2.28), 0, 1),
        "cadmium_dichloride": (linear(["Cl", "Cd", "Cl"], [2.28, 2.28]), 0, 1),
        "mercury_dichloride": (linear(["Cl", "Hg", "Cl"], [2.25, 2.25]), 0, 1),
        "tetrachloropalladate": (square_planar("Pd", "Cl", 2.31), -2, 1),
        "tetrachloroplatinate": (square_planar("Pt", "Cl", 2.32), -2, 1),
        "gold_chloride": (diatomic("Au", "Cl", 2.20), 0, 1),
        "nickel_tetracarbonyl": (carbonyl("Ni", 1.84, 1.13, "tetrahedral"), 0, 1),
        "iron_pentacarbonyl": (carbonyl("Fe", 1.81, 1.15, "bipyramidal"), 0, 1),
        "chromium_hexacarbonyl": (carbonyl("Cr", 1.92, 1.14, "octahedral"), 0, 1),
        "molybdenum_hexacarbonyl": (carbonyl("Mo", 2.06, 1.15, "octahedral"), 0, 1),
        "tungsten_hexacarbonyl": (carbonyl("W", 2.06, 1.15, "octahedral"), 0, 1),
        "ruthenium_pentacarbonyl": (carbonyl("Ru", 1.94, 1.14, "bipyramidal"), 0, 1),
        "osmium_pentacarbonyl": (carbonyl("Os", 1.95, 1.14, "bipyramidal"), 0, 1),
        "manganese_carbonyl_hydride": (carbonyl("Mn", 1.85, 1.14, "octahedral")[: 1 + 2 * 5]
                                       + [("H", 0.0, 0.0, -1.58)], 0, 1),
        "cobalt_tetracarbonyl_anion": (carbonyl("Co", 1.80, 1.15, "tetrahedral"), -1, 1),
        "rhodium_dicarbonyl_chloride": ([("Rh", 0.0, 0.0, 0.0), ("C", 1.85, 0.0, 0.0), ("O", 2.99, 0.0, 0.0),
                                         ("C", 0.0, 1.85, 0.0), ("O", 0.0, 2.99, 0.0), ("Cl", -2.32, 0.0, 0.0),
                                         ("Cl", 0.0, -2.32, 0.0)], -1, 1),
        "iridium_dicarbonyl_dichloride": ([("Ir", 0.0, 0.0, 0.0), ("C", 1.87, 0.0, 0.0), ("O", 3.01, 0.0, 0.0),
                                           ("C", 0.0, 1.87, 0.0), ("O", 0.0, 3.01, 0.0), ("Cl", -2.35, 0.0, 0.0),
                                           ("Cl", 0.0, -2.35, 0.0)], -1, 1),
        "ferrocene": (metallocene("Fe", 1.66, 1.43, 1.08), 0, 1),
        "methylmagnesium_chloride": ([("Mg", 0.0, 0.0, 0.0), ("C", 2.10, 0.0, 0.0), ("Cl", -2.20, 0.0, 0.0)]
                                     + cap((2.10, 0.0, 0.0), (0.0, 0.0, 0.0), "H", 1.09, 3), 0, 1),
        "titanocene_dichloride": (metallocene("Ti", 2.06, 1.42, 1.08, tilt=1.30)
                                  + [("Cl", 0.0, 2.10, -1.55), ("Cl", 0.0, -2.10, -1.55)], 0, 1),
        "phosphorus_atom_quartet": (xyz(["P 0.0 0.0 0.0"]), 0, 4),
        "sulfur_atom_triplet": (xyz(["S 0.0 0.0 0.0"]), 0, 3),
        "silicon_atom_triplet": (xyz(["Si 0.0 0.0 0.0"]), 0, 3),
        "tetramethyltin": ([("Sn", 0.0, 0.0, 0.0)]
                           + [("C", 2.14 * x / math.sqrt(3.0), 2.14 * y / math.sqrt(3.0), 2.14 * z / math.sqrt(3.0))
                              for x, y, z in TET_DIRS]
                           + [a for x, y, z in TET_DIRS
                              for a in cap((2.14 * x / math.sqrt(3.0), 2.14 * y / math.sqrt(3.0),
                                            2.14 * z / math.sqrt(3.0)), (0.0, 0.0, 0.0), "H", 1.09, 3)], 0, 1),
    }
    return c


def main() -> None:
    ref = json.loads(REF_JSON.read_text())["cases"]
    out = {}
    for name, (atoms, charge, mult) in cases().items():
        if name not in ref:
            print("skip (not in reference.json):", name, file=sys.stderr)
            continue
        r = ref[name]
        if "pm7_rs" not in r:
            print("skip (no pm7-rs result recorded):", name, file=sys.stderr)
            continue
        atoms = fmt(atoms)
        if r["atoms"] != len(atoms) or r["charge"] != charge or r["multiplicity"] != mult:
            raise SystemExit(f"{name}: builder disagrees with reference.json header")
        pm = r["pm7_rs"]
        out[name] = {
            "atoms": [[e, x, y, z] for e, x, y, z in atoms],
            "charge": charge,
            "multiplicity": mult,
            "energy_ev": pm["energy_ev"],
            "heat_of_formation_kcal": pm["heat_of_formation_kcal"],
            "homo_ev": pm.get("homo_ev"),
            "lumo_ev": pm.get("lumo_ev"),
            "homo_ev_beta": pm.get("homo_ev_beta"),
            "lumo_ev_beta": pm.get("lumo_ev_beta"),
            "charges": pm["charges"],
            "dipole_debye": pm["dipole_debye"],
            "dipole_point_charge_debye": pm["dipole_point_charge_debye"],
            "dipole_sp_hybrid_debye": pm["dipole_sp_hybrid_debye"],
            "mopac_hof_kcal": r["mopac"]["hof_kcal"],
        }
    OUT.write_text(json.dumps({"source": "pm7-rs 0.2.4 tools/oracle/baselines/reference.json",
                               "cases": out}, indent=1))
    print(f"wrote {len(out)} cases to {OUT}")


if __name__ == "__main__":
    main()
