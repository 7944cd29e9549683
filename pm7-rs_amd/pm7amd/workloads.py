# SPDX-License-Identifier: GPL-3.0-or-later
"""Seeded synthetic workloads for the PM7 engine.

Every generator is deterministic across platforms: a splitmix64-seeded xoshiro256** stream,
uniform doubles from the top 53 bits, and Gaussians from an explicitly coded Box-Muller pair
(no library distribution whose algorithm is implementation-defined).

  water()                 one water, O at the origin, O-H 0.96 A, H-O-H 104.5 deg, H1 on +x
  alkane_chain(32, seed)  all-trans CnH2n+2, C-C 1.54, C-H 1.09, tetrahedral, N(0, 0.02 A) jitter
  random_organic(seed)    a closed-shell H/C/N/O molecule grown by a seeded random walk
  organic_batch(256, 3)   256 such molecules, heavy-atom count uniform in 1..25, <= 60 atoms
  water_cluster(500, 4)   waters at random orientation in a 1 g/cm^3 cube, O-O >= 2.6 A
  serving_stream(...)     Poisson arrivals of a mixed request stream
All coordinates are Angstrom lists [[symbol, x, y, z], ...].
"""

from __future__ import annotations

import math
from typing import List

MASK = (1 << 64) - 1


class Rng:
    """xoshiro256** seeded through splitmix64."""

    def __init__(self, seed: int):
        x = seed & MASK
        s = []
        for _ in range(4):
            x = (x + 0x9E3779B97F4A7C15) & MASK
            z = x
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
            s.append(z ^ (z >> 31))
        self.s = s
        self._spare = None

    def next_u64(self) -> int:
        s = self.s
        r = (((s[1] * 5) & MASK) << 7 | ((s[1] * 5) & MASK) >> 57) & MASK
        r = (r * 9) & MASK
        t = (s[1] << 17) & MASK
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = ((s[3] << 45) | (s[3] >> 19)) & MASK
        return r

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1))

    def gauss(self) -> float:
        if self._spare is not None:
            g, self._spare = self._spare, None
            return g
        u1 = self.uniform()
        while u1 <= 0.0:
            u1 = self.uniform()
        u2 = self.uniform()
        r = math.sqrt(-2.0 * math.log(u1))
        self._spare = r * math.sin(2.0 * math.pi * u2)
        return r * math.cos(2.0 * math.pi * u2)


def _unit(v):
    n = math.sqrt(sum(c * c for c in v))
    return [c / n for c in v]


def _cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def _perp(u):
    seed = [0.0, 0.0, 1.0] if abs(u[2]) < 0.9 else [1.0, 0.0, 0.0]
    return _unit(_cross(u, seed))


TET_COS = -1.0 / 3.0  # cos(109.47 deg)


def _cone(axis, count, twist_deg):
    """`count` directions at the tetrahedral angle from -axis (i.e. continuing along +axis)."""
    u = _unit(axis)
    v = _perp(u)
    w = _cross(u, v)
    c, s = 1.0 / 3.0, math.sqrt(8.0) / 3.0
    out = []
    for k in range(count):
        phi = math.radians(twist_deg + 360.0 * k / count)
        out.append([c * u[i] + s * (math.cos(phi) * v[i] + math.sin(phi) * w[i]) for i in range(3)])
    return out


def water():
    a = math.radians(104.5)
    return [["O", 0.0, 0.0, 0.0], ["H", 0.96, 0.0, 0.0], ["H", 0.96 * math.cos(a), 0.96 * math.sin(a), 0.0]]


def alkane_chain(n_carbon: int = 32, seed: int = 2, jitter: float = 0.02):
    theta = math.acos(TET_COS)
    dx, dz = 1.54 * math.sin(theta / 2.0), 1.54 * math.cos(theta / 2.0)
    cs = [[i * dx, 0.0, (i % 2) * dz] for i in range(n_carbon)]
    atoms = [["C", *c] for c in cs]
    for i, c in enumerate(cs):
        nbrs = [cs[j] for j in (i - 1, i + 1) if 0 <= j < n_carbon]
        if len(nbrs) == 2:
            a = _unit([c[k] - nbrs[0][k] for k in range(3)])
            b = _unit([c[k] - nbrs[1][k] for k in range(3)])
            bis = _unit([a[k] + b[k] for k in range(3)])
            nrm = _unit(_cross(a, b))
            half = theta / 2.0
            dirs = [[math.cos(half) * bis[k] + sg * math.sin(half) * nrm[k] for k in range(3)] for sg in (1.0, -1.0)]
        else:
            dirs = _cone([c[k] - nbrs[0][k] for k in range(3)], 3, 0.0)
        for d in dirs:
            atoms.append(["H", c[0] + 1.09 * d[0], c[1] + 1.09 * d[1], c[2] + 1.09 * d[2]])
    rng = Rng(seed)
    for a in atoms:
        for k in (1, 2, 3):
            a[k] += jitter * rng.gauss()
    return atoms


BOND = {("C", "C"): 1.54, ("C", "N"): 1.47, ("C", "O"): 1.43}
XH = {"C": 1.09, "N": 1.01, "O": 0.96}
VALENCE = {"C": 4, "N": 3, "O": 2}


def flat_alkane(n_carbon: int):
    """Linear alkane C_nH_2n+2 in the flat all-anti zig-zag of pm7-rs's divide-and-conquer tests
    and scaling record (tests/dandc.rs:95-127, tests/dandc_scaling.rs:24-44): C at (1.26 i, 0 or
    0.44, 0) Angstrom, two H per carbon at z = +-0.89, one cap H on each end."""
    dx, dy = 1.26, 0.44
    atoms = []
    for i in range(n_carbon):
        x = i * dx
        y = 0.0 if i % 2 == 0 else dy
        hy = y + (-0.5 if i % 2 == 0 else 0.5)
        atoms += [["C", x, y, 0.0], ["H", x, hy, 0.89], ["H", x, hy, -0.89]]
    atoms.append(["H", -1.09, 0.0, 0.0])
    atoms.append(["H", (n_carbon - 1) * dx + 1.09, 0.0 if (n_carbon - 1) % 2 == 0 else dy, 0.0])
    return atoms


def _random_rotation(rng: Rng):
    q = _unit([rng.gauss() for _ in range(4)])
    w, x, y, z = q
    return [[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
            [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
            [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]]


def _apply(R, v):
    return [sum(R[i][k] * v[k] for k in range(3)) for i in range(3)]


def random_organic(rng: Rng, n_heavy: int, max_atoms: int = 60, weights=(0.50, 0.25, 0.25)):
    """Grow a closed-shell H/C/N/O molecule; returns None when a clash forces a restart."""
    tet = [_unit(v) for v in ([1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1])]
    R = _random_rotation(rng)
    heavy = [["C", 0.0, 0.0, 0.0]]
    slots = [[_apply(R, t) for t in tet]]  # free bond directions per heavy atom
    used = [0]
    for _ in range(n_heavy - 1):
        placed = False
        for _attempt in range(30):
            u = rng.uniform()
            el = "C" if u < weights[0] else ("N" if u < weights[0] + weights[1] else "O")
            cands = [i for i, h in enumerate(heavy) if used[i] < VALENCE[h[0]] and (el == "C" or h[0] == "C")]
            if not cands:
                continue
            p = cands[rng.randint(0, len(cands) - 1)]
            if el != "C" and heavy[p][0] != "C":
                continue
            d = slots[p][used[p]]
            r = BOND.get((heavy[p][0], el)) or BOND[(el, heavy[p][0])]
            pos = [heavy[p][1 + k] + r * d[k] for k in range(3)]
            if any(math.dist(pos, h[1:]) < 1.0 for j, h in enumerate(heavy) if j != p):
                continue
            used[p] += 1
            heavy.append([el, *pos])
            slots.append([[-c for c in d]] + _cone(d, 3, 360.0 * rng.uniform()))
            used.append(1)
            placed = True
            break
        if not placed:
            break
    atoms = [list(h) for h in heavy]
    for i, h in enumerate(heavy):
        for k in range(used[i], VALENCE[h[0]]):
            d = slots[i][k]
            pos = [h[1 + q] + XH[h[0]] * d[q] for q in range(3)]
            if any(math.dist(pos, a[1:]) < 1.0 for a in atoms):
                return None
            atoms.append(["H", *pos])
    if len(atoms) > max_atoms:
        return None
    return atoms


def organic_batch(n: int = 256, seed: int = 3, h_min: int = 1, h_max: int = 25, max_atoms: int = 60):
    """Heavy-atom count uniform in [h_min, h_max]; a draw whose saturated molecule exceeds
    `max_atoms` is regrown with the same count (20 tries), then with one heavy atom fewer."""
    rng = Rng(seed)
    out = []
    while len(out) < n:
        h = rng.randint(h_min, h_max)
        mol = None
        while mol is None and h >= 1:
            for _ in range(20):
                mol = random_organic(rng, h, max_atoms)
                if mol is not None and len(mol) >= 3:
                    break
                mol = None
            h -= 1
        if mol is not None:
            out.append(mol)
    return out


def water_cluster(n_water: int = 500, seed: int = 4, density_g_cm3: float = 1.0, min_oo: float = 2.6):
    avogadro = 6.02214076e23
    side = (n_water * 18.015 / avogadro / density_g_cm3) ** (1.0 / 3.0) * 1.0e8
    rng = Rng(seed)
    a = math.radians(104.5)
    local = [[0.0, 0.0, 0.0], [0.96, 0.0, 0.0], [0.96 * math.cos(a), 0.96 * math.sin(a), 0.0]]
    cell = min_oo
    grid = {}
    oxy = []
    atoms = []
    while len(oxy) < n_water:
        p = [side * rng.uniform() for _ in range(3)]
        key = tuple(int(math.floor(c / cell)) for c in p)
        clash = False
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for q in grid.get((key[0] + dx, key[1] + dy, key[2] + dz), ()):
                        if math.dist(p, oxy[q]) < min_oo:
                            clash = True
        if clash:
            continue
        grid.setdefault(key, []).append(len(oxy))
        oxy.append(p)
        R = _random_rotation(rng)
        for sym, v in zip(("O", "H", "H"), local):
            w = _apply(R, v)
            atoms.append([sym, p[0] + w[0], p[1] + w[1], p[2] + w[2]])
    return atoms


def hetero_chain(rng: Rng, n_carbon: int, hetero: str):
    """A jittered all-trans alkane with one heteroatom: Cl in place of a terminal H (C-Cl 1.78 A),
    or an interior CH2 turned into S (thioether, both H removed) or into PH (one H removed).
    3 n_carbon + 2 atoms for Cl, one or two fewer for P and S."""
    atoms = alkane_chain(n_carbon, seed=rng.randint(1, 1 << 30), jitter=0.02)
    n = n_carbon

    def hs(i):  # hydrogens of carbon i (appended after the carbons, in carbon order)
        if i == 0:
            return [n, n + 1, n + 2]
        s0 = n + 3 + 2 * (i - 1)
        return [s0, s0 + 1, s0 + 2] if i == n - 1 else [s0, s0 + 1]

    if hetero == "Cl":
        h = hs(0)[0]
        c = atoms[0][1:]
        d = _unit([atoms[h][1 + k] - c[k] for k in range(3)])
        atoms[h] = ["Cl"] + [c[k] + 1.78 * d[k] for k in range(3)]
        return atoms
    k = rng.randint(1, n - 2)
    drop = set(hs(k) if hetero == "S" else hs(k)[:1])
    atoms[k][0] = hetero
    return [a for j, a in enumerate(atoms) if j not in drop]


def remove_one_h(atoms):
    """A doublet radical: the last hydrogen removed."""
    idx = max(i for i, a in enumerate(atoms) if a[0] == "H")
    return atoms[:idx] + atoms[idx + 1:]


def serving_stream(n_requests: int = 2000, seed: int = 5, rate_per_s: float = 1000.0,
                   frac_sp: float = 0.7, frac_d: float = 0.2):
    """Poisson arrivals of (arrival_s, atoms, charge, multiplicity, kind) requests.

    kind "sp": closed-shell H/C/N/O as in organic_batch; "d": a molecule of 12-200 atoms holding
    S, P or Cl (every pair through the MNDO/d kernels; `hetero_chain` with 4-66 carbons); "uhf":
    an sp molecule with one H removed (doublet).
    """
    rng = Rng(seed)
    t = 0.0
    out = []
    while len(out) < n_requests:
        u = rng.uniform()
        t += -math.log(1.0 - rng.uniform()) / rate_per_s
        if u < frac_sp or u >= frac_sp + frac_d:
            mol = None
            while mol is None or len(mol) < 3:
                mol = random_organic(rng, rng.randint(1, 25), 60)
            if u < frac_sp:
                out.append((t, mol, 0.0, 1, "sp"))
            else:
                out.append((t, remove_one_h(mol), 0.0, 2, "uhf"))
        else:
            n_carbon = rng.randint(4, 66)
            mol = hetero_chain(rng, n_carbon, ("S", "P", "Cl")[rng.randint(0, 2)])
            out.append((t, mol, 0.0, 1, "d"))
    return out


def to_molecule(atoms: List, charge: float = 0.0, multiplicity: int = 1):
    from . import Molecule
    return Molecule.from_symbols_angstrom(atoms, charge, multiplicity)
