"""Wire and disk formats of pm7-rs for molecules: the (extended-)XYZ reader and writer
(src/system.rs:149-305, 314-345) and the ``energy --json`` record (src/bin/pm7_rs.rs:374-441).

Only molecules are in scope: an extended-XYZ comment line carrying ``Lattice=`` describes a
periodic cell, which this engine does not solve, so such input is refused as UnsupportedFeature
rather than silently read as a molecule.
"""
from __future__ import annotations

import math
from pathlib import Path
from typing import Optional, Union

import numpy as np

from . import ANGSTROM_TO_BOHR, SYMBOLS, Molecule, Pm7Error, Pm7Result

BOHR_TO_ANGSTROM = 1.0 / ANGSTROM_TO_BOHR


def symbol_to_z(sym: str) -> Optional[int]:
    """src/system.rs:314-327: an atomic number 1..102, or a symbol in any letter case."""
    try:
        z = int(sym)
    except ValueError:
        z = None
    if z is not None and 1 <= z < len(SYMBOLS):
        return z
    if not sym:
        return None
    norm = sym[0].upper() + sym[1:].lower()
    try:
        z = SYMBOLS.index(norm)
    except ValueError:
        return None
    return z if z >= 1 else None


def z_to_symbol(z: int) -> Optional[str]:
    return SYMBOLS[z] if 1 <= z < len(SYMBOLS) else None


def _quoted_value(text: str, key: str) -> Optional[str]:
    """Value of ``key="..."`` (or a bare ``key=value``) in an extended-XYZ comment line, where the
    key starts the line or follows whitespace (src/system.rs:267-293)."""
    rest = text
    while True:
        idx = rest.find(key)
        if idx < 0:
            return None
        before_ok = idx == 0 or rest[idx - 1].isspace()
        after = rest[idx + len(key):]
        if before_ok and after.startswith("="):
            after = after[1:]
            if after.startswith('"'):
                return after[1:].split('"', 1)[0]
            parts = after.split()
            return parts[0] if parts else ""
        rest = after


def _parse_float(token: str, line_no: int) -> float:
    try:
        v = float(token)
    except ValueError:
        raise Pm7Error(1, f"line {line_no}: invalid floating point value: {token}") from None
    return v


def read_xyz(text: str, charge: float = 0.0, multiplicity: int = 1) -> Molecule:
    """``Molecule::from_xyz_str`` (src/system.rs:149-196): atom count, comment, then one
    ``symbol x y z`` line per atom in Angstrom (extra columns ignored); positions are stored in Bohr.
    A leading UTF-8 byte-order mark is skipped."""
    if text.startswith("\ufeff"):
        text = text[1:]
    lines = text.splitlines()
    if not lines:
        raise Pm7Error(1, "empty XYZ")
    try:
        natoms = int(lines[0].strip())
        if natoms < 0:
            raise ValueError
    except ValueError:
        raise Pm7Error(1, f"invalid XYZ atom count: {lines[0]}") from None
    comment = lines[1] if len(lines) > 1 else ""
    if _quoted_value(comment, "Lattice") is not None:
        raise Pm7Error(4, "extended-XYZ Lattice describes a periodic cell; only molecules are supported")
    z = np.empty(natoms, dtype=np.int32)
    xyz = np.empty((natoms, 3), dtype=np.float64)
    for idx in range(natoms):
        line_no = idx + 3
        if 2 + idx >= len(lines):
            raise Pm7Error(1, f"XYZ ended before atom {idx}")
        parts = lines[2 + idx].split()
        if len(parts) < 4:
            raise Pm7Error(1, f"XYZ line {line_no} has fewer than 4 fields")
        zz = symbol_to_z(parts[0])
        if zz is None:
            raise Pm7Error(1, f"unknown element on line {line_no}: {parts[0]}")
        z[idx] = zz
        for k in range(3):
            xyz[idx, k] = _parse_float(parts[1 + k], line_no) * ANGSTROM_TO_BOHR
    return Molecule.from_arrays(z, xyz, charge, multiplicity)


def read_xyz_file(path: Union[str, Path], charge: float = 0.0, multiplicity: int = 1) -> Molecule:
    return read_xyz(Path(path).read_text(encoding="utf-8"), charge, multiplicity)


def to_xyz_string(molecule: Molecule, title: str = "") -> str:
    """``Molecule::to_xyz_string`` for a molecule (src/system.rs:198-232): Angstrom, 12 decimals;
    quotes and newlines in the title become spaces."""
    comment = title.replace('"', " ").replace("\n", " ")
    out = [f"{len(molecule.z)}", comment]
    for zz, p in zip(molecule.z, molecule.xyz):
        sym = z_to_symbol(int(zz)) or "X"
        out.append(f"{sym} {p[0] * BOHR_TO_ANGSTROM:.12f} {p[1] * BOHR_TO_ANGSTROM:.12f} "
                   f"{p[2] * BOHR_TO_ANGSTROM:.12f}")
    return "\n".join(out) + "\n"


def gap_ev(result: Pm7Result) -> Optional[float]:
    """``Pm7Result::gap_ev`` (src/scf.rs:462-472): the higher occupied and the lower virtual
    frontier over both spin channels."""
    homos = [e for e in (result.homo_ev, result.homo_ev_beta) if e is not None]
    lumos = [e for e in (result.lumo_ev, result.lumo_ev_beta) if e is not None]
    if not homos or not lumos:
        return None
    return min(lumos) - max(homos)


def spin_squared(result: Pm7Result) -> Optional[float]:
    """``Pm7Result::spin_squared`` (src/scf.rs:423-437): <S^2> of an unrestricted determinant with
    the NDDO identity overlap, S_z(S_z + 1) + n_beta - Tr(P^a P^b), Tr(P^a P^b) =
    (Tr(P P) - Tr(P^s P^s)) / 4. None for a restricted result."""
    if not result.unrestricted or result.spin_density is None or result.n_occ_beta is None:
        return None
    nb = int(result.n_occ_beta)
    sz = 0.5 * (int(result.n_occ) - nb)
    p = np.asarray(result.density, dtype=np.float64).reshape(-1)
    s = np.asarray(result.spin_density, dtype=np.float64).reshape(-1)
    overlap = 0.25 * (float(p @ p) - float(s @ s))
    return sz * (sz + 1.0) + nb - overlap


def _f12(x: float) -> str:
    # Rust's `{:.12}` on f64: fixed notation; non-finite values print as NaN / inf / -inf.
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    return f"{x:.12f}"


def energy_json(result: Pm7Result, method: str = "pm7") -> str:
    """The one-line JSON record of ``pm7_rs energy --json`` for a molecule
    (src/bin/pm7_rs.rs:374-441): same keys, same order, twelve decimals. The periodic-only keys
    (field, k-points, Fermi level, entropy) never appear for a molecule."""
    d = result.dipole_debye
    pc = result.dipole_point_charge_debye
    sp = result.dipole_sp_hybrid_debye
    pd = result.dipole_pd_hybrid_debye
    charges = ",".join(_f12(float(q)) for q in result.charges)
    f = (f'"method":"{method}","energy_ev":{_f12(result.total_ev)},'
         f'"heat_of_formation_kcal":{_f12(result.heat_of_formation_kcal)},"charges":[{charges}]')
    f += (f',"dipole_debye":[{_f12(d[0])},{_f12(d[1])},{_f12(d[2])}]'
          f',"dipole_point_charge_debye":[{_f12(pc[0])},{_f12(pc[1])},{_f12(pc[2])}]'
          f',"dipole_sp_hybrid_debye":[{_f12(sp[0])},{_f12(sp[1])},{_f12(sp[2])}]'
          f',"dipole_pd_hybrid_debye":[{_f12(pd[0])},{_f12(pd[1])},{_f12(pd[2])}]'
          f',"orbital_source":"molecular"')
    for key in ("homo_ev", "lumo_ev", "homo_ev_beta", "lumo_ev_beta"):
        v = getattr(result, key)
        if v is not None:
            f += f',"{key}":{_f12(v)}'
    g = gap_ev(result)
    if g is not None:
        f += f',"gap_ev":{_f12(g)}'
    s2 = spin_squared(result)
    if s2 is not None:
        f += f',"spin_squared":{_f12(s2)}'
    return "{" + f + "}"


def energy_text(result: Pm7Result, method: str = "pm7") -> str:
    """The human-readable ``pm7_rs energy`` lines for a molecule (src/bin/pm7_rs.rs:442-470)."""
    d = result.dipole_debye
    lines = [
        f"method: {method}",
        f"total energy: {result.total_ev:.12f} eV",
        f"heat of formation: {result.heat_of_formation_kcal:.12f} kcal/mol",
        f"SCF iterations: {result.iterations}",
        f"dipole: {d[0]:.6f} {d[1]:.6f} {d[2]:.6f} Debye  (|mu| = {result.dipole_magnitude:.6f})",
    ]
    if result.homo_ev is not None and result.lumo_ev is not None:
        g = gap_ev(result)
        lines.append(f"HOMO / LUMO: {result.homo_ev:.6f} / {result.lumo_ev:.6f} eV  "
                     f"(gap {g if g is not None else float('nan'):.6f} eV, from the molecular orbitals)")
    s2 = spin_squared(result)
    if s2 is not None:
        nb = result.n_occ_beta if result.n_occ_beta is not None else result.n_occ
        sz = 0.5 * (result.n_occ - nb)
        lines.append(f"(SZ) = {sz:.6f}   <S^2> = {s2:.6f}   (exact {sz * (sz + 1.0):.6f})")
    return "\n".join(lines)
