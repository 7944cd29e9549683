# SPDX-License-Identifier: GPL-3.0-or-later
"""ctypes binding to the CPU FP64 oracle (``oracle/liboracle_pm7.so``).

Test infrastructure: imported by tests, by ``__graft_entry__.smoke`` and by ``bench.py``'s
``cpu_baseline`` leg only. The engine never loads it.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
LIB_PATH = ROOT / "oracle" / "liboracle_pm7.so"
DATA_DIR = str(ROOT / "pm7-rs_amd" / "data")

ANGSTROM_TO_BOHR = 1.0 / 0.529177210903

SYMBOLS = ["X", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P", "S", "Cl",
           "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As",
           "Se", "Br", "Kr", "Rb", "Sr", "Y", "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In",
           "Sn", "Sb", "Te", "I", "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb",
           "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W", "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl",
           "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac", "Th", "Pa", "U", "Np", "Pu", "Am", "Cm", "Bk",
           "Cf", "Es", "Fm", "Md", "No"]
Z_OF = {s: i for i, s in enumerate(SYMBOLS)}


class OracleOptions(C.Structure):
    _fields_ = [("method", C.c_int), ("multiplicity", C.c_int), ("max_scf", C.c_int),
                ("use_diis", C.c_int), ("accelerator", C.c_int), ("reference", C.c_int),
                ("threads", C.c_int), ("charge", C.c_double), ("e_tol", C.c_double),
                ("p_tol", C.c_double), ("adiis_switch", C.c_double), ("level_shift", C.c_double),
                ("force_dpath", C.c_int), ("smooth_dispersion", C.c_int), ("dipole_origin", C.c_int),
                ("pad_", C.c_int), ("initial_density", C.POINTER(C.c_double)),
                ("initial_spin_density", C.POINTER(C.c_double))]


class OracleResult(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("converged", C.c_int),
                ("unrestricted", C.c_int), ("nao", C.c_int), ("n_occ", C.c_int),
                ("n_occ_beta", C.c_int), ("n_atoms", C.c_int), ("electronic_ev", C.c_double),
                ("core_ev", C.c_double), ("total_ev", C.c_double), ("hof_kcal", C.c_double),
                ("correction_kcal", C.c_double), ("density_error", C.c_double),
                ("homo_ev", C.c_double), ("lumo_ev", C.c_double), ("homo_ev_beta", C.c_double),
                ("lumo_ev_beta", C.c_double), ("dipole", C.c_double * 3),
                ("dipole_point", C.c_double * 3), ("dipole_sp", C.c_double * 3),
                ("dipole_pd", C.c_double * 3), ("message", C.c_char * 256)]


class OracleDandcOptions(C.Structure):
    _fields_ = [("buffer_bohr", C.c_double), ("core_size", C.c_int), ("max_scf", C.c_int),
                ("fermi_width_ev", C.c_double), ("p_tol", C.c_double), ("mixing", C.c_double)]


class OracleDandcResult(C.Structure):
    _fields_ = [("status", C.c_int), ("converged", C.c_int), ("iterations", C.c_int), ("unrestricted", C.c_int),
                ("subsystems", C.c_int), ("largest_subsystem", C.c_int), ("largest_subsystem_nao", C.c_int),
                ("nao", C.c_int), ("stored_pairs", C.c_longlong), ("stored_elements", C.c_longlong),
                ("fermi_ev", C.c_double), ("electronic_ev", C.c_double), ("density_error", C.c_double),
                ("message", C.c_char * 256)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"oracle library missing: build it with `make -C oracle` ({LIB_PATH})")
        _lib = C.CDLL(str(LIB_PATH))
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        _lib.oracle_run.argtypes = [C.c_char_p, C.c_int, ip, dp, C.POINTER(OracleOptions),
                                    C.POINTER(OracleResult), dp, dp, dp, dp]
        _lib.oracle_nao.argtypes = [C.c_char_p, C.c_int, ip]
        _lib.oracle_hcore.argtypes = [C.c_char_p, C.c_int, ip, dp, dp]
        _lib.oracle_fock.argtypes = [C.c_char_p, C.c_int, ip, dp, dp, dp, dp]
        _lib.oracle_pair_block.argtypes = [C.c_char_p, C.c_int, C.c_int, dp, dp, dp, dp, dp]
        _lib.oracle_core_core.argtypes = [C.c_char_p, C.c_int, ip, dp]
        _lib.oracle_core_core.restype = C.c_double
        _lib.oracle_corrections.argtypes = [C.c_int, ip, dp, dp]
        _lib.oracle_corrections2.argtypes = [C.c_int, ip, dp, dp]
        _lib.oracle_eigh.argtypes = [C.c_int, dp, dp, dp]
        _lib.oracle_element.argtypes = [C.c_char_p, C.c_int, C.c_int, dp]
        _lib.oracle_default_options.argtypes = [C.POINTER(OracleOptions)]
        _lib.oracle_overlap_locals.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, C.c_double, C.c_double,
                                               C.c_double, C.c_int, dp]
        _lib.oracle_default_dandc_options.argtypes = [C.POINTER(OracleDandcOptions)]
        _lib.oracle_run_dandc.argtypes = [C.c_char_p, C.c_int, ip, dp, C.POINTER(OracleOptions),
                                          C.POINTER(OracleDandcOptions), C.POINTER(OracleDandcResult), dp, dp, dp]
        _lib.oracle_dandc_partition.argtypes = [C.c_int, dp, C.c_int, C.c_double, ip, ip, ip, ip, C.c_int]
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def default_options(**kw) -> OracleOptions:
    o = OracleOptions()
    lib().oracle_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def atoms_to_arrays(atoms):
    """[[symbol, x, y, z] (Angstrom)] -> (int32 z, float64 xyz_bohr[n,3])."""
    z = np.array([Z_OF[a[0]] for a in atoms], dtype=np.int32)
    xyz = np.array([[a[1], a[2], a[3]] for a in atoms], dtype=np.float64) * ANGSTROM_TO_BOHR
    return z, np.ascontiguousarray(xyz)


def nao(z) -> int:
    z = np.ascontiguousarray(z, dtype=np.int32)
    return lib().oracle_nao(DATA_DIR.encode(), len(z), _i(z))


def run(z, xyz_bohr, charge=0.0, multiplicity=1, initial_density=None, initial_spin_density=None, **kw):
    z = np.ascontiguousarray(z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
    n = len(z)
    m = max(nao(z), 1)
    o = default_options(charge=float(charge), multiplicity=int(multiplicity), **kw)
    keep = []
    for name, arr in (("initial_density", initial_density), ("initial_spin_density", initial_spin_density)):
        if arr is not None:
            a = np.ascontiguousarray(arr, dtype=np.float64)
            assert a.shape == (m, m), (name, a.shape)
            keep.append(a)
            setattr(o, name, _d(a))
    res = OracleResult()
    dens = np.zeros((m, m))
    spin = np.zeros((m, m))
    eps = np.zeros(m)
    q = np.zeros(n)
    lib().oracle_run(DATA_DIR.encode(), n, _i(z), _d(xyz), C.byref(o), C.byref(res), _d(dens), _d(spin),
                     _d(eps), _d(q))
    return res, dens, spin, eps, q


def hcore(z, xyz_bohr):
    z = np.ascontiguousarray(z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
    m = nao(z)
    h = np.zeros((m, m))
    st = lib().oracle_hcore(DATA_DIR.encode(), len(z), _i(z), _d(xyz), _d(h))
    if st:
        raise RuntimeError(f"oracle_hcore status {st}")
    return h


def fock(z, xyz_bohr, p_tot, p_spin):
    z = np.ascontiguousarray(z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
    m = nao(z)
    f = np.zeros((m, m))
    pt = np.ascontiguousarray(p_tot, dtype=np.float64)
    ps = np.ascontiguousarray(p_spin, dtype=np.float64)
    st = lib().oracle_fock(DATA_DIR.encode(), len(z), _i(z), _d(xyz), _d(pt), _d(ps), _d(f))
    if st:
        raise RuntimeError(f"oracle_fock status {st}")
    return f


def pair_block(za, zb, d_bohr):
    d = np.ascontiguousarray(d_bohr, dtype=np.float64)
    w = np.zeros(100)
    e1b = np.zeros(81)
    e2a = np.zeros(81)
    s = np.zeros(81)
    st = lib().oracle_pair_block(DATA_DIR.encode(), za, zb, _d(d), _d(w), _d(e1b), _d(e2a), _d(s))
    if st:
        raise RuntimeError(f"oracle_pair_block status {st}")
    return w, e1b.reshape(9, 9), e2a.reshape(9, 9), s.reshape(9, 9)


def core_core(z, xyz_bohr):
    z = np.ascontiguousarray(z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
    return lib().oracle_core_core(DATA_DIR.encode(), len(z), _i(z), _d(xyz))


def corrections2(z, xyz_bohr):
    """[smooth dispersion, PM7-HH H-H repulsion] in kcal/mol."""
    z = np.ascontiguousarray(z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
    out = np.zeros(2)
    lib().oracle_corrections2(len(z), _i(z), _d(xyz), _d(out))
    return out


def corrections(z, xyz_bohr):
    z = np.ascontiguousarray(z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
    out = np.zeros(4)
    lib().oracle_corrections(len(z), _i(z), _d(xyz), _d(out))
    return out


def eigh(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.shape[0]
    w = np.zeros(n)
    v = np.zeros((n, n))
    lib().oracle_eigh(n, _d(a), _d(w), _d(v))
    return w, v


def element(z, method=0):
    out = np.zeros(9)
    st = lib().oracle_element(DATA_DIR.encode(), method, z, _d(out))
    if st:
        raise RuntimeError(f"oracle_element status {st}")
    return out


def overlap_locals(na, zsa, zpa, nb, zsb, zpb, r, numeric=False):
    out = np.zeros(5)
    st = lib().oracle_overlap_locals(int(na), float(zsa), float(zpa), int(nb), float(zsb), float(zpb), float(r),
                                     1 if numeric else 0, _d(out))
    if st:
        raise RuntimeError("no closed form for these principal quantum numbers")
    return out


def dandc_partition(xyz_bohr, core_size, buffer):
    """[(members, n_core)] per subsystem (src/dandc/partition.rs:97-133)."""
    xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
    n = xyz.shape[0]
    ns = C.c_int(0)
    total = lib().oracle_dandc_partition(n, _d(xyz), int(core_size), float(buffer), C.byref(ns), None, None, None, 0)
    ncore = np.zeros(ns.value, np.int32)
    start = np.zeros(ns.value + 1, np.int32)
    members = np.zeros(max(total, 1), np.int32)
    lib().oracle_dandc_partition(n, _d(xyz), int(core_size), float(buffer), C.byref(ns), _i(ncore), _i(start),
                                 _i(members), len(members))
    return [(list(members[start[k]:start[k + 1]]), int(ncore[k])) for k in range(ns.value)]


def run_dandc(z, xyz_bohr, charge=0.0, multiplicity=1, buffer=15.0, core_size=12, max_scf=200, fermi_width_ev=0.1,
              p_tol=1.0e-6, mixing=0.4, **kw):
    """Divide-and-conquer SCF: (result, dense density, dense spin density, charges)."""
    z = np.ascontiguousarray(z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
    n = len(z)
    m = max(nao(z), 1)
    o = default_options(charge=float(charge), multiplicity=int(multiplicity), **kw)
    d = OracleDandcOptions(float(buffer), int(core_size), int(max_scf), float(fermi_width_ev), float(p_tol),
                           float(mixing))
    res = OracleDandcResult()
    dens = np.zeros((m, m))
    spin = np.zeros((m, m))
    q = np.zeros(n)
    lib().oracle_run_dandc(DATA_DIR.encode(), n, _i(z), _d(xyz), C.byref(o), C.byref(d), C.byref(res), _d(dens),
                           _d(spin), _d(q))
    return res, dens, spin, q
