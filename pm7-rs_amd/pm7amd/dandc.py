# SPDX-License-Identifier: GPL-3.0-or-later
"""Divide-and-conquer SCF, mirroring pm7-rs ``pm7_rs::dandc`` (src/dandc/mod.rs,
src/dandc/partition.rs, src/dandc/sparse.rs): ``DandcOptions``, ``run_dandc`` -> ``DandcResult``
with a ``SparseDensity`` of atom-pair blocks, ``partition`` -> ``Partitioning``. Runs on the HIP
engine through ``pm7_run_dandc`` / ``pm7_dandc_partition`` (include/pm7_amd.h)."""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import (Engine, Molecule, Pm7Error, Pm7Options, Pm7Parameters, METHODS, _Molecule, _Options,
               default_engine, library)


class _DandcOptions(C.Structure):
    _fields_ = [("buffer_bohr", C.c_double), ("core_size", C.c_int32), ("max_scf", C.c_int32),
                ("fermi_width_ev", C.c_double), ("p_tol", C.c_double), ("mixing", C.c_double)]


class _DandcResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("converged", C.c_int32), ("iterations", C.c_int32),
                ("unrestricted", C.c_int32), ("subsystems", C.c_int32), ("largest_subsystem", C.c_int32),
                ("largest_subsystem_nao", C.c_int32), ("n_atoms", C.c_int32), ("stored_pairs", C.c_int64),
                ("stored_elements", C.c_int64), ("fermi_ev", C.c_double), ("electronic_ev", C.c_double),
                ("density_error", C.c_double), ("device_ms", C.c_double), ("needed_mb", C.c_double),
                ("budget_mb", C.c_double), ("row_start", C.POINTER(C.c_int64)), ("columns", C.POINTER(C.c_int32)),
                ("offsets", C.POINTER(C.c_int64)), ("norb", C.POINTER(C.c_int32)),
                ("density", C.POINTER(C.c_double)), ("spin_density", C.POINTER(C.c_double)),
                ("charges", C.POINTER(C.c_double)), ("message", C.c_char * 512)]


class _Partition(C.Structure):
    _fields_ = [("n_sub", C.c_int32), ("n_core", C.POINTER(C.c_int32)), ("start", C.POINTER(C.c_int64)),
                ("members", C.POINTER(C.c_int32))]


_bound = False


def _lib():
    global _bound
    lib = library()
    if not _bound:
        vp = C.c_void_p
        lib.pm7_dandc_default_options.argtypes = [C.POINTER(_DandcOptions)]
        lib.pm7_dandc_default_options.restype = None
        lib.pm7_run_dandc.argtypes = [vp, C.POINTER(_Molecule), C.POINTER(_Options), C.POINTER(_DandcOptions),
                                      C.POINTER(_DandcResult)]
        lib.pm7_dandc_result_free.argtypes = [C.POINTER(_DandcResult)]
        lib.pm7_dandc_result_free.restype = None
        lib.pm7_dandc_partition.argtypes = [C.POINTER(_Molecule), C.c_int32, C.c_double, C.POINTER(_Partition)]
        lib.pm7_partition_free.argtypes = [C.POINTER(_Partition)]
        lib.pm7_partition_free.restype = None
        _bound = True
    return lib


@dataclass
class DandcOptions:
    """pm7-rs ``DandcOptions`` (src/dandc/mod.rs:57-89); ``buffer`` in Bohr."""

    buffer: float = 15.0
    core_size: int = 12
    fermi_width_ev: float = 0.1
    max_scf: int = 200
    p_tol: float = 1.0e-6
    mixing: float = 0.4

    def to_c(self) -> _DandcOptions:
        d = _DandcOptions()
        d.buffer_bohr = float(self.buffer)
        d.core_size = int(self.core_size)
        d.max_scf = int(self.max_scf)
        d.fermi_width_ev = float(self.fermi_width_ev)
        d.p_tol = float(self.p_tol)
        d.mixing = float(self.mixing)
        return d


class SparseDensity:
    """pm7-rs ``SparseDensity`` (src/dandc/sparse.rs): atom-pair blocks, CSR over atoms with sorted
    columns; block (a, b) is norb[a] x norb[b] row-major."""

    def __init__(self, row_start, columns, offsets, norb, data):
        self.row_start = row_start
        self.columns = columns
        self.offsets = offsets
        self.norb = norb
        self.data = data

    def n_atoms(self) -> int:
        return len(self.norb)

    def stored_pairs(self) -> int:
        return len(self.columns)

    def stored_elements(self) -> int:
        return len(self.data)

    def pair_slot(self, a: int, b: int) -> Optional[int]:
        lo, hi = int(self.row_start[a]), int(self.row_start[a + 1])
        k = lo + int(np.searchsorted(self.columns[lo:hi], b))
        return k if k < hi and self.columns[k] == b else None

    def block(self, a: int, b: int) -> np.ndarray:
        k = self.pair_slot(a, b)
        if k is None:
            return np.zeros((0, 0))
        return self.data[self.offsets[k]:self.offsets[k + 1]].reshape(int(self.norb[a]), int(self.norb[b]))

    def population(self, a: int) -> float:
        blk = self.block(a, a)
        return float(np.trace(blk)) if blk.size else 0.0

    def trace(self) -> float:
        return sum(self.population(a) for a in range(self.n_atoms()))

    def to_dense(self) -> np.ndarray:
        off = np.concatenate([[0], np.cumsum(self.norb)]).astype(np.int64)
        m = np.zeros((off[-1], off[-1]))
        for a in range(self.n_atoms()):
            for k in range(int(self.row_start[a]), int(self.row_start[a + 1])):
                b = int(self.columns[k])
                m[off[a]:off[a + 1], off[b]:off[b + 1]] = self.data[self.offsets[k]:self.offsets[k + 1]].reshape(
                    int(self.norb[a]), int(self.norb[b]))
        return m


@dataclass
class DandcResult:
    """pm7-rs ``DandcResult`` (src/dandc/mod.rs:117-135) plus the Mulliken charges
    (``mulliken_charges``, src/dandc/mod.rs:823-835) and the device time of the SCF steps."""

    density: SparseDensity
    spin_density: Optional[SparseDensity]
    fermi_ev: float
    electronic_ev: float
    converged: bool
    iterations: int
    density_error: float
    subsystems: int
    largest_subsystem: int
    largest_subsystem_nao: int
    unrestricted: bool
    charges: np.ndarray
    device_ms: float


@dataclass
class Subsystem:
    """pm7-rs ``Subsystem`` (src/dandc/partition.rs:17-67): core atoms first."""

    parent: List[int]
    n_core: int

    def __len__(self) -> int:
        return len(self.parent)

    def weight(self, i: int, j: int) -> float:
        return {(True, True): 1.0, (True, False): 0.5, (False, True): 0.5}.get((i < self.n_core, j < self.n_core),
                                                                                0.0)


@dataclass
class Partitioning:
    subsystems: List[Subsystem]

    def largest(self) -> int:
        return max((len(s) for s in self.subsystems), default=0)

    def mean_size(self) -> float:
        return float(np.mean([len(s) for s in self.subsystems])) if self.subsystems else 0.0


def partition(molecule: Molecule, core_size: int, buffer: float) -> Partitioning:
    """pm7-rs ``partition`` (src/dandc/partition.rs:97-133), molecular case; host only."""
    lib = _lib()
    z, xyz = molecule.arrays()
    cm = _Molecule(len(z), z.ctypes.data_as(C.POINTER(C.c_int32)), xyz.ctypes.data_as(C.POINTER(C.c_double)))
    p = _Partition()
    st = lib.pm7_dandc_partition(C.byref(cm), int(core_size), float(buffer), C.byref(p))
    if st != 0:
        raise Pm7Error(st, "invalid partition request")
    try:
        ns = p.n_sub
        start = np.ctypeslib.as_array(p.start, shape=(ns + 1,)).copy()
        n_core = np.ctypeslib.as_array(p.n_core, shape=(ns,)).copy()
        members = np.ctypeslib.as_array(p.members, shape=(int(start[-1]),)).copy() if start[-1] else np.zeros(0, int)
    finally:
        lib.pm7_partition_free(C.byref(p))
    return Partitioning([Subsystem([int(x) for x in members[start[s]:start[s + 1]]], int(n_core[s]))
                         for s in range(ns)])


def _run(engine: Engine, molecule: Molecule, options: Pm7Options, dandc: DandcOptions) -> DandcResult:
    lib = _lib()
    z, xyz = molecule.arrays()
    cm = _Molecule(len(z), z.ctypes.data_as(C.POINTER(C.c_int32)), xyz.ctypes.data_as(C.POINTER(C.c_double)))
    co = options.to_c()
    cd = dandc.to_c()
    r = _DandcResult()
    st = lib.pm7_run_dandc(engine._h, C.byref(cm), C.byref(co), C.byref(cd), C.byref(r))
    if st != 0:
        err = Pm7Error(st, r.message.decode(errors="replace"))
        err.needed_mb, err.budget_mb = r.needed_mb, r.budget_mb
        err.subsystems, err.largest_subsystem = r.subsystems, r.largest_subsystem
        err.largest_subsystem_nao = r.largest_subsystem_nao
        raise err
    try:
        n = r.n_atoms
        slots, nel = int(r.stored_pairs), int(r.stored_elements)
        row_start = np.ctypeslib.as_array(r.row_start, shape=(n + 1,)).copy()
        columns = np.ctypeslib.as_array(r.columns, shape=(slots,)).copy() if slots else np.zeros(0, np.int32)
        offsets = np.ctypeslib.as_array(r.offsets, shape=(slots + 1,)).copy()
        norb = np.ctypeslib.as_array(r.norb, shape=(n,)).copy()
        data = np.ctypeslib.as_array(r.density, shape=(nel,)).copy()
        spin = np.ctypeslib.as_array(r.spin_density, shape=(nel,)).copy() if r.spin_density else None
        charges = np.ctypeslib.as_array(r.charges, shape=(n,)).copy()
    finally:
        lib.pm7_dandc_result_free(C.byref(r))
    dens = SparseDensity(row_start, columns, offsets, norb, data)
    sdens = SparseDensity(row_start, columns, offsets, norb, spin) if spin is not None else None
    return DandcResult(dens, sdens, r.fermi_ev, r.electronic_ev, bool(r.converged), r.iterations, r.density_error,
                       r.subsystems, r.largest_subsystem, r.largest_subsystem_nao, bool(r.unrestricted), charges,
                       r.device_ms)


def run_dandc(molecule: Molecule, params: Optional[Pm7Parameters] = None, options: Optional[Pm7Options] = None,
              dandc: Optional[DandcOptions] = None, engine: Optional[Engine] = None) -> DandcResult:
    """Drop-in for pm7-rs ``run_dandc`` (src/dandc/mod.rs:138-143): raises Pm7Error like ``Err(..)``;
    an unconverged run returns with ``converged`` False, as the reference does."""
    options = options or Pm7Options(charge=molecule.charge, multiplicity=molecule.multiplicity)
    if params is not None and METHODS[params.method.lower()] != METHODS[options.method.lower()]:
        raise Pm7Error(1, f"options method `{options.method}` does not match parameter method `{params.method}`")
    eng = engine or default_engine()
    if METHODS[options.method.lower()] != eng.method:
        eng.load(options.method)
    return _run(eng, molecule, options, dandc or DandcOptions())


def mulliken_charges(result: DandcResult) -> np.ndarray:
    """pm7-rs ``dandc::mulliken_charges`` (src/dandc/mod.rs:823-835)."""
    return result.charges


__all__ = ["DandcOptions", "DandcResult", "SparseDensity", "Subsystem", "Partitioning", "partition", "run_dandc",
           "mulliken_charges"]
