# SPDX-License-Identifier: GPL-3.0-or-later
"""pm7amd: the PM7 NDDO SCF single point on one MI355X, mirroring pm7-rs's API.

The Python surface follows pm7-rs names and semantics (``Molecule``, ``Pm7Parameters``,
``Pm7Options``, ``run_pm7`` -> ``Pm7Result``, ``Pm7Error``; src/scf.rs:73-342, 512-867) and calls
the HIP engine through the C ABI declared in ``include/pm7_amd.h``. There is no CPU fallback: if
the library or a HIP device is missing, every call raises ``Pm7Error``.
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libpm7amd.so"
DATA_DIR = _HERE.parent / "data"

ANGSTROM_TO_BOHR = 1.0 / 0.529177210903
EV_TO_KCAL = 23.06054783061903

STATUS_NAMES = {
    0: "Ok", 1: "InvalidInput", 2: "MissingElement", 3: "MissingParameter", 4: "UnsupportedFeature",
    5: "LinearAlgebra", 6: "ScfNotConverged", 7: "InsufficientMemory", 8: "Device", 9: "NotReady",
}

METHODS = {"pm7": 0, "pm7-ts": 1, "pm7ts": 1, "pm7-sparkle": 2, "pm7+sparkle": 2, "sparkle": 2,
           "pm7-minus": 3, "pm7-": 3, "pm7minus": 3, "pm7-hh": 4, "pm7hh": 4}
DIPOLE_ORIGINS = {"centre-of-mass": 0, "center-of-mass": 0, "com": 0, "coordinates": 1, "origin": 1,
                  "centre-of-charge": 2, "center-of-charge": 2}
ACCELERATORS = {"none": 0, "cdiis": 1, "adiis-cdiis": 2, "adiis_cdiis": 2, "hybrid": 2}
REFERENCES = {"auto": 0, "rhf": 1, "restricted": 1, "r": 1, "uhf": 2, "unrestricted": 2, "u": 2}
PRECISIONS = {"fp64": 0, "mixed": 1}

SYMBOLS = ["X", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P", "S", "Cl",
           "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As",
           "Se", "Br", "Kr", "Rb", "Sr", "Y", "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In",
           "Sn", "Sb", "Te", "I", "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb",
           "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W", "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl",
           "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac", "Th", "Pa", "U", "Np", "Pu", "Am", "Cm", "Bk",
           "Cf", "Es", "Fm", "Md", "No"]
Z_OF = {s.lower(): i for i, s in enumerate(SYMBOLS)}


class Pm7Error(RuntimeError):
    """Mirror of pm7-rs ``Pm7Error`` (src/error.rs:13-48): ``kind`` is the variant name."""

    def __init__(self, status: int, message: str, result: Optional["Pm7Result"] = None):
        self.status = int(status)
        self.kind = STATUS_NAMES.get(self.status, f"Status{status}")
        self.result = result
        super().__init__(f"{self.kind}: {message}")


# ---- C ABI structs (include/pm7_amd.h) -----------------------------------------------------

class _Molecule(C.Structure):
    _fields_ = [("n_atoms", C.c_int32), ("z", C.POINTER(C.c_int32)), ("xyz_bohr", C.POINTER(C.c_double))]


class _ServerConfig(C.Structure):
    _fields_ = [("lanes", C.c_int32), ("max_batch", C.c_int32), ("max_memory_mb", C.c_int64)]


class _Options(C.Structure):
    _fields_ = [("method", C.c_int32), ("charge", C.c_double), ("multiplicity", C.c_int32),
                ("max_scf", C.c_int32), ("e_tol", C.c_double), ("p_tol", C.c_double),
                ("use_diis", C.c_int32), ("accelerator", C.c_int32), ("adiis_switch", C.c_double),
                ("level_shift", C.c_double), ("reference", C.c_int32), ("precision", C.c_int32),
                ("mixed_switch", C.c_double), ("max_memory_mb", C.c_int64), ("force_dpath", C.c_int32),
                ("smooth_dispersion", C.c_int32), ("dipole_origin", C.c_int32),
                ("initial_density_nao", C.c_int32), ("initial_density", C.POINTER(C.c_double)),
                ("initial_spin_density", C.POINTER(C.c_double))]


class _Result(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int32), ("converged", C.c_int32),
                ("unrestricted", C.c_int32), ("nao", C.c_int32), ("n_atoms", C.c_int32),
                ("n_occ", C.c_int32), ("n_occ_beta", C.c_int32), ("fp32_iterations", C.c_int32),
                ("jacobi_sweeps", C.c_int32), ("electronic_ev", C.c_double), ("core_ev", C.c_double),
                ("total_ev", C.c_double), ("heat_of_formation_kcal", C.c_double),
                ("correction_kcal", C.c_double), ("density_error", C.c_double),
                ("homo_ev", C.c_double), ("lumo_ev", C.c_double), ("homo_ev_beta", C.c_double),
                ("lumo_ev_beta", C.c_double), ("dipole_debye", C.c_double * 3),
                ("dipole_point_charge_debye", C.c_double * 3), ("dipole_sp_hybrid_debye", C.c_double * 3),
                ("dipole_pd_hybrid_debye", C.c_double * 3), ("dipole_magnitude", C.c_double),
                ("device_phase_ms", C.c_double * 6), ("eig_phase_ms", C.c_double * 3),
                ("err_vs_fp64_ev", C.c_double), ("err_vs_fp64_density", C.c_double),
                ("missing_element", C.c_int32), ("reserved_", C.c_int32), ("needed_mb", C.c_double),
                ("budget_mb", C.c_double), ("density", C.POINTER(C.c_double)), ("spin_density", C.POINTER(C.c_double)),
                ("mo_energies", C.POINTER(C.c_double)), ("mo_coeff", C.POINTER(C.c_double)),
                ("mo_energies_beta", C.POINTER(C.c_double)), ("mo_coeff_beta", C.POINTER(C.c_double)),
                ("charges", C.POINTER(C.c_double)), ("message", C.c_char * 256)]


TIMING_SLOTS = 14

ABI_FUNCTIONS = [
    "pm7_ctx_create", "pm7_ctx_destroy", "pm7_last_error", "pm7_abi_version", "pm7_params_load",
    "pm7_default_options", "pm7_run", "pm7_run_batch", "pm7_result_free", "pm7_core_hamiltonian",
    "pm7_fock_spin", "pm7_core_core_energy", "pm7_last_timings", "pm7_server_default_config",
    "pm7_server_start", "pm7_server_submit", "pm7_server_poll", "pm7_server_stop", "pm7_gradient_batch",
    "pm7_gradient", "pm7_response_fock_batch", "pm7_response_fock_spin_batch", "pm7_fock_spin_noalloc",
    "pm7_dandc_default_options", "pm7_run_dandc", "pm7_dandc_result_free", "pm7_dandc_partition",
    "pm7_partition_free", "pm7_batch_submit", "pm7_batch_wait", "pm7_autotune",
]

_lib = None


class _TuneResult(C.Structure):
    _fields_ = [("eig_block_mask", C.c_int32), ("mixed_sweeps", C.c_int32), ("batches_in_flight", C.c_int32),
                ("candidates", C.c_int32), ("solves_per_s", C.c_double), ("step_ms", C.c_double),
                ("default_solves_per_s", C.c_double), ("tune_ms", C.c_double)]


def library() -> C.CDLL:
    """Load the engine library (raises Pm7Error when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise Pm7Error(8, f"engine library missing: build it with `make -C pm7-rs_amd` ({LIB_PATH})")
    lib = C.CDLL(str(LIB_PATH))
    vp = C.c_void_p
    lib.pm7_ctx_create.argtypes = [C.c_int32, C.c_char_p, C.POINTER(vp)]
    lib.pm7_ctx_destroy.argtypes = [vp]
    lib.pm7_last_error.argtypes = [vp]
    lib.pm7_last_error.restype = C.c_char_p
    lib.pm7_params_load.argtypes = [vp, C.c_int32]
    lib.pm7_default_options.argtypes = [C.POINTER(_Options)]
    lib.pm7_run.argtypes = [vp, C.POINTER(_Molecule), C.POINTER(_Options), C.POINTER(_Result)]
    lib.pm7_run_batch.argtypes = [vp, C.c_int32, C.POINTER(_Molecule), C.POINTER(_Options), C.c_int32,
                                  C.POINTER(_Result)]
    lib.pm7_result_free.argtypes = [C.POINTER(_Result)]
    lib.pm7_gradient_batch.argtypes = [vp, C.c_int32, C.POINTER(_Molecule), C.POINTER(_Options), C.c_int32,
                                       C.POINTER(_Result), C.POINTER(C.c_double)]
    lib.pm7_gradient.argtypes = [vp, C.POINTER(_Molecule), C.POINTER(_Options), C.POINTER(_Result),
                                 C.POINTER(C.c_double)]
    lib.pm7_core_hamiltonian.argtypes = [vp, C.POINTER(_Molecule), C.POINTER(C.c_double), C.POINTER(C.c_int32)]
    lib.pm7_autotune.argtypes = [vp, C.c_int32, C.POINTER(_Molecule), C.POINTER(_Options), C.c_int32, C.c_int32,
                                 C.c_double, C.c_char_p, C.POINTER(_TuneResult)]
    lib.pm7_fock_spin.argtypes = [vp, C.POINTER(_Molecule), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]
    lib.pm7_fock_spin_noalloc.argtypes = [vp, C.POINTER(_Molecule), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                          C.POINTER(C.c_double)]
    lib.pm7_core_core_energy.argtypes = [vp, C.POINTER(_Molecule), C.POINTER(C.c_double)]
    lib.pm7_response_fock_batch.argtypes = [vp, C.POINTER(_Molecule), C.c_int32, C.POINTER(C.c_double),
                                            C.POINTER(C.c_double)]
    lib.pm7_response_fock_spin_batch.argtypes = [vp, C.POINTER(_Molecule), C.c_int32, C.POINTER(C.c_double),
                                                 C.POINTER(C.c_double), C.POINTER(C.c_double),
                                                 C.POINTER(C.c_double)]
    lib.pm7_last_timings.argtypes = [vp, C.POINTER(C.c_double), C.c_int32]
    lib.pm7_server_default_config.argtypes = [C.POINTER(_ServerConfig)]
    lib.pm7_server_start.argtypes = [vp, C.POINTER(_ServerConfig), C.POINTER(vp)]
    lib.pm7_server_submit.argtypes = [vp, C.POINTER(_Molecule), C.POINTER(_Options), C.POINTER(C.c_int64)]
    lib.pm7_server_poll.argtypes = [vp, C.c_int64, C.c_int32, C.POINTER(C.c_int64), C.POINTER(_Result),
                                    C.POINTER(C.c_double)]
    lib.pm7_batch_submit.argtypes = [vp, C.c_int32, C.POINTER(_Molecule), C.POINTER(_Options), C.c_int32,
                                     C.POINTER(_Result), C.POINTER(C.c_int64)]
    lib.pm7_batch_wait.argtypes = [vp, C.c_int64]
    lib.pm7_server_stop.argtypes = [vp]
    lib.pm7_server_stop.restype = None
    _lib = lib
    return lib


# ---- reference-shaped Python types ---------------------------------------------------------

@dataclass
class Atom:
    z: int
    position: tuple  # Bohr


class Molecule:
    """pm7-rs ``Molecule`` (src/system.rs:20-39): atoms with Bohr positions.

    Held as the two arrays the C-ABI takes (atomic numbers int32[n], positions float64[n, 3] in
    Bohr), so handing a molecule to the engine copies nothing; ``atoms`` is a derived view."""

    __slots__ = ("z", "xyz", "charge", "multiplicity")

    def __init__(self, atoms=(), charge: float = 0.0, multiplicity: int = 1, *, z=None, xyz=None):
        if z is None:
            atoms = list(atoms)
            z = [a.z for a in atoms]
            xyz = [a.position for a in atoms]
        self.z = np.ascontiguousarray(z, dtype=np.int32).reshape(-1)
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        if self.xyz.shape[0] != self.z.shape[0]:
            raise Pm7Error(1, "atomic numbers and positions differ in length")
        self.charge = charge
        self.multiplicity = multiplicity

    @classmethod
    def from_arrays(cls, z, xyz_bohr, charge: float = 0.0, multiplicity: int = 1) -> "Molecule":
        return cls(charge=charge, multiplicity=multiplicity, z=z, xyz=xyz_bohr)

    @property
    def atoms(self):
        return [Atom(int(zz), tuple(float(v) for v in p)) for zz, p in zip(self.z, self.xyz)]

    @classmethod
    def from_symbols_angstrom(cls, atoms: Sequence, charge: float = 0.0, multiplicity: int = 1) -> "Molecule":
        z = [Z_OF[sym.lower()] if isinstance(sym, str) else int(sym) for sym, _, _, _ in atoms]
        xyz = np.array([(x, y, zz) for _, x, y, zz in atoms], dtype=np.float64).reshape(-1, 3) * ANGSTROM_TO_BOHR
        return cls(charge=charge, multiplicity=multiplicity, z=z, xyz=xyz)

    @classmethod
    def from_xyz_str(cls, text: str, charge: float = 0.0) -> "Molecule":
        """pm7-rs ``Molecule::from_xyz_str`` (src/system.rs:149-196); see ``formats.read_xyz``."""
        from .formats import read_xyz
        return read_xyz(text, charge)

    def to_xyz_string(self, title: str = "") -> str:
        from .formats import to_xyz_string
        return to_xyz_string(self, title)

    def arrays(self):
        return self.z, self.xyz


@dataclass
class Pm7Parameters:
    """Selects the bundled parameter tables (pm7-rs ``Pm7Parameters::method``)."""

    method: str = "pm7"

    @classmethod
    def standard(cls) -> "Pm7Parameters":
        return cls("pm7")


@dataclass
class Pm7Options:
    """pm7-rs ``Pm7Options`` (src/scf.rs:73-234) plus the engine's precision switch."""

    method: str = "pm7"
    charge: float = 0.0
    multiplicity: int = 1
    max_scf: int = 200
    e_tol: float = 1.0e-8
    p_tol: float = 1.0e-7
    use_diis: bool = True
    accelerator: str = "adiis-cdiis"
    adiis_switch: float = 0.1
    level_shift: float = 0.0
    reference: str = "auto"
    precision: str = "fp64"
    mixed_switch: float = 1.0e-3
    max_memory_mb: Optional[int] = None
    force_dpath: bool = False
    smooth_dispersion: bool = False
    dipole_origin: str = "centre-of-mass"
    initial_density: Optional[np.ndarray] = None
    initial_spin_density: Optional[np.ndarray] = None

    def to_c(self) -> _Options:
        o = _Options()
        library().pm7_default_options(C.byref(o))
        o.method = METHODS[self.method.lower()]
        o.charge = float(self.charge)
        o.multiplicity = int(self.multiplicity)
        o.max_scf = int(self.max_scf)
        o.e_tol = float(self.e_tol)
        o.p_tol = float(self.p_tol)
        o.use_diis = 1 if self.use_diis else 0
        o.accelerator = ACCELERATORS[self.accelerator.lower()]
        o.adiis_switch = float(self.adiis_switch)
        o.level_shift = float(self.level_shift)
        o.reference = REFERENCES[self.reference.lower()]
        o.precision = PRECISIONS[self.precision.lower()]
        o.mixed_switch = float(self.mixed_switch)
        o.max_memory_mb = int(self.max_memory_mb or 0)
        o.force_dpath = 1 if self.force_dpath else 0
        o.smooth_dispersion = 1 if self.smooth_dispersion else 0
        o.dipole_origin = DIPOLE_ORIGINS[self.dipole_origin.lower()]
        keep = []
        if self.initial_density is not None:
            d = np.ascontiguousarray(self.initial_density, dtype=np.float64)
            if d.ndim != 2 or d.shape[0] != d.shape[1]:
                raise Pm7Error(1, f"the supplied initial density has shape {d.shape}")
            o.initial_density_nao = d.shape[0]
            o.initial_density = d.ctypes.data_as(C.POINTER(C.c_double))
            keep.append(d)
        if self.initial_spin_density is not None:
            sd = np.ascontiguousarray(self.initial_spin_density, dtype=np.float64)
            if sd.ndim != 2 or sd.shape[0] != sd.shape[1] or (self.initial_density is not None
                                                                 and sd.shape != keep[0].shape):
                raise Pm7Error(1, f"the supplied initial spin density has shape {sd.shape}")
            if self.initial_density is None:
                o.initial_density_nao = sd.shape[0]
            o.initial_spin_density = sd.ctypes.data_as(C.POINTER(C.c_double))
            keep.append(sd)
        o._keep = keep  # the arrays must outlive the call that reads them
        return o


@dataclass
class Pm7Result:
    """pm7-rs ``Pm7Result`` (src/scf.rs:262-342), molecular fields."""

    status: int
    density: np.ndarray
    spin_density: Optional[np.ndarray]
    mo_energies: np.ndarray
    mo_coeff: np.ndarray
    n_occ: int
    mo_energies_beta: Optional[np.ndarray]
    mo_coeff_beta: Optional[np.ndarray]
    n_occ_beta: Optional[int]
    electronic_ev: float
    core_ev: float
    total_ev: float
    heat_of_formation_kcal: float
    correction_kcal: float
    charges: np.ndarray
    dipole_debye: np.ndarray
    dipole_magnitude: float
    dipole_point_charge_debye: np.ndarray
    dipole_sp_hybrid_debye: np.ndarray
    dipole_pd_hybrid_debye: np.ndarray
    homo_ev: Optional[float]
    lumo_ev: Optional[float]
    homo_ev_beta: Optional[float]
    lumo_ev_beta: Optional[float]
    iterations: int
    converged: bool
    unrestricted: bool
    density_error: float
    fp32_iterations: int = 0
    jacobi_sweeps: int = 0
    device_phase_ms: tuple = ()
    eig_phase_ms: tuple = ()
    message: str = ""
    err_vs_fp64_ev: float = 0.0
    err_vs_fp64_density: float = 0.0
    missing_element: int = 0
    needed_mb: float = 0.0
    budget_mb: float = 0.0

    orbital_source: str = "molecular"

    @property
    def gap_ev(self) -> Optional[float]:
        """pm7-rs ``Pm7Result::gap_ev`` (src/scf.rs:462-472), resolved over both spin channels."""
        from .formats import gap_ev
        return gap_ev(self)

    def spin_squared(self) -> Optional[float]:
        """pm7-rs ``Pm7Result::spin_squared`` (src/scf.rs:423-437); None for restricted runs."""
        from .formats import spin_squared
        return spin_squared(self)


def _opt(x: float) -> Optional[float]:
    return None if math.isnan(x) else float(x)


class _ResultArrays:
    """Owns one result's library-allocated arrays; frees them when the last numpy view goes."""

    __slots__ = ("_free", "_res")

    def __init__(self, free, res):
        self._free = free
        self._res = res

    def __del__(self):
        self._free(C.byref(self._res))


def _view(owner, addr, k):
    """Zero-copy numpy view of k doubles that the library allocated at `addr` (None for 0)."""
    if not addr:
        return None
    buf = (C.c_double * k).from_address(addr)
    buf._owner = owner
    return np.frombuffer(buf, dtype=np.float64)


def _struct_dtype(cls):
    """numpy dtype with the exact field offsets of a ctypes Structure (pointers as addresses)."""
    names, formats, offsets = [], [], []
    for name, tp in cls._fields_:
        if issubclass(tp, C._Pointer):
            fmt = np.uintp
        elif issubclass(tp, C.Array):
            fmt = ("S%d" % tp._length_) if tp._type_ is C.c_char else (np.dtype(tp._type_), (tp._length_,))
        else:
            fmt = np.dtype(tp)
        names.append(name)
        formats.append(fmt)
        offsets.append(getattr(cls, name).offset)
    return np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": C.sizeof(cls)})


_RESULT_DTYPE = _struct_dtype(_Result)


def _convert_all(c_res, count: int, free) -> list:
    """Pm7Results over a filled `_Result` array without copying the matrices: every scalar field of
    the batch is read in one pass; each result's arrays are views kept alive by an owner that
    calls `free` (pm7_result_free) once no view of that result is referenced any more."""
    cols = np.frombuffer(c_res, dtype=_RESULT_DTYPE, count=count)
    col = {name: cols[name].tolist() for name in _RESULT_DTYPE.names}
    out = []
    for i in range(count):
        n = col["nao"][i]
        na = col["n_atoms"][i]
        owner = _ResultArrays(free, c_res[i])
        dens = col["density"][i]
        nob = col["n_occ_beta"][i]

        def mat(addr, rows=n, cols_=n):
            a = _view(owner, addr, rows * cols_)
            return None if a is None else a.reshape(rows, cols_)

        out.append(Pm7Result(
            status=col["status"][i], density=mat(dens) if dens else np.zeros((0, 0)),
            spin_density=mat(col["spin_density"][i]),
            mo_energies=_view(owner, col["mo_energies"][i], n) if col["mo_energies"][i] else np.zeros(0),
            mo_coeff=mat(col["mo_coeff"][i]) if col["mo_coeff"][i] else np.zeros((0, 0)),
            n_occ=col["n_occ"][i], mo_energies_beta=_view(owner, col["mo_energies_beta"][i], n),
            mo_coeff_beta=mat(col["mo_coeff_beta"][i]), n_occ_beta=nob if nob >= 0 else None,
            electronic_ev=col["electronic_ev"][i], core_ev=col["core_ev"][i], total_ev=col["total_ev"][i],
            heat_of_formation_kcal=col["heat_of_formation_kcal"][i], correction_kcal=col["correction_kcal"][i],
            charges=_view(owner, col["charges"][i], na) if col["charges"][i] else np.zeros(0),
            dipole_debye=np.array(col["dipole_debye"][i]), dipole_magnitude=col["dipole_magnitude"][i],
            dipole_point_charge_debye=np.array(col["dipole_point_charge_debye"][i]),
            dipole_sp_hybrid_debye=np.array(col["dipole_sp_hybrid_debye"][i]),
            dipole_pd_hybrid_debye=np.array(col["dipole_pd_hybrid_debye"][i]),
            homo_ev=_opt(col["homo_ev"][i]), lumo_ev=_opt(col["lumo_ev"][i]),
            homo_ev_beta=_opt(col["homo_ev_beta"][i]), lumo_ev_beta=_opt(col["lumo_ev_beta"][i]),
            iterations=col["iterations"][i], converged=bool(col["converged"][i]),
            unrestricted=bool(col["unrestricted"][i]), density_error=col["density_error"][i],
            fp32_iterations=col["fp32_iterations"][i], jacobi_sweeps=col["jacobi_sweeps"][i],
            device_phase_ms=tuple(col["device_phase_ms"][i]), eig_phase_ms=tuple(col["eig_phase_ms"][i]),
            message=col["message"][i].decode(errors="replace"),
            err_vs_fp64_ev=col["err_vs_fp64_ev"][i], err_vs_fp64_density=col["err_vs_fp64_density"][i],
            missing_element=col["missing_element"][i], needed_mb=col["needed_mb"][i],
            budget_mb=col["budget_mb"][i],
        ))
    return out


def _convert(r: _Result, free) -> Pm7Result:
    """One result (see _convert_all)."""
    arr = (_Result * 1)()
    C.memmove(arr, C.byref(r), C.sizeof(_Result))
    return _convert_all(arr, 1, free)[0]


class Engine:
    """A device context (pm7_ctx): device, streams, parameter tables, pooled HBM."""

    def __init__(self, device: int = 0, data_dir: Optional[str] = None, method: str = "pm7"):
        lib = library()
        h = C.c_void_p()
        st = lib.pm7_ctx_create(int(device), str(data_dir or DATA_DIR).encode(), C.byref(h))
        if st != 0:
            raise Pm7Error(st, f"pm7_ctx_create failed on device {device}")
        self._h = h
        self._lib = lib
        self.method = 0  # pm7_ctx_create loads the PM7 tables
        if METHODS[method.lower()] != 0:
            self.load(method)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pm7_ctx_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - interpreter teardown order varies
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def last_error(self) -> str:
        return self._lib.pm7_last_error(self._h).decode(errors="replace")

    def load(self, method: str):
        """pm7-rs ``Pm7Parameters::method`` (src/params.rs:130-155) on this context."""
        code = METHODS[method.lower()]
        st = self._lib.pm7_params_load(self._h, code)
        if st != 0:
            raise Pm7Error(st, self.last_error())
        self.method = code

    @staticmethod
    def _mol(m: Molecule):
        z, xyz = m.arrays()
        cm = _Molecule(len(z), z.ctypes.data_as(C.POINTER(C.c_int32)), xyz.ctypes.data_as(C.POINTER(C.c_double)))
        return cm, (z, xyz)

    def run(self, molecule: Molecule, options: Optional[Pm7Options] = None, raise_on_error: bool = True) -> Pm7Result:
        return self.run_batch([molecule], options, raise_on_error=raise_on_error)[0]

    def run_batch(self, molecules: Sequence[Molecule], options=None, raise_on_error: bool = False):
        """Many molecules in one device pass; ``options`` is one Pm7Options or one per molecule."""
        return self._batch(molecules, options, raise_on_error, None)

    def gradient(self, molecule: Molecule, options: Optional[Pm7Options] = None, raise_on_error: bool = True):
        """pm7-rs ``closed_form_gradient`` (src/gradient.rs:500-555): (result, gradient[n_atoms, 3]
        in eV/Bohr at the converged density, geometric corrections included)."""
        res, grads = self.gradient_batch([molecule], options, raise_on_error)
        return res[0], grads[0]

    def gradient_batch(self, molecules: Sequence[Molecule], options=None, raise_on_error: bool = False):
        """Batched SCF + gradients: (results, [gradient[n_atoms_i, 3]]); NaN rows and status
        UnsupportedFeature for a molecule whose SCF left no density."""
        flat = np.full(3 * sum(len(m.z) for m in molecules), np.nan)
        res = self._batch(molecules, options, raise_on_error, flat)
        grads, off = [], 0
        for m in molecules:
            grads.append(flat[off:off + 3 * len(m.z)].reshape(-1, 3))
            off += 3 * len(m.z)
        return res, grads

    def autotune(self, molecules: Sequence[Molecule], options=None, steps: int = 4, latency_budget: float = 2.0,
                 cache_path: Optional[str] = None) -> dict:
        """pm7_autotune: time this context's launch shapes on the given batch (eigensolver per size
        class, reduced-precision sweep cap, batches in flight) and keep the fastest; returns the
        choice and its measured throughput."""
        n = len(molecules)
        if options is None:
            options = Pm7Options()
        opts_list = options if isinstance(options, (list, tuple)) else [options]
        c_mols = (_Molecule * n)()
        keep = []
        for i, m in enumerate(molecules):
            cm, k = self._mol(m)
            c_mols[i] = cm
            keep.append(k)
        c_list = [o.to_c() for o in opts_list]
        c_opts = (_Options * len(opts_list))(*c_list)
        out = _TuneResult()
        st = self._lib.pm7_autotune(self._h, n, c_mols, c_opts, len(opts_list), int(steps), float(latency_budget),
                                    cache_path.encode() if cache_path else None, C.byref(out))
        if st != 0:
            raise Pm7Error(st, self.last_error())
        return {f: getattr(out, f) for f, _ in _TuneResult._fields_}

    def _batch(self, molecules, options, raise_on_error, grad_flat):
        n = len(molecules)
        if n == 0:
            return []
        if options is None:
            options = Pm7Options()
        opts_list = options if isinstance(options, (list, tuple)) else [options]
        c_mols = (_Molecule * n)()
        keep = []
        for i, m in enumerate(molecules):
            cm, k = self._mol(m)
            c_mols[i] = cm
            keep.append(k)
        # One method for the whole call selects the matching tables, as a caller of run_pm7 passes
        # Pm7Parameters::method(options.method); mixed methods are checked per molecule.
        methods = {METHODS[o.method.lower()] for o in opts_list}
        if len(methods) == 1 and next(iter(methods)) != self.method and next(iter(methods)) in (0, 1, 3, 4):
            self.load(opts_list[0].method)
        c_list = [o.to_c() for o in opts_list]  # keeps seeded densities alive for the call
        c_opts = (_Options * len(opts_list))(*c_list)
        c_res = (_Result * n)()
        if grad_flat is None:
            st = self._lib.pm7_run_batch(self._h, n, c_mols, c_opts, len(opts_list), c_res)
        else:
            st = self._lib.pm7_gradient_batch(self._h, n, c_mols, c_opts, len(opts_list), c_res,
                                              grad_flat.ctypes.data_as(C.POINTER(C.c_double)))
        if st != 0:
            raise Pm7Error(st, self.last_error())
        out = _convert_all(c_res, n, self._lib.pm7_result_free)
        if raise_on_error:
            for r in out:
                if r.status != 0:
                    raise Pm7Error(r.status, r.message, r)
        return out

    def submit_batch(self, molecules: Sequence[Molecule], options=None) -> int:
        """pm7_batch_submit: launch a batch and return its ticket at once; wait_batch(ticket) returns
        what run_batch would. Consecutive batches overlap on the device."""
        n = len(molecules)
        if options is None:
            options = Pm7Options()
        opts_list = options if isinstance(options, (list, tuple)) else [options]
        methods = {METHODS[o.method.lower()] for o in opts_list}
        if len(methods) == 1 and next(iter(methods)) != self.method and next(iter(methods)) in (0, 1, 3, 4):
            self.load(opts_list[0].method)
        c_mols = (_Molecule * max(n, 1))()
        keep = []
        for i, m in enumerate(molecules):
            cm, k = self._mol(m)
            c_mols[i] = cm
            keep.append(k)
        c_list = [o.to_c() for o in opts_list]
        c_opts = (_Options * len(opts_list))(*c_list)
        c_res = (_Result * max(n, 1))()
        ticket = C.c_int64(0)
        st = self._lib.pm7_batch_submit(self._h, n, c_mols, c_opts, len(opts_list), c_res, C.byref(ticket))
        if st != 0:
            raise Pm7Error(st, self.last_error())
        # The result array is written at the wait; the molecule and option arrays only at submit.
        if not hasattr(self, "_pending"):
            self._pending = {}
        self._pending[ticket.value] = (c_res, n)
        return ticket.value

    def wait_batch(self, ticket: int, raise_on_error: bool = False):
        c_res, n = self._pending.pop(ticket)
        st = self._lib.pm7_batch_wait(self._h, int(ticket))
        if st != 0:
            raise Pm7Error(st, self.last_error())
        out = _convert_all(c_res, n, self._lib.pm7_result_free)
        if raise_on_error:
            for r in out:
                if r.status != 0:
                    raise Pm7Error(r.status, r.message, r)
        return out

    def core_hamiltonian(self, molecule: Molecule) -> np.ndarray:
        cm, keep = self._mol(molecule)
        nao = C.c_int32(0)
        st = self._lib.pm7_core_hamiltonian(self._h, C.byref(cm), None, C.byref(nao))
        if st != 0:
            raise Pm7Error(st, self.last_error())
        h = np.zeros((nao.value, nao.value))
        st = self._lib.pm7_core_hamiltonian(self._h, C.byref(cm), h.ctypes.data_as(C.POINTER(C.c_double)),
                                            C.byref(nao))
        if st != 0:
            raise Pm7Error(st, self.last_error())
        return h

    def fock_spin(self, molecule: Molecule, p_tot: np.ndarray, p_spin: np.ndarray) -> np.ndarray:
        cm, keep = self._mol(molecule)
        pt = np.ascontiguousarray(p_tot, dtype=np.float64)
        ps = np.ascontiguousarray(p_spin, dtype=np.float64)
        f = np.zeros_like(pt)
        st = self._lib.pm7_fock_spin(self._h, C.byref(cm), pt.ctypes.data_as(C.POINTER(C.c_double)),
                                     ps.ctypes.data_as(C.POINTER(C.c_double)),
                                     f.ctypes.data_as(C.POINTER(C.c_double)))
        if st != 0:
            raise Pm7Error(st, self.last_error())
        return f

    def response_fock(self, molecule: Molecule, densities: np.ndarray) -> np.ndarray:
        """G(X) = build_fock_x(X) - H_core = J(X) - K(X/2) for each symmetric density X of one
        molecule, the two-electron response kernel of the CPHF solver (pm7-rs cphf_ov,
        src/hessian.rs:1074-1081). ``densities`` is (k, nao, nao) or (nao, nao); the result has the
        same shape. The integrals are built once for all k."""
        cm, keep = self._mol(molecule)
        x = np.ascontiguousarray(densities, dtype=np.float64)
        single = x.ndim == 2
        if single:
            x = x[None]
        nao = C.c_int32(0)
        st = self._lib.pm7_core_hamiltonian(self._h, C.byref(cm), None, C.byref(nao))
        if st != 0:
            raise Pm7Error(st, self.last_error())
        if x.ndim != 3 or x.shape[1:] != (nao.value, nao.value):
            raise ValueError(f"densities must be (k, {nao.value}, {nao.value}), got {densities.shape}")
        g = np.zeros_like(x)
        st = self._lib.pm7_response_fock_batch(self._h, C.byref(cm), x.shape[0],
                                               x.ctypes.data_as(C.POINTER(C.c_double)),
                                               g.ctypes.data_as(C.POINTER(C.c_double)))
        if st != 0:
            raise Pm7Error(st, self.last_error())
        return g[0] if single else g

    def response_fock_spin(self, molecule: Molecule, dpa: np.ndarray, dpb: np.ndarray):
        """(G^a, G^b) with G^s = build_fock_spin_x(dPa + dPb, dP^s) - H_core for each pair of symmetric
        alpha/beta response densities, the open-shell CPHF kernel (pm7-rs analytic_hessian_uhf,
        src/hessian.rs:1850-1880). Arrays are (k, nao, nao) or (nao, nao)."""
        cm, keep = self._mol(molecule)
        a = np.ascontiguousarray(dpa, dtype=np.float64)
        b = np.ascontiguousarray(dpb, dtype=np.float64)
        single = a.ndim == 2
        if single:
            a, b = a[None], b[None]
        nao = C.c_int32(0)
        st = self._lib.pm7_core_hamiltonian(self._h, C.byref(cm), None, C.byref(nao))
        if st != 0:
            raise Pm7Error(st, self.last_error())
        if a.ndim != 3 or a.shape[1:] != (nao.value, nao.value) or b.shape != a.shape:
            raise ValueError(f"densities must be two (k, {nao.value}, {nao.value}) arrays, got {dpa.shape}, {dpb.shape}")
        ga, gb = np.zeros_like(a), np.zeros_like(b)
        ptr = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
        st = self._lib.pm7_response_fock_spin_batch(self._h, C.byref(cm), a.shape[0], ptr(a), ptr(b), ptr(ga), ptr(gb))
        if st != 0:
            raise Pm7Error(st, self.last_error())
        return (ga[0], gb[0]) if single else (ga, gb)

    def core_core_energy(self, molecule: Molecule) -> float:
        cm, keep = self._mol(molecule)
        e = C.c_double(0.0)
        st = self._lib.pm7_core_core_energy(self._h, C.byref(cm), C.byref(e))
        if st != 0:
            raise Pm7Error(st, self.last_error())
        return e.value

    def last_timings_ms(self):
        """[integrals, scf (all classes), gather, wall, scf class 0, 1, 2, 3, host prepare,
        host launch, device wait, host result assembly] in ms."""
        buf = (C.c_double * TIMING_SLOTS)()
        self._lib.pm7_last_timings(self._h, buf, TIMING_SLOTS)
        return list(buf)


_default_engine: Optional[Engine] = None


class Server:
    """Continuous-batching request server on an Engine (pm7_server_*): submit() queues a
    molecule and returns its id; poll() claims a finished result with its latency (submit to
    completion, ms). Each result equals Engine.run of that molecule alone."""

    NOT_READY = 9

    def __init__(self, engine: Engine, lanes: int = 4, max_batch: int = 256, max_memory_mb: int = 0):
        self._engine = engine
        self._lib = engine._lib
        cfg = _ServerConfig(int(lanes), int(max_batch), int(max_memory_mb))
        h = C.c_void_p()
        st = self._lib.pm7_server_start(engine._h, C.byref(cfg), C.byref(h))
        if st != 0:
            raise Pm7Error(st, engine.last_error())
        self._h = h
        self._keep = {}

    def submit(self, molecule: Molecule, options: Optional["Pm7Options"] = None) -> int:
        cm, keep = Engine._mol(molecule)
        opts = (options or Pm7Options(charge=molecule.charge, multiplicity=molecule.multiplicity)).to_c()
        rid = C.c_int64(-1)
        st = self._lib.pm7_server_submit(self._h, C.byref(cm), C.byref(opts), C.byref(rid))
        if st != 0:
            raise Pm7Error(st, "submit refused")
        return rid.value

    def poll(self, request_id: int = -1, timeout_ms: int = -1):
        """(id, Pm7Result, latency_ms) of a finished request, or None when none finished in time."""
        res = (_Result * 1)()
        rid = C.c_int64(-1)
        lat = C.c_double(0.0)
        st = self._lib.pm7_server_poll(self._h, int(request_id), int(timeout_ms), C.byref(rid), res, C.byref(lat))
        if st == self.NOT_READY:
            return None
        if st != 0:
            raise Pm7Error(st, f"poll of request {request_id} failed")
        return rid.value, _convert_all(res, 1, self._lib.pm7_result_free)[0], lat.value

    def stop(self):
        if getattr(self, "_h", None):
            self._lib.pm7_server_stop(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.stop()

    def __del__(self):  # pragma: no cover
        try:
            self.stop()
        except Exception:
            pass


def default_engine() -> Engine:
    global _default_engine
    if _default_engine is None:
        _default_engine = Engine(0)
    return _default_engine


def run_pm7(molecule: Molecule, params: Optional[Pm7Parameters] = None,
            options: Optional[Pm7Options] = None) -> Pm7Result:
    """Drop-in for pm7-rs ``run_pm7`` (src/scf.rs:512-516): raises Pm7Error like ``Err(..)``."""
    options = options or Pm7Options(charge=molecule.charge, multiplicity=molecule.multiplicity)
    if params is not None and METHODS[params.method.lower()] != METHODS[options.method.lower()]:
        raise Pm7Error(1, f"options method `{options.method}` does not match parameter method `{params.method}`")
    eng = default_engine()
    if params is not None and METHODS[params.method.lower()] != eng.method:
        eng.load(params.method)
    return eng.run(molecule, options, raise_on_error=True)


def run_pm7_batch(molecules: Sequence[Molecule], options=None) -> list:
    """Batched single points; per-molecule failures are reported in ``result.status``."""
    return default_engine().run_batch(molecules, options)


__all__ = [
    "Atom", "Molecule", "Pm7Parameters", "Pm7Options", "Pm7Result", "Pm7Error", "Engine", "Server", "run_pm7",
    "run_pm7_batch", "default_engine", "library", "ABI_FUNCTIONS", "LIB_PATH", "DATA_DIR",
]
