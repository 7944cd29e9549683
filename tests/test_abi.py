# SPDX-License-Identifier: GPL-3.0-or-later
"""CPU-side checks of the drop-in boundary: the engine library loads and exports every entry
point include/pm7_amd.h declares, and the Python mirror's structs match the header's fields."""

import ctypes as C
import re
from pathlib import Path

import pytest

import pm7amd

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "pm7_amd.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(pm7_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_abi_list():
    assert declared_functions() == sorted(pm7amd.ABI_FUNCTIONS)


@pytest.mark.parametrize("name", declared_functions())
def test_library_exports_symbol(name):
    lib = pm7amd.library()
    assert hasattr(lib, name), name


def test_abi_version_and_defaults_without_device():
    lib = pm7amd.library()
    lib.pm7_abi_version.restype = C.c_int32
    assert lib.pm7_abi_version() == 2
    o = pm7amd.Pm7Options().to_c()
    assert o.max_scf == 200 and o.e_tol == 1e-8 and o.p_tol == 1e-7 and o.accelerator == 2
    assert o.adiis_switch == 0.1 and o.use_diis == 1 and o.level_shift == 0.0
    assert o.force_dpath == 0 and o.smooth_dispersion == 0 and o.dipole_origin == 0
    assert not o.initial_density and not o.initial_spin_density


def test_struct_field_names_follow_header():
    text = HEADER.read_text()
    for struct, fields in (("pm7_options_v2", pm7amd._Options._fields_),
                           ("pm7_result_v2", pm7amd._Result._fields_)):
        body = text[:text.index(f"}} {struct};")]
        body = body[body.rindex("typedef struct {"):]
        for name, _ in fields:
            assert re.search(rf"\b{name}\b", body), (struct, name)


def test_context_creation_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(pm7amd.Pm7Error):
        pm7amd.Engine(0)


def test_struct_layout_matches_the_compiled_header(tmp_path):
    """Offsets and sizes of the ctypes mirrors equal what a C compiler makes of the header."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        pytest.skip("no C compiler")
    lines = ["#include <stddef.h>", "#include <stdio.h>", f'#include "{HEADER}"', "int main(void) {"]
    for struct, cls in (("pm7_options_v2", pm7amd._Options), ("pm7_result_v2", pm7amd._Result),
                        ("pm7_molecule_v1", pm7amd._Molecule), ("pm7_server_config_v1", pm7amd._ServerConfig),
                        ("pm7_tune_result_v1", pm7amd._TuneResult)):
        lines.append(f'  printf("{struct} size %zu\\n", sizeof({struct}));')
        for name, _ in cls._fields_:
            lines.append(f'  printf("{struct} {name} %zu\\n", offsetof({struct}, {name}));')
    lines.append("  return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-o", str(exe), str(src)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        struct, name, value = line.split()
        got[(struct, name)] = int(value)
    for struct, cls in (("pm7_options_v2", pm7amd._Options), ("pm7_result_v2", pm7amd._Result),
                        ("pm7_molecule_v1", pm7amd._Molecule), ("pm7_server_config_v1", pm7amd._ServerConfig),
                        ("pm7_tune_result_v1", pm7amd._TuneResult)):
        assert got[(struct, "size")] == C.sizeof(cls), struct
        for name, _ in cls._fields_:
            assert got[(struct, name)] == getattr(cls, name).offset, (struct, name)
