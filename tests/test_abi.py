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
    assert lib.pm7_abi_version() == 1
    o = pm7amd.Pm7Options().to_c()
    assert o.max_scf == 200 and o.e_tol == 1e-8 and o.p_tol == 1e-7 and o.accelerator == 2
    assert o.adiis_switch == 0.1 and o.use_diis == 1 and o.level_shift == 0.0


def test_struct_field_names_follow_header():
    text = HEADER.read_text()
    for struct, fields in (("pm7_options_v1", pm7amd._Options._fields_),
                           ("pm7_result_v1", pm7amd._Result._fields_)):
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
