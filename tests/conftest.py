# SPDX-License-Identifier: GPL-3.0-or-later
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "tests"))
sys.path.insert(0, str(ROOT / "pm7-rs_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    _ensure_oracle()


def _ensure_oracle():
    """The CPU oracle is test infrastructure; build it in place when missing or stale."""
    lib = ROOT / "oracle" / "liboracle_pm7.so"
    srcs = [ROOT / "oracle" / "pm7_oracle.cpp", ROOT / "oracle" / "pm7_oracle.h"]
    if lib.exists() and all(lib.stat().st_mtime >= s.stat().st_mtime for s in srcs):
        return
    try:
        subprocess.run(["make", "-C", str(ROOT / "oracle")], check=True, capture_output=True)
    except (OSError, subprocess.CalledProcessError) as exc:  # pragma: no cover - reported by tests
        sys.stderr.write(f"oracle build failed: {exc}\n")
