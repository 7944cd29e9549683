# SPDX-License-Identifier: GPL-3.0-or-later
"""Host-only sanitizer builds (`make -C pm7-rs_amd sanitize`, g++, no device code): the engine's
host sources (parameter loading, geometric corrections, divide-and-conquer partition) under
AddressSanitizer + UndefinedBehaviorSanitizer, and the lockstep rounds of the grouped large-system
path under ThreadSanitizer with the device work stubbed."""

import os
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SAN = ROOT / "pm7-rs_amd" / "_build" / "san"


@pytest.fixture(scope="module")
def built():
    r = subprocess.run(["make", "-C", str(ROOT / "pm7-rs_amd"), "sanitize"], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(f"sanitizer build failed:\n{r.stdout}\n{r.stderr}")
    return SAN


def run(exe, *args, env_extra=None):
    env = dict(os.environ)
    env.pop("LD_PRELOAD", None)
    env.update(env_extra or {})
    return subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=600, env=env, cwd=ROOT)


def test_host_code_under_asan_and_ubsan(built):
    r = run(built / "host_check", str(ROOT / "pm7-rs_amd" / "data"),
            env_extra={"ASAN_OPTIONS": "detect_leaks=1:halt_on_error=1", "UBSAN_OPTIONS": "halt_on_error=1"})
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
    assert "host_check: ok" in r.stdout


def test_lockstep_rounds_under_tsan(built):
    r = run(built / "lockstep_check", env_extra={"TSAN_OPTIONS": "halt_on_error=1"})
    assert r.returncode == 0, r.stdout + r.stderr
    assert "WARNING: ThreadSanitizer" not in r.stderr
    assert "lockstep_check: ok" in r.stdout
