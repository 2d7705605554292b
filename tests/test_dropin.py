"""The drop-in shim's copies between GIZMO's particle arrays and the library's flat arrays
(include/gizmo_dropin.h, used by integration/gz_dropin.cc): tests/dropin_check.cpp compiled with
g++ against the header on particle types with GIZMO's member names and shapes, checking pack,
per-stage unpack, the wakeup report and the active flags entry by entry (CPU only)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
def test_dropin_pack_unpack(tmp_path):
    exe = tmp_path / "dropin_check"
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "dropin_check.cpp"), "-o", str(exe)]
    subprocess.run(cmd, check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "dropin ok" in r.stdout
