"""Device scratch layouts checked on the host: a C++ program built from
tests/layout_check.cpp against the library's own kernel header runs every
case (no GPU)."""
import shutil
import subprocess

import pytest

from conftest import ROOT


def test_eppstein_scratch_layout_aligned(tmp_path):
    """Every region of Eppstein's per-automaton scratch starts on an 8-byte
    boundary (its 64-bit keys take 64-bit atomics; a 4-byte offset faulted
    the GPU once) and the total is a multiple of 256, for all n <= 512 and
    k <= 8."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no host C++ compiler")
    exe = tmp_path / "layout_check"
    subprocess.run([gxx, "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    str(ROOT / "tests" / "layout_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert "checked 4096 layouts, 0 bad" in r.stdout
