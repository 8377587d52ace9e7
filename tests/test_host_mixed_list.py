"""Host check of the single-launch pass-B chunk list (csrc/lbm_lists.hpp):
tests/c_abi/mixed_list_check.cpp is built with g++ and run (no device)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_mixed_chunk_list_properties(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "mixed_list_check"
    subprocess.run([gxx, "-O1", "-std=c++17", os.path.join(HERE, "c_abi", "mixed_list_check.cpp"),
                    "-o", str(exe)], check=True, capture_output=True, timeout=120)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:]
    assert "mixed_list ok" in r.stdout
