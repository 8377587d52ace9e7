"""The C ABI used from a plain C++ host program (tests/c_abi/abi_driver.cpp,
built by build() with g++ against the library and the HIP runtime, no Python
or torch in the process): life cycle, mass, determinism, state image,
injection and error codes, checked inside the program."""
import os
import subprocess

import pytest
import torch

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c_abi", "abi_driver")


def test_c_abi_driver():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    assert os.path.exists(DRIVER), "tests/c_abi/abi_driver is built by __graft_entry__.build()"
    print("[gpu] C ABI driver", flush=True)
    out = subprocess.run([DRIVER], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0, out.stderr
    assert "abi_driver ok" in out.stdout
