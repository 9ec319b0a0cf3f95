"""GPU: the C++ layer that mirrors the reference API (include/chainamps_amd.hpp)
drives the library from a host C++ program: tests/cpp/spin_half_idmrg.cpp
reproduces the reference's spin_half_example growth energies and entropies
(README.md:31-40) and checks contract / reorder / SVD through the same
layer.  The program is built next to the library by chainamps_amd.build.
"""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_spin_half_example_through_cpp_layer():
    print("case cpp_layer spin_half_idmrg", flush=True)
    exe = os.path.join(ROOT, "chainamps_amd", "lib", "spin_half_idmrg")
    assert os.path.exists(exe), "build the library first (python -m chainamps_amd.build)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASSED" in r.stdout
