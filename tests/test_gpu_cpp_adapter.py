"""Runs the C++ adapter parity test (tests/cpp/test_kinship_batch_gpu.cpp) on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_kinship_batch_gpu")


def test_cpp_kinship_batch_through_adapter():
    assert os.path.exists(BIN), "build() compiles the C++ adapter test"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "BATCH_KOMEGA_PASS" in r.stdout
