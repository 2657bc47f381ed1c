"""The command-line drivers (snc_amd/bin, built by snc_amd/csrc/Makefile):
plain C++ programs over the C ABI. On the CPU: they exist, print usage, and
refuse to run without a gfx950 device."""
import os
import subprocess

import pytest

from snc_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "snc_amd", "bin")
TOOLS = ["snc_hip_bench", "snc_hip_train", "snc_hip_cotrain"]


@pytest.mark.parametrize("tool", TOOLS)
def test_driver_prints_usage(tool):
    exe = os.path.join(BIN, tool)
    assert os.access(exe, os.X_OK), f"{exe} not built (make -C snc_amd/csrc)"
    r = subprocess.run([exe, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and tool in r.stdout
    r = subprocess.run([exe, "--no-such-option"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2


@pytest.mark.skipif(L.lib().snc_hip_available() == 1, reason="a device is present")
@pytest.mark.parametrize("tool", TOOLS)
def test_driver_refuses_without_device(tool):
    r = subprocess.run([os.path.join(BIN, tool)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 3 and "no usable gfx950 device" in r.stderr
