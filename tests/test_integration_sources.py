"""The maintainer-side C++ boundary (oidn-protoray_amd/integration: DeviceHip over the C ABI, and
HipDenoiser) compiles against the reference's own Device, Denoiser, Props and tone-mapper headers,
and implements every pure virtual of Device and Denoiser (tests/integration/instantiate.cpp
instantiates both).  Needs the reference tree (read in place, never copied); skipped without it."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
FLAGS = ["-std=gnu++17", "-mavx2", "-mfma", "-mf16c", "-mlzcnt", "-mbmi", "-mbmi2", "-fsyntax-only", "-Wall",
         "-Wno-class-memaccess", "-I" + REF, "-I" + os.path.join(ROOT, "include"),
         "-I" + os.path.join(ROOT, "oidn-protoray_amd", "integration")]


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "render")), reason="reference headers not present")
@pytest.mark.parametrize("src", ["oidn-protoray_amd/integration/device_hip.cpp", "tests/integration/instantiate.cpp",
                                 "tests/integration/device_hip_driver.cpp"])
def test_integration_compiles_against_reference_headers(src):
    r = subprocess.run(["g++"] + FLAGS + [os.path.join(ROOT, src)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "warning" not in r.stderr, r.stderr
