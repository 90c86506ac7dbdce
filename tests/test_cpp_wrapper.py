"""Builds and runs tests/cpp/test_wrapper.cpp: the C++ layer over the C-ABI
(include/beamformer/amd.hpp) rethrows the reference's exception types.
CPU only (every call fails validation before any device work)."""
import os
import shutil
import subprocess

import pytest

from beamform_project_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_cpp_wrapper_maps_status_to_exceptions(tmp_path):
    exe = tmp_path / "test_wrapper"
    libdir = os.path.dirname(_capi.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror",
                    "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "test_wrapper.cpp"),
                    "-L", libdir, "-lbeamform_hip", f"-Wl,-rpath,{libdir}", "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK" in r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_reference_named_cpp_api_compiles_and_links(tmp_path):
    """include/beamformer/reference_api.hpp compiles with the reference's
    names and links against the library (the GPU test runs it)."""
    libdir = os.path.dirname(_capi.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror",
                    "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "test_reference_api.cpp"),
                    "-L", libdir, "-lbeamform_hip", f"-Wl,-rpath,{libdir}",
                    "-o", str(tmp_path / "t")], check=True)
