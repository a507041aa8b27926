"""The drop-in's reference side (caps_amd/adapter/caps_gpu.cpp): CAPS's own ForceField, EnergyOptions, Cell, System and
DynamicsOptions turned into the capsgpu C ABI, compiled here against the reference's headers (core/include) as CAPS's
core would compile it, and every capsgpu entry point it calls found among the library's exports. Host code only; skipped
where the reference's headers are not on this machine."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF_INCLUDE = "/root/reference/core/include"
SRC = os.path.join(ROOT, "caps_amd", "adapter", "caps_gpu.cpp")
LIB = os.path.join(ROOT, "caps_amd", "libcapsgpu.so")

pytestmark = pytest.mark.skipif(not os.path.isdir(REF_INCLUDE) or shutil.which("g++") is None,
                                reason="the reference's headers or g++ are not on this machine")


def _symbols(path, undefined):
    """the undefined symbols of an object file, or the defined dynamic symbols of a shared library"""
    cmd = ["nm", "-u", path] if undefined else ["nm", "-D", "--defined-only", path]
    out = subprocess.run(cmd, capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_the_adapter_compiles_against_the_reference_headers(tmp_path):
    obj = str(tmp_path / "caps_gpu.o")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-c", "-I", REF_INCLUDE, "-I", os.path.join(ROOT, "include"),
                        SRC, "-o", obj], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    called = {s for s in _symbols(obj, True) if s.startswith("capsgpu_")}
    # the entry points the stand-ins for Evaluator and run_dynamics need
    assert {"capsgpu_create", "capsgpu_destroy", "capsgpu_compute", "capsgpu_set_state", "capsgpu_get_state",
            "capsgpu_run_dynamics", "capsgpu_last_error", "capsgpu_list_builds"} <= called
    if os.path.exists(LIB):
        missing = called - _symbols(LIB, False)
        assert not missing, missing
    # what it takes from the reference's core is the reference's own (resolved when linked into it)
    ref = {s for s in _symbols(obj, True) if s.startswith("_ZN4caps")}
    assert any("mixed_pair" in s for s in ref)
