"""Runs the C++ API tests (macis_amd/cpp_tests/test_api.cpp), which use the
C++ mirror of the reference API the way the reference's own tests do."""
import os
import subprocess

import pytest

from conftest import GOLDEN, ROOT

EXE = os.path.join(ROOT, "macis_amd", "cpp_tests", "test_api")


def _run(*extra):
    assert os.path.exists(EXE), "C++ tests not built (run __graft_entry__.build())"
    p = subprocess.run([EXE, GOLDEN, *extra], capture_output=True, text=True, timeout=600)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr


def test_cpp_api_host_cases():
    _run("--cpu-only")


@pytest.mark.gpu
def test_cpp_api_device_cases():
    _run()
