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
    # the rotating asci_grow case compares with the oracle's energy
    import numpy as np

    from macis_amd import dets as D
    from macis_amd.fcidump import read_fcidump
    from oracle import oracle as O

    _, _, T, V = read_fcidump(os.path.join(GOLDEN, "h2o.ccpvdz.fci.dat"))
    H = O.Ham(T, V)
    s = O.ASCISettings.defaults()
    s.ntdets_max = 2000
    hf = D.canonical_hf_determinant(1, 5, 5)
    E, w, _ = H.asci_grow_rot(1, s, H.matrix_element(1, hf, hf), hf, np.array([1.0]),
                              rot_size_start=100)
    assert len(w) == 2000
    _run(f"--rot-ref={E!r}")
