import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def h2o():
    """H2O cc-pVDZ integrals from the reference fixture h2o.ccpvdz.fci.dat."""
    from macis_amd.fcidump import read_fcidump

    norb, E_core, T, V = read_fcidump(os.path.join(GOLDEN, "h2o.ccpvdz.fci.dat"))
    return norb, E_core, T, V


@pytest.fixture(scope="session")
def h2o_oracle(h2o):
    from oracle import oracle as O

    norb, E_core, T, V = h2o
    return O.Ham(T, V)


@pytest.fixture(scope="session")
def h2o_cisd(h2o):
    from macis_amd import dets as D
    from oracle import oracle as O

    norb = h2o[0]
    hf = D.canonical_hf_determinant(1, 5, 5)
    return O.cisd_space(1, norb, hf)


@pytest.fixture(scope="session")
def ctx():
    """One HIP context on cuda:0 for the GPU tests (C-ABI, no fallback)."""
    from macis_amd import capi

    C = capi()
    c = C.Context(0)
    yield c
    c.close()


def announce(msg):
    """Print the case name before GPU work so a crash can be attributed."""
    print(f"\n[case] {msg}", flush=True)
