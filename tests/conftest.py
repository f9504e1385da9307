import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def bundle():
    from rrea_feedback_warpx_amrex_geant4_amd.transport import load_bundle
    return load_bundle()


@pytest.fixture(scope="session")
def otables(bundle):
    from oracle import binding as ob
    ob.build()
    return ob.Tables(bundle)


@pytest.fixture(scope="session")
def ref_available():
    return os.path.isdir("/root/reference/native_amrex_rrea")
