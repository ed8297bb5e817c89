import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "orac-porting_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


def pytest_runtest_logstart(nodeid, location):
    """With ORAC_TEST_TRACE=1 each test's id goes to stderr (flushed) before it
    runs, so a run that ends abruptly still names the test it was in."""
    if os.environ.get("ORAC_TEST_TRACE") == "1":
        sys.stderr.write(f"start {nodeid}\n")
        sys.stderr.flush()


def relrms(a, b):
    """Relative RMS error of a against reference b (the reference's check_force, pme_fftw3.f90:390-405)."""
    a, b = np.asarray(a), np.asarray(b)
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.fixture(scope="session")
def water216():
    from orac_hip import synth
    return synth.config("water216")


@pytest.fixture(scope="session")
def small_solute():
    from orac_hip import synth
    return synth.small_solvated(rcut=7.0)


@pytest.fixture(scope="session")
def gpu_ctx():
    from orac_hip.api import Context
    c = Context(0)
    yield c
    c.close()
