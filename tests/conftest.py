"""Shared test setup: the `gpu` marker, library builds and common inputs."""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O
    O.lib()  # builds with make if the library is missing
    return O


@pytest.fixture(scope="session")
def pdmk():
    import periodicdmk_amd as P
    if not os.path.exists(P.LIB_PATH):
        from periodicdmk_amd import build
        build.build(verbose=False)
    P.lib()
    return P


@pytest.fixture(scope="session")
def cuda(pdmk):
    """GPU tests go through libpdmk.so's own HIP runtime (no torch in the process)."""
    if pdmk.device_count() < 1:
        pytest.fail("a GPU test ran without a visible GPU")
    return pdmk
