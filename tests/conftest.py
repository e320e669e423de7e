import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the libycn kernels")
    config.addinivalue_line("markers", "slow: long-running statistical test")


def _have_gpu():
    try:
        import yukawacolourednoise_amd as Y
    except ImportError:
        return False
    p = Y.default_params(Nt=4, Nx=4)
    try:
        Y.Context(p).close()
        return True
    except Y.YcnError:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not _have_gpu():
        pytest.fail("no HIP device / libycn.so not loadable: GPU tests must run on the GPU")
    import yukawacolourednoise_amd as Y
    return Y
