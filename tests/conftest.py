import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running parity case")


def _built():
    lib = os.path.join(ROOT, "natureuserinterfacestudio_amd", "lib")
    return all(os.path.exists(os.path.join(lib, n)) for n in ("libnuif.so", "libnuif_kinfu.so", "libnuif_synth.so")) \
        and os.path.exists(os.path.join(ROOT, "oracle", "_build", "liboracle.so"))


@pytest.fixture(scope="session", autouse=True)
def native_built():
    if not _built():
        from natureuserinterfacestudio_amd import _build
        _build.build_all()
    yield


@pytest.fixture(scope="session")
def gpu():
    """Initialises the device; the GPU tests must not silently fall back."""
    from natureuserinterfacestudio_amd import nuif
    nuif.init(0)
    return nuif
