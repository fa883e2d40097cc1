import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")


@pytest.fixture(scope="session")
def hip():
    """The loaded HIP library (GPU tests only); fails loudly if absent."""
    import logosq_amd
    lib = logosq_amd.load()
    name = __import__("ctypes").create_string_buffer(256)
    cu = __import__("ctypes").c_int()
    rc = lib.lq_device_info(name, 256, __import__("ctypes").byref(cu), None, None)
    assert rc == 0, lib.lq_last_error()
    print("device:", name.value.decode(), "CUs:", cu.value, flush=True)
    return lib


@pytest.fixture(scope="session", autouse=True)
def forced_paths():
    """LQ_TEST_FORCE_PATHS=1: run every test with each fused pass on its
    run-time specialised kernel (compiled at first sight, any size) and every
    expectation's flip groups from LDS tiles (from 8 qubits), so the whole
    suite checks those paths against the oracle too."""
    if os.environ.get("LQ_TEST_FORCE_PATHS") == "1":
        from tests import modes
        modes.restore_defaults()
    yield
