import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


C_CLIENT = os.path.join(ROOT, "tests", "c", "run_system_client")


def _ensure_built():
    from milkywayathome_client_amd import build as b
    if not (os.path.exists(b.LIB_PATH) and os.path.exists(b.MODELS_LIB_PATH)):
        b.build()
    from oracle import oracle
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    if not os.path.exists(C_CLIENT):
        import subprocess
        subprocess.run(["make", "-s", "-C", os.path.dirname(C_CLIENT)], check=True)


_ensure_built()


@pytest.fixture(scope="session")
def gpu_available():
    import ctypes
    from milkywayathome_client_amd.nbody import load_library
    lib = load_library()
    assert lib is not None
    return True
