import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def _ensure_built():
    lib = os.path.join(ROOT, "nelson_amd", "libnelson_amd.so")
    if not os.path.exists(lib):
        from nelson_amd import build as b
        b.build()
    orc = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(orc):
        from oracle import oracle as o
        o.build()


_ensure_built()


@pytest.fixture(scope="session")
def cfg1():
    from nelson_amd import CONFIGS, generate_ba
    return generate_ba(CONFIGS["cfg1"])
