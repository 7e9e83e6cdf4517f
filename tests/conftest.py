import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("", "nbody-fmm_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI of libnbfmm.so)")


@pytest.fixture(scope="session")
def ref_lib():
    import orc
    if not orc.ref_available():
        pytest.skip("reference build oracle/_ref/libnbref.so not present")
    return orc.ref()
