import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def built():
    """Native pieces built in-tree (oracle, synth host lib); HIP lib must exist."""
    import oracle
    from supergwas_amd import build as b
    if not os.path.exists(os.path.join(ROOT, "oracle", "build", "libkgoracle.so")):
        oracle.build()
    if not os.path.exists(os.path.join(b.LIB, "libkgsynth.so")):
        b.build_synth()
    return True
