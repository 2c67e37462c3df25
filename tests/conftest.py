import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libmmr_recon.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import binding

    return binding.lib()


@pytest.fixture(scope="session")
def ctx():
    import mmrecode_amd

    c = mmrecode_amd.Context(device=0)
    yield c
    c.close()
