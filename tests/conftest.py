import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libebivgpu.so")


def pytest_collection_modifyitems(config, items):
    # GPU tests run in file order, the allocation-free entry point's test last.
    pass


@pytest.fixture(scope="session")
def gpu_codec_factory():
    from govid_amd.codec import Codec

    def make(ring=3):
        return Codec(ring=ring, device=0)

    return make
