import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libwfs.so")


def pytest_collection_modifyitems(config, items):
    # GPU tests need a visible device. The library itself must load: a
    # missing libwfs.so fails loudly instead of skipping.
    gpu_items = [it for it in items if "gpu" in it.keywords]
    if not gpu_items:
        return
    from water_flow_simulation_amd import _native
    if _native.device_count() == 0:
        skip = pytest.mark.skip(reason="no HIP device visible")
        for it in gpu_items:
            it.add_marker(skip)


def golden(name):
    return np.load(os.path.join(GOLDEN, name))


def unpack_bool(bits, shape):
    n = int(np.prod(shape))
    return np.unpackbits(bits)[:n].reshape(tuple(int(s) for s in shape)).astype(bool)
