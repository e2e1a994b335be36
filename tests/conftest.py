import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def pytest_collection_modifyitems(config, items):
    try:
        import torch
        have_gpu = torch.cuda.is_available()
    except Exception:
        have_gpu = False
    if have_gpu:
        return
    skip = pytest.mark.skip(reason="no HIP device visible")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def al_sim_pattern():
    """The reference's own test pattern (pyebsdindex/tests/data/al_sim_20kv/al_sim_20kv.png):
    100 x 120 uint8, EMsoft-simulated Al at 20 kV."""
    from PIL import Image
    return np.array(Image.open(os.path.join(GOLDEN, "al_sim_20kv.png")))
