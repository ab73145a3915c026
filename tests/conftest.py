import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def pytest_collection_modifyitems(config, items):
    # GPU tests skip (rather than fail) where no HIP device is visible.
    if not _has_gpu():
        skip = pytest.mark.skip(reason="no HIP device visible")
        for it in items:
            if "gpu" in it.keywords:
                it.add_marker(skip)


_GPU = None


def _has_gpu():
    global _GPU
    if _GPU is None:
        _GPU = os.path.exists("/dev/kfd")
        if _GPU:
            try:
                import torch
                _GPU = torch.cuda.device_count() > 0
            except Exception:
                _GPU = False
    return _GPU


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
