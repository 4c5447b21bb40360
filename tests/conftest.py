import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    # torch's own HIP runtime must initialise the device before the one
    # libagdeblend links, or torch sees no device in this process
    try:
        import torch
        torch.cuda.is_available()
    except ImportError:
        pass
