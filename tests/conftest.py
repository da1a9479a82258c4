import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "fermi.jl_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libfermihip.so")
    config.addinivalue_line("markers", "slow: oracle runs of more than a few seconds")


def gpu_available():
    try:
        import torch  # device query only
        return torch.cuda.is_available()
    except Exception:
        return False
