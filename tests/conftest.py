import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(REPO, "evoc-rs_amd"), os.path.join(REPO, "oracle"), os.path.join(REPO, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: longer CPU-side checks")


def pytest_sessionstart(session):
    """torch bundles its own HIP runtime; it only finds the GPU if it initialises it before
    libevoc_hip.so (linked against /opt/rocm's runtime) does. Tests that generate data
    with torch rely on this; on a machine without a GPU nothing happens."""
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.zeros(1, device="cuda")
        torch.cuda.synchronize()
