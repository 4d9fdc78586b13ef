"""Shared pytest configuration.

Tests marked ``gpu`` need an MI355X (gfx950) visible to torch; everything
else runs on the CPU.  GPU tests skip only when no GPU is visible at all.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs a visible AMD GPU (runs the HIP kernels)")
    config.addinivalue_line(
        "markers", "slow: long-running CPU test")


def _gpu_visible():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _gpu_visible():
        return
    skip = pytest.mark.skip(reason="no GPU visible to torch")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)
