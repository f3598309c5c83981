"""Shared pytest configuration.

Tests marked ``gpu`` run the HIP library on a device and compare it with the
CPU oracle (``oracle/``); everything else runs on the CPU.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built library")
    config.addinivalue_line("markers", "slow: longer CPU oracle checks")


def pytest_collection_modifyitems(config, items):
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:  # pragma: no cover
        has_gpu = False
    if has_gpu:
        return
    skip = pytest.mark.skip(reason="no HIP device visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


def rel_l2(a, b) -> float:
    import numpy as np
    a = np.asarray(a)
    b = np.asarray(b)
    nb = float(np.linalg.norm(b.ravel()))
    return float(np.linalg.norm((a - b).ravel())) / (nb if nb > 0 else 1.0)
