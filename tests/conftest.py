"""Shared test setup.

* puts the package directory (``dense-evolution_amd``) and the repo root on
  ``sys.path`` so ``import desv`` and ``import oracle`` work;
* registers the ``gpu`` marker: GPU tests need a visible HIP device and the
  built ``libdesv.so``; CPU tests never launch a kernel.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "dense-evolution_amd")
for p in (PKG, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running GPU parity case")


def _gpu_available() -> bool:
    # Decided from the kernel driver node, not from the library: on a GPU
    # machine a missing or broken libdesv.so must fail the GPU tests loudly.
    return os.path.exists("/dev/kfd")


def pytest_collection_modifyitems(config, items):
    if _gpu_available():
        return
    skip = pytest.mark.skip(reason="no HIP device visible")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
