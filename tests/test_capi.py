"""The C ABI library loads, exports every declared symbol, and validates
arguments on the host (no kernel launches here)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import desv
from desv import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    with open(os.path.join(ROOT, "include", "desv.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(desv_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    syms = header_symbols()
    assert len(syms) >= 30
    lib = C.CDLL(_capi.LIB_PATH)
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(_capi.SYMBOLS) == syms


def test_version_and_error_string():
    assert _capi.lib.desv_version() >= 1
    assert isinstance(_capi.last_error(), str)


def test_program_validation_errors():
    p = desv.Program(3)
    with pytest.raises(ValueError, match="unknown gate"):
        p.add_ops(np.array([[17, 0, 0, 0.0]]))
    with pytest.raises(ValueError, match="out of range"):
        p.add_ops(np.array([[1, 3, 0, 0.0]]))
    with pytest.raises(ValueError, match="must differ"):
        p.add_ops(np.array([[20, 1, 1, 0.0]]))
    with pytest.raises(ValueError):
        p.add_matrix(np.eye(64), [0, 1, 2, 0, 1, 2])
    with pytest.raises(ValueError, match="not compiled"):
        p.stats()
    p.close()


def test_program_create_rejects_bad_sizes():
    with pytest.raises(ValueError, match=r"\[1, 34\]"):
        desv.Program(35)
    with pytest.raises(ValueError):
        desv.Program(0)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the no-device path")
def test_create_without_device_is_a_runtime_error():
    with pytest.raises(RuntimeError, match="no HIP device"):
        desv.DenseSVSimulator(3)


def test_simulator_validates_before_the_device():
    with pytest.raises(ValueError, match=r"\[1, 34\]"):
        desv.DenseSVSimulator(0)
    with pytest.raises(ValueError, match=r"\[1, 34\]"):
        desv.DenseSVSimulator(35)
