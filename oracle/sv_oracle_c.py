"""ctypes binding of the C restatement of the oracle (``oracle/c``).

TEST INFRASTRUCTURE ONLY: used by tests/ as a checker for sizes where the
NumPy oracle is slow, and by bench.py's ``cpu_baseline`` leg (all host cores,
OpenMP).  Same conventions as :mod:`oracle.sv_oracle`, which pins it
(tests/test_oracle_c.py).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "c", "libsvoracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing; build it with `make -C oracle/c`")
        _lib = C.CDLL(LIB_PATH)
        _lib.svo_apply_matrix.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double),
                                          C.POINTER(C.c_int), C.c_int, C.c_int]
        _lib.svo_apply_matrix.restype = C.c_int
        _lib.svo_run_ops.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double), C.c_int64, C.c_int]
        _lib.svo_run_ops.restype = C.c_int
        _lib.svo_max_threads.restype = C.c_int
    return _lib


def default_threads() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def _state(sv):
    if sv.dtype not in (np.complex64, np.complex128) or not sv.flags.c_contiguous:
        raise ValueError("state must be a contiguous complex64/complex128 array")
    n = sv.shape[0].bit_length() - 1
    if sv.shape != (1 << n,):
        raise ValueError("state length must be a power of 2")
    return n, int(sv.dtype == np.complex64)


def apply_matrix(sv: np.ndarray, mat, qubits, threads: int = 0) -> np.ndarray:
    """In place: sv <- U sv on `qubits` (qubits[0] = most significant matrix bit)."""
    n, single = _state(sv)
    k = len(qubits)
    m = np.ascontiguousarray(np.asarray(mat, dtype=np.complex128).reshape(1 << k, 1 << k))
    q = np.ascontiguousarray(qubits, dtype=np.int32)
    rc = lib().svo_apply_matrix(sv.ctypes.data, single, n, m.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)),
                                q.ctypes.data_as(C.POINTER(C.c_int)), k, threads or default_threads())
    if rc:
        raise ValueError("bad matrix application")
    return sv


def run_ops(sv: np.ndarray, ops, threads: int = 0) -> np.ndarray:
    """In place: apply an op table [g_id, q1, q2, param] row by row."""
    n, single = _state(sv)
    o = np.ascontiguousarray(ops, dtype=np.float64).reshape(-1, 4)
    rc = lib().svo_run_ops(sv.ctypes.data, single, n, o.ctypes.data_as(C.POINTER(C.c_double)), len(o),
                           threads or default_threads())
    if rc:
        raise ValueError("unknown gate id in op table")
    return sv
