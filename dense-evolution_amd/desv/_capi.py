"""ctypes binding of ``libdesv.so`` (the C ABI declared in ``include/desv.h``).

The loader follows the reference's native-library convention
(``dense_evolution/native_hf/libcint_bridge.py:59-83``): an environment
override (``DESV_LIB``), then the in-tree build, with argtypes/restype set
once.  There is no fallback: if the library is missing the import fails and
names the build command.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("DESV_LIB") or os.path.join(_HERE, "lib", "libdesv.so")

DESV_OK, DESV_EINVAL, DESV_ENOMEM, DESV_EDEVICE, DESV_EUNSUPPORTED = 0, 1, 2, 3, 4
C64, C128 = 0, 1

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"libdesv.so not found at {LIB_PATH}; build it with "
        f"`make -C dense-evolution_amd/csrc` (or __graft_entry__.build())")

# One HIP runtime per process: PyTorch-ROCm ships its own libamdhip64.so.7 and
# HSA runtime under the same sonames as /opt/rocm's.  Whichever loads first
# serves every later user of that soname, and a process that ends up with the
# system HIP runtime under torch's HSA runtime sees no GPU from torch.  Loading
# torch first (when it is installed) makes the engine and torch share torch's
# runtime, so device tensors, streams and events pass between them.
if os.environ.get("DESV_NO_TORCH", "") != "1":
    try:
        import torch  # noqa: F401
    except ImportError:
        pass

lib = C.CDLL(LIB_PATH)

_p = C.c_void_p
_i = C.c_int
_i64 = C.c_int64
_d = C.c_double
_pd = C.POINTER(C.c_double)
_pi64 = C.POINTER(C.c_int64)
_pu64 = C.POINTER(C.c_uint64)
_pi = C.POINTER(C.c_int)


class CompileOpts(C.Structure):
    """``desv_compile_opts``."""
    _fields_ = [("fuse_bits", _i), ("tile_bits", _i), ("low_bits", _i),
                ("per_gate", _i), ("use_mfma", _i), ("dtype", _i), ("compensated", _i),
                ("reg_groups", _i), ("plan_search", _i)]


_SIGS = {
    "desv_last_error": ([], C.c_char_p),
    "desv_version": ([], _i),
    "desv_device_count": ([_pi], _i),
    "desv_create": ([_i, _i, _i, C.POINTER(_p)], _i),
    "desv_destroy": ([_p], _i),
    "desv_set_stream": ([_p, _p], _i),
    "desv_synchronize": ([_p], _i),
    "desv_set_double_buffer": ([_p, _i], _i),
    "desv_set_compensated": ([_p, _i], _i),
    "desv_batch_set_compensated": ([_p, _i], _i),
    "desv_batch_completion_us": ([_p, _pd], _i),
    "desv_batch_get_states": ([_p, _p], _i),
    "desv_batch_set_forward_only": ([_p, _i], _i),
    "desv_batch_run_reverse": ([_p, _pd], _i),
    "desv_batch_state_device_ptr": ([_p, _i64, C.POINTER(_p)], _i),
    "desv_state_device_ptr": ([_p, C.POINTER(_p)], _i),
    "desv_norm2_device": ([_p, _p], _i),
    "desv_reset": ([_p], _i),
    "desv_set_state": ([_p, _p], _i),
    "desv_get_state": ([_p, _p], _i),
    "desv_copy": ([_p, _p], _i),
    "desv_scale": ([_p, _d, _d], _i),
    "desv_norm2": ([_p, _pd], _i),
    "desv_inner": ([_p, _p, _pd, _pd], _i),
    "desv_inner_many": ([_p, C.POINTER(_p), _i, _pd], _i),
    "desv_combine": ([C.POINTER(_p), _i, C.POINTER(_p), _i, _pd], _i),
    "desv_fill_random": ([_p, C.c_uint64], _i),
    "desv_probabilities": ([_p, _pd], _i),
    "desv_qubit_probs": ([_p, _i, _pd, _pd], _i),
    "desv_collapse": ([_p, _i, _i, _d], _i),
    "desv_sample": ([_p, _pd, _i64, _pi64], _i),
    "desv_run_ops": ([_p, _pd, _i64], _i),
    "desv_apply_matrix": ([_p, _pd, _pi, _i], _i),
    "desv_default_opts": ([_i, C.POINTER(CompileOpts)], _i),
    "desv_program_create": ([_i, C.POINTER(_p)], _i),
    "desv_program_destroy": ([_p], _i),
    "desv_program_add_ops": ([_p, _pd, _i64], _i),
    "desv_program_add_matrix": ([_p, _pd, _pi, _i], _i),
    "desv_program_compile": ([_p, C.POINTER(CompileOpts)], _i),
    "desv_program_stats": ([_p, _pi64, _pi64, _pi64, _pi64], _i),
    "desv_run_program": ([_p, _p], _i),
    "desv_run_program_timed": ([_p, _p, C.POINTER(C.c_float)], _i),
    "desv_program_plan_sizes": ([_p, _pi64, _pi64, _pi64], _i),
    "desv_program_plan_dump": ([_p, _pu64, _pi64, _pi64, _pd], _i),
    "desv_program_reg_sizes": ([_p, _i, _pi64, _pi64], _i),
    "desv_program_reg_dump": ([_p, _i, _pi, _pi, _pi, C.POINTER(C.c_float)], _i),
    "desv_program_tail_dump": ([_p, _i, _pi64, _pi64, _pu64, _pi, _pu64, _pi64, _pi, _pi], _i),
    "desv_pauli_expect": ([_p, _pu64, _pu64, _i64, _pd], _i),
    "desv_pauli_matvec": ([_p, _p, _pu64, _pu64, _pd, _pd, _i64], _i),
    "desv_pauli_sum_expect": ([_p, _pu64, _pu64, _pd, _i64, _pd], _i),
    "desv_apply1q_probs": ([_p, _i, _pd, _i, _pd, _pd], _i),
    "desv_apply_monomial_layer": ([_p, _i, _pi, _pd, _i, _pi, _pd], _i),
    "desv_dense_matvec": ([_p, _p, _pd], _i),
    "desv_adjoint": ([_p, _pd, _i64, _pi64, _i64, _pu64, _pu64, _pd, _i64, _pd, _pd, _pd], _i),
    "desv_batch_create": ([_i, _i, C.POINTER(_p)], _i),
    "desv_batch_destroy": ([_p], _i),
    "desv_batch_add": ([_p, _i, _pd, _i64, _pi64, _i64, _pu64, _pu64, _pd, _i64, _pi64], _i),
    "desv_batch_set_params": ([_p, _i64, _pd], _i),
    "desv_batch_set_initial_state": ([_p, _i64, _p], _i),
    "desv_batch_get_state": ([_p, _i64, _p], _i),
    "desv_batch_configure": ([_p, _i, _i, _i, _i], _i),
    "desv_batch_run": ([_p, _pd, _pd], _i),
    "desv_batch_last_time_us": ([_p, _pd], _i),
}

#: every entry point the header declares (checked by the symbol test)
SYMBOLS = tuple(_SIGS)

for _name, (_args, _res) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.argtypes = _args
    _f.restype = _res


def last_error() -> str:
    msg = lib.desv_last_error()
    return msg.decode() if msg else ""


def check(rc: int) -> None:
    """Map a status code to the reference's exception types."""
    if rc == DESV_OK:
        return
    msg = last_error()
    if rc == DESV_EINVAL:
        raise ValueError(msg)
    if rc == DESV_ENOMEM:
        raise MemoryError(msg)
    if rc == DESV_EUNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)


def device_count() -> int:
    c = C.c_int(0)
    check(lib.desv_device_count(C.byref(c)))
    return c.value


# ---- small helpers for array arguments -------------------------------------

def dptr(a: np.ndarray):
    return a.ctypes.data_as(_pd)


def i64ptr(a: np.ndarray):
    return a.ctypes.data_as(_pi64)


def u64ptr(a: np.ndarray):
    return a.ctypes.data_as(_pu64)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_pi)


def vptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)
