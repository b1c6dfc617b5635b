"""Readout conveniences: shot sampling and pure-state fidelity.

Same behaviour as the reference's ``utils/measurement.py:17-94``: shots are
drawn with the caller's NumPy generator (the same uniforms its
``rng.choice`` draws) and located in the state's cumulative probabilities on
the GPU (fixed-point integer prefix sums, so the CDF is nondecreasing and an
exactly-zero amplitude is never drawn; one binary search per shot); the
overlap of two states is accumulated in FP64 on the GPU.
"""
import ctypes as C

import numpy as np

from . import _capi
from ._capi import check, lib
from .simulator import DeviceState, StateHandle


def _device(statevector):
    if isinstance(statevector, DeviceState) and isinstance(statevector.handle, StateHandle):
        return statevector.handle, False
    a = np.asarray(statevector)
    n = a.shape[0].bit_length() - 1
    if 1 << n != a.shape[0]:
        raise ValueError(f"statevector length {a.shape[0]} is not a power of 2")
    h = StateHandle(n, a.dtype != np.complex64)
    h.upload(a)
    return h, True


def sample_counts(statevector, n_shots, rng=None):
    """Qiskit-style counts {bitstring (qubit 0 first): hits} of n_shots samples."""
    dim = statevector.shape[0] if hasattr(statevector, 'shape') else len(statevector)
    n = int(dim).bit_length() - 1
    if 1 << n != dim:
        raise ValueError(f"statevector length {dim} is not a power of 2")
    if n_shots < 1:
        raise ValueError(f"n_shots must be >= 1, got {n_shots}")
    if rng is None:
        rng = np.random.default_rng()
    # the same draws as the reference's rng.choice(dim, n_shots, p=probs):
    # n_shots uniforms, searched in the normalised CDF (built on the GPU)
    u = np.ascontiguousarray(rng.random(n_shots), dtype=np.float64)
    outcomes = np.empty(n_shots, dtype=np.int64)
    h, owned = _device(statevector)
    try:
        check(lib.desv_sample(h.ptr, _capi.dptr(u), n_shots, _capi.i64ptr(outcomes)))
    finally:
        if owned:
            h.close()
    values, freqs = np.unique(outcomes, return_counts=True)
    return {format(int(v), f'0{n}b'): int(f) for v, f in zip(values, freqs)}


def statevector_fidelity(statevector_a, statevector_b):
    """|<a|b>|^2 between two pure states (overlap accumulated in FP64)."""
    sa = statevector_a.shape if hasattr(statevector_a, 'shape') else np.asarray(statevector_a).shape
    sb = statevector_b.shape if hasattr(statevector_b, 'shape') else np.asarray(statevector_b).shape
    if tuple(sa) != tuple(sb):
        raise ValueError(f"statevector shapes differ: {tuple(sa)} vs {tuple(sb)}")
    a = statevector_a
    b = statevector_b
    # both operands in one precision: complex128 unless both are complex64
    both64 = all(getattr(x, 'dtype', np.complex128) == np.complex64 for x in (a, b))
    ha, oa = _device(a if both64 or (isinstance(a, DeviceState) and a.dtype == np.complex128)
                     else np.asarray(a, dtype=np.complex128))
    hb, ob = _device(b if both64 or (isinstance(b, DeviceState) and b.dtype == np.complex128)
                     else np.asarray(b, dtype=np.complex128))
    try:
        re, im = C.c_double(), C.c_double()
        check(lib.desv_inner(ha.ptr, hb.ptr, C.byref(re), C.byref(im)))
        return float(re.value ** 2 + im.value ** 2)
    finally:
        if oa:
            ha.close()
        if ob:
            hb.close()
