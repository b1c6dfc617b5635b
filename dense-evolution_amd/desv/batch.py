"""Heterogeneous VQE batches: many instances of uneven qubit counts, one call.

Generalises ``run_batch_jit`` (reference ``statevector.py:496-584``, one
circuit vmapped over parameter rows) to independent circuits, Hamiltonians
and qubit counts.  Each instance is an ansatz circuit (QASMCircuit), its
theta and a Pauli-sum Hamiltonian; :meth:`Batch.run` returns every energy
and adjoint gradient.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Tuple

import numpy as np

from . import _capi
from ._capi import check, lib
from .autodiff import _build_template
from .observables import encode_pauli_terms


class Batch:
    """``compensated`` (complex64): forward and reverse sweeps with
    compensated matrices (FP32 hi + lo pairs, lo part first); None takes the
    process default (:func:`desv.config.set_compensation`)."""

    def __init__(self, dtype="complex64", device: int = 0, compensated=None):
        self.code = _capi.C64 if np.dtype(dtype) == np.complex64 else _capi.C128
        self.device = int(device)
        p = C.c_void_p()
        check(lib.desv_batch_create(self.code, device, C.byref(p)))
        self.ptr = p
        if compensated is None:
            from .config import default_compensation
            compensated = default_compensation()
        self.set_compensated(compensated)
        self.n_params: List[int] = []
        self.slots: List[np.ndarray] = []
        self._shapes: List[Tuple[int, int, int]] = []
        self.config = {"tile_bits": 0, "fuse_bits": 2, "use_mfma": False, "use_graph": True}

    def close(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            lib.desv_batch_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass

    def add(self, circuit, n_qubits: int, theta, terms) -> int:
        tmpl = _build_template(circuit, n_qubits)
        slots = np.nonzero(tmpl[:, 3] == -1.0)[0]
        pmap = np.full(len(tmpl), -1, dtype=np.int64)
        pmap[slots] = np.arange(len(slots))
        ops = tmpl.copy()
        th = np.asarray(theta, dtype=np.float64).reshape(-1)[:len(slots)]
        ops[slots, 3] = th
        xm, zm, cs = encode_pauli_terms(terms, n_qubits)
        if np.any(np.imag(cs) != 0):
            raise ValueError("batch Hamiltonians must have real coefficients")
        coeff = np.ascontiguousarray(np.real(cs))
        ops = np.ascontiguousarray(ops)
        idx = C.c_int64()
        check(lib.desv_batch_add(self.ptr, n_qubits, _capi.dptr(ops), len(ops), _capi.i64ptr(pmap), len(slots),
                                 _capi.u64ptr(xm), _capi.u64ptr(zm), _capi.dptr(coeff), len(xm), C.byref(idx)))
        self.n_params.append(len(slots))
        self._shapes.append((int(n_qubits), len(ops), len(xm)))
        return idx.value

    def set_params(self, index: int, theta) -> None:
        th = np.ascontiguousarray(np.asarray(theta, dtype=np.float64).reshape(-1)[:self.n_params[index]])
        check(lib.desv_batch_set_params(self.ptr, index, _capi.dptr(th)))

    def set_initial_state(self, index: int, state=None) -> None:
        """Initial state of one instance (host array) or None for |0...0>."""
        if state is None:
            check(lib.desv_batch_set_initial_state(self.ptr, index, None))
            return
        a = np.ascontiguousarray(state, dtype=np.complex64 if self.code == _capi.C64 else np.complex128)
        check(lib.desv_batch_set_initial_state(self.ptr, index, _capi.vptr(a)))

    def set_forward_only(self, enable: bool = True) -> None:
        """Only apply the circuits (no Hamiltonian, no gradient sweep)."""
        check(lib.desv_batch_set_forward_only(self.ptr, int(bool(enable))))

    def run_forward(self) -> None:
        """One forward run of every instance (forward-only batches)."""
        check(lib.desv_batch_run(self.ptr, None, None))

    def get_states(self) -> np.ndarray:
        """Every instance's final state, concatenated in instance order, in
        one device-to-host copy; shaped (B, 2^n) when all widths agree."""
        ns = [s[0] for s in self._shapes]
        out = np.empty(sum(1 << n for n in ns), dtype=np.complex64 if self.code == _capi.C64 else np.complex128)
        check(lib.desv_batch_get_states(self.ptr, _capi.vptr(out)))
        if ns and all(n == ns[0] for n in ns):
            return out.reshape(len(ns), 1 << ns[0])
        return out

    def run_reverse(self) -> List[np.ndarray]:
        """Gradients for the last energies-only run (:meth:`energies`), from
        the psi copy and lambda it left on the device: the reverse sweep alone."""
        g = np.zeros(max(1, sum(self.n_params)), dtype=np.float64)
        check(lib.desv_batch_run_reverse(self.ptr, _capi.dptr(g)))
        out, off = [], 0
        for k in self.n_params:
            out.append(g[off:off + k].copy())
            off += k
        return out

    def state_device_ptr(self, index: int) -> int:
        p = C.c_void_p()
        check(lib.desv_batch_state_device_ptr(self.ptr, index, C.byref(p)))
        return int(p.value or 0)

    def get_state(self, index: int, n_qubits: int) -> np.ndarray:
        out = np.empty(1 << n_qubits, dtype=np.complex64 if self.code == _capi.C64 else np.complex128)
        check(lib.desv_batch_get_state(self.ptr, index, _capi.vptr(out)))
        return out

    def set_compensated(self, enable: bool) -> None:
        """complex64: compensated matrices in both sweeps (rebuilds on the next run)."""
        self.compensated = bool(enable) and self.code == _capi.C64
        check(lib.desv_batch_set_compensated(self.ptr, int(self.compensated)))

    def configure(self, tile_bits: int = 0, use_graph: bool = True, fuse_bits: int = 0,
                  use_mfma: bool = False) -> None:
        """Launch shape for the next run (the instances are recompiled)."""
        check(lib.desv_batch_configure(self.ptr, int(tile_bits), int(fuse_bits), int(bool(use_mfma)),
                                       int(use_graph)))
        self.config = {"tile_bits": int(tile_bits), "fuse_bits": int(fuse_bits) or 2,
                       "use_mfma": bool(use_mfma), "use_graph": bool(use_graph)}

    def signature(self) -> str:
        """Workload class of the batch for the autotuner's cache."""
        return (f"c{'64' if self.code == _capi.C64 else '128'}{'-comp' if self.compensated else ''}"
                f"-inst{len(self.n_params)}-" + _shape_sig(self._shapes))

    def autotune(self, candidates=None, reps: int = 2, cache=True):
        """Time candidate launch shapes and keep the fastest (see desv.tune)."""
        from .tune import autotune_batch
        return autotune_batch(self, candidates=candidates, reps=reps, cache=cache)

    def add_ops(self, n_qubits: int, ops, param_of_op, n_params: int, xmask, zmask, coeff) -> int:
        """Add an instance from an op table and encoded Pauli terms."""
        ops = np.ascontiguousarray(ops, dtype=np.float64).reshape(-1, 4)
        pmap = np.ascontiguousarray(param_of_op, dtype=np.int64)
        xm = np.ascontiguousarray(xmask, dtype=np.uint64)
        zm = np.ascontiguousarray(zmask, dtype=np.uint64)
        cf = np.ascontiguousarray(coeff, dtype=np.float64)
        idx = C.c_int64()
        check(lib.desv_batch_add(self.ptr, n_qubits, _capi.dptr(ops), len(ops), _capi.i64ptr(pmap), n_params,
                                 _capi.u64ptr(xm), _capi.u64ptr(zm), _capi.dptr(cf), len(xm), C.byref(idx)))
        self.n_params.append(int(n_params))
        self._shapes.append((int(n_qubits), len(ops), len(xm)))
        return idx.value

    def energies(self) -> np.ndarray:
        """Forward pass and energies only."""
        e = np.zeros(len(self.n_params), dtype=np.float64)
        check(lib.desv_batch_run(self.ptr, _capi.dptr(e), None))
        return e

    def run(self) -> Tuple[np.ndarray, List[np.ndarray]]:
        e = np.zeros(len(self.n_params), dtype=np.float64)
        g = np.zeros(max(1, sum(self.n_params)), dtype=np.float64)
        check(lib.desv_batch_run(self.ptr, _capi.dptr(e), _capi.dptr(g)))
        out, off = [], 0
        for k in self.n_params:
            out.append(g[off:off + k].copy())
            off += k
        return e, out

    def completion_us(self) -> np.ndarray:
        """Per-instance completion times (microseconds from the start) of the
        last run, from device wall-clock stamps written as each instance's
        results are reduced."""
        out = np.zeros(len(self.n_params), dtype=np.float64)
        check(lib.desv_batch_completion_us(self.ptr, _capi.dptr(out)))
        return out

    def last_time_us(self) -> float:
        v = C.c_double()
        check(lib.desv_batch_last_time_us(self.ptr, C.byref(v)))
        return v.value


def _shape_sig(shapes) -> str:
    """Compact digest of (qubits, ops, terms) per instance."""
    import hashlib
    h = hashlib.sha1(repr(sorted(shapes)).encode()).hexdigest()[:12]
    ns = [x[0] for x in shapes]
    return f"n{min(ns) if ns else 0}-{max(ns) if ns else 0}-{h}"
