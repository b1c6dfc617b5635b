"""Compiled gate programs: build once, replay many times.

A :class:`Program` collects op-table rows and explicit matrices, is compiled
by the engine's host compiler (gate fusion into blocks of up to
``fuse_bits`` qubits, then cache-blocked stages of ``2^tile_bits``-amplitude
LDS tiles) and replayed on a state with one kernel launch per stage.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional

import numpy as np

from . import _capi
from ._capi import check, lib
from .simulator import StateHandle, encode_op_table
from .transpiler import QuantumTranspiler


class Program:
    def __init__(self, n_qubits: int):
        self.n = int(n_qubits)
        p = C.c_void_p()
        check(lib.desv_program_create(self.n, C.byref(p)))
        self.ptr = p

    def close(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            lib.desv_program_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass

    def add_ops(self, ops) -> "Program":
        ops = np.ascontiguousarray(ops, dtype=np.float64).reshape(-1, 4)
        check(lib.desv_program_add_ops(self.ptr, _capi.dptr(ops), len(ops)))
        return self

    def add_circuit(self, circuit) -> "Program":
        return self.add_ops(encode_op_table(QuantumTranspiler.transpile(circuit), self.n))

    def add_matrix(self, mat, qubits) -> "Program":
        k = len(qubits)
        m = np.ascontiguousarray(np.asarray(mat, dtype=np.complex128).reshape(1 << k, 1 << k))
        q = np.ascontiguousarray(qubits, dtype=np.int32)
        check(lib.desv_program_add_matrix(self.ptr, m.view(np.float64).ctypes.data_as(_capi._pd),
                                          _capi.iptr(q), k))
        return self

    def compile(self, dtype="complex64", fuse_bits: Optional[int] = None, tile_bits: Optional[int] = None,
                low_bits: Optional[int] = None, per_gate: bool = False, use_mfma: bool = False) -> "Program":
        code = _capi.C64 if np.dtype(dtype) == np.complex64 else _capi.C128
        o = _capi.CompileOpts()
        check(lib.desv_default_opts(code, C.byref(o)))
        if fuse_bits is not None:
            o.fuse_bits = int(fuse_bits)
        if tile_bits is not None:
            o.tile_bits = int(tile_bits)
        if low_bits is not None:
            o.low_bits = int(low_bits)
        o.per_gate = int(bool(per_gate))
        o.use_mfma = int(bool(use_mfma))
        o.dtype = code
        check(lib.desv_program_compile(self.ptr, C.byref(o)))
        return self

    def stats(self) -> Dict[str, int]:
        g, p, b, m = (C.c_int64() for _ in range(4))
        check(lib.desv_program_stats(self.ptr, C.byref(g), C.byref(p), C.byref(b), C.byref(m)))
        return {"gates": g.value, "passes": p.value, "blocks": b.value, "mfma_blocks": m.value}

    def plan(self):
        """(tile_masks u64[S], stage_ops i64[S,18] = (begin, end, sperm[16]),
        ops i64[O,11], mats complex128[M])."""
        s, o, m = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib.desv_program_plan_sizes(self.ptr, C.byref(s), C.byref(o), C.byref(m)))
        tm = np.zeros(s.value, dtype=np.uint64)
        so = np.zeros((s.value, 18), dtype=np.int64)
        ops = np.zeros((o.value, 11), dtype=np.int64)
        mats = np.zeros(2 * m.value, dtype=np.float64)
        check(lib.desv_program_plan_dump(self.ptr, _capi.u64ptr(tm), _capi.i64ptr(so), _capi.i64ptr(ops),
                                         _capi.dptr(mats)))
        return tm, so, ops, mats.view(np.complex128)

    def run(self, target) -> None:
        h = target.handle if not isinstance(target, StateHandle) else target
        check(lib.desv_run_program(h.ptr, self.ptr))
