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
                low_bits: Optional[int] = None, per_gate: bool = False, use_mfma: bool = False,
                compensated: bool = False, reg_groups: bool = False,
                plan_search: Optional[int] = None) -> "Program":
        """Compile for a state of `dtype`.  ``compensated`` (complex64):
        matrices uploaded as FP32 hi + lo pairs and applied lo part first, for
        long programs whose repeated matrices would otherwise accumulate their
        FP32 rounding coherently; about twice the gate arithmetic.
        ``reg_groups`` (complex64 with ``use_mfma``): 1 = eligible stages keep
        the tile in registers between LDS round trips, 2 = eligible stages of
        2^11- or 2^12-amplitude tiles hold one tile per wave in registers from
        HBM load to HBM store (``desv_compile_opts``).  ``plan_search``:
        randomised planner passes tried beside the greedy one (None: the
        library default, 0: greedy only)."""
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
        o.compensated = int(bool(compensated))
        o.reg_groups = int(reg_groups)
        if plan_search is not None:
            o.plan_search = int(plan_search)
        o.dtype = code
        check(lib.desv_program_compile(self.ptr, C.byref(o)))
        self._reg_mode = int(reg_groups) if use_mfma and code == _capi.C64 else 0
        return self

    def stats(self) -> Dict[str, int]:
        """gates, passes (HBM sweeps per run), blocks, mfma_blocks, and
        passes_in_place: the sweeps on a handle without a second buffer (a
        relabelled tail then runs in its in-place form)."""
        g, p, b, m = (C.c_int64() for _ in range(4))
        check(lib.desv_program_stats(self.ptr, C.byref(g), C.byref(p), C.byref(b), C.byref(m)))
        fb_from, n_fb = C.c_int64(), C.c_int64()
        check(lib.desv_program_tail_dump(self.ptr, 0, C.byref(fb_from), C.byref(n_fb), None, None, None, None,
                                         None, None))
        in_place = p.value if fb_from.value < 0 else fb_from.value + n_fb.value
        return {"gates": g.value, "passes": p.value, "blocks": b.value, "mfma_blocks": m.value,
                "passes_in_place": in_place}

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

    def reg_program(self, compensated: bool = False):
        """The register-group program the upload would build: (stage_word
        i32[S] (-1: LDS path), stage_frag i32[S], words i32[W], frags
        f32[F/256, 64, 4])."""
        w, f = C.c_int64(), C.c_int64()
        check(lib.desv_program_reg_sizes(self.ptr, int(compensated), C.byref(w), C.byref(f)))
        s, o, m = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib.desv_program_plan_sizes(self.ptr, C.byref(s), C.byref(o), C.byref(m)))
        sw = np.zeros(max(1, s.value), dtype=np.int32)
        sf = np.zeros(max(1, s.value), dtype=np.int32)
        words = np.zeros(max(1, w.value), dtype=np.int32)
        frags = np.zeros(max(1, f.value), dtype=np.float32)
        check(lib.desv_program_reg_dump(self.ptr, int(compensated), _capi.iptr(sw), _capi.iptr(sf),
                                        _capi.iptr(words), frags.ctypes.data_as(C.POINTER(C.c_float))))
        return sw[:s.value], sf[:s.value], words[:w.value], frags[:f.value].reshape(-1, 64, 4)

    def tail(self, compensated: bool = False):
        """The relabelled tail of a wave-tile program (``desv_program_tail_dump``):
        dict with fb_from (-1: none), store_masks u64[S], omaps i32[S, 64], and
        for the in-place form of the tail fb_tile_masks u64[F], fb_stage_ops
        i64[F, 18], fb_stage_word / fb_stage_frag i32[F] (offsets into
        ``reg_program``'s words and fragment blocks)."""
        s, o, m = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib.desv_program_plan_sizes(self.ptr, C.byref(s), C.byref(o), C.byref(m)))
        fb_from, n_fb = C.c_int64(), C.c_int64()
        check(lib.desv_program_tail_dump(self.ptr, int(compensated), C.byref(fb_from), C.byref(n_fb), None, None,
                                         None, None, None, None))
        S, F = s.value, n_fb.value
        sm = np.zeros(max(1, S), dtype=np.uint64)
        om = np.zeros((max(1, S), 64), dtype=np.int32)
        ftm = np.zeros(max(1, F), dtype=np.uint64)
        fso = np.zeros((max(1, F), 18), dtype=np.int64)
        fsw = np.zeros(max(1, F), dtype=np.int32)
        fsf = np.zeros(max(1, F), dtype=np.int32)
        check(lib.desv_program_tail_dump(self.ptr, int(compensated), C.byref(fb_from), C.byref(n_fb),
                                         _capi.u64ptr(sm), _capi.iptr(om), _capi.u64ptr(ftm), _capi.i64ptr(fso),
                                         _capi.iptr(fsw), _capi.iptr(fsf)))
        return {"fb_from": fb_from.value, "store_masks": sm[:S], "omaps": om[:S], "fb_tile_masks": ftm[:F],
                "fb_stage_ops": fso[:F], "fb_stage_word": fsw[:F], "fb_stage_frag": fsf[:F]}

    def run(self, target) -> None:
        h = target.handle if not isinstance(target, StateHandle) else target
        check(lib.desv_run_program(h.ptr, self.ptr))

    def run_timed(self, target) -> np.ndarray:
        """Run once with an event between stages; per-stage milliseconds."""
        h = target.handle if not isinstance(target, StateHandle) else target
        s, o, m = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib.desv_program_plan_sizes(self.ptr, C.byref(s), C.byref(o), C.byref(m)))
        out = np.zeros(max(1, s.value), dtype=np.float32)
        check(lib.desv_run_program_timed(h.ptr, self.ptr, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out[:s.value].astype(np.float64)

    def stage_kernels(self):
        """Per stage: (kernel name, algorithmic FLOPs, algorithmic HBM bytes)
        for a state of this program's width in `dtype` (8 or 16 bytes per
        amplitude).  FLOPs: N 2^k 8 per dense k-qubit block, 6 N per diagonal
        or scalar op; bytes: one read and one write of the state per stage."""
        tm, so, ops, _ = self.plan()
        mode = getattr(self, "_reg_mode", 0)
        regs = self.reg_program()[0] if mode else np.full(len(tm), -1)
        out = []
        for i in range(len(tm)):
            name, flops, mflops = _stage_kernel(so[i], ops, self.n)
            if regs[i] >= 0 and name != "k_direct":
                # register-resident tile: every block on MFMA
                name = "k_wstage" if mode == 2 else "k_stage<regs>"
                mflops = flops
            out.append((name, flops, mflops))
        return out


_OPK_DENSE1, _OPK_DENSE2, _OPK_DENSEK, _OPK_DIAG, _OPK_SCALAR, _OPK_GRAD, _OPK_GRADDIAG, _OPK_MFMA = range(8)


def _stage_kernel(stage, ops, n):
    """Which kernel runs a stage (mirrors the launch choice of the engine) and
    its algorithmic FLOPs per amplitude-count N = 2^n."""
    b, e = int(stage[0]), int(stage[1])
    sperm = [int(x) for x in stage[2:] if x >= 0]
    permuted = any(sperm[j] != j for j in range(len(sperm)))
    N = 1 << n
    flops = 0
    kinds = set()
    for o in ops[b:e]:
        kind, k = int(o[0]), int(o[1])
        kinds.add(kind)
        if kind in (_OPK_DENSE1, _OPK_DENSE2, _OPK_DENSEK, _OPK_MFMA):
            flops += N * (1 << k) * 8
        else:
            flops += 6 * N
    if e - b == 1 and not permuted and int(ops[b][0]) in (_OPK_DENSE1, _OPK_DENSE2, _OPK_DIAG) \
            and 1 <= int(ops[b][1]) <= 2:
        name = "k_direct"
    elif _OPK_MFMA in kinds:
        name = "k_stage<mfma>"
    elif _OPK_DENSEK in kinds:
        name = "k_stage<valu-k>"
    else:
        name = "k_stage<light>"
    mfma_flops = sum(N * (1 << int(o[1])) * 8 for o in ops[b:e] if int(o[0]) == _OPK_MFMA)
    return name, flops, mfma_flops
