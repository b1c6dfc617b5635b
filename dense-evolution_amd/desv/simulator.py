"""DenseSVSimulator: the drop-in engine class, state resident on one GPU.

Same constructor, attributes, methods, validation and error messages as the
reference class (``dense_evolution/backends/statevector.py:70-616``), backed
by the HIP engine through the C ABI.  The amplitudes live in device memory
owned by a ``desv_handle``; ``sim.sv`` is a :class:`DeviceState` view with
``__array__`` (host copy on demand) and stays assignable (``sim.sv = arr``),
as the reference's callers expect.  Gates update the state in place, which
matches the reference's donated-buffer semantics for ``run_circuit_jit``.
"""
from __future__ import annotations

import ctypes as C
import warnings
from typing import List, Optional, Tuple

import numpy as np

from . import _capi
from ._capi import check, lib
from .config import default_compensation
from .gates import GATE_IDS, GATES, PARAMETRIC_GATES, _TWO_QUBIT_PARAMETRIC_GATES
from .transpiler import QuantumTranspiler


class StateHandle:
    """Owner of one device state (``desv_create`` / ``desv_destroy``)."""

    def __init__(self, n: int, complex128: bool, device: int = 0):
        self.n = int(n)
        self.dtype = np.dtype(np.complex128 if complex128 else np.complex64)
        self.code = _capi.C128 if complex128 else _capi.C64
        self.device = int(device)
        h = C.c_void_p()
        check(lib.desv_create(self.n, self.code, self.device, C.byref(h)))
        self.ptr = h

    @property
    def dim(self) -> int:
        return 1 << self.n

    def close(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            lib.desv_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass

    # transfers
    def upload(self, host: np.ndarray) -> None:
        a = np.ascontiguousarray(host, dtype=self.dtype)
        if a.shape != (self.dim,):
            raise ValueError(f"State vector length {a.size} != 2**{self.n} = {self.dim}")
        check(lib.desv_set_state(self.ptr, _capi.vptr(a)))

    def download(self) -> np.ndarray:
        out = np.empty(self.dim, dtype=self.dtype)
        check(lib.desv_get_state(self.ptr, _capi.vptr(out)))
        return out

    def norm2(self) -> float:
        v = C.c_double()
        check(lib.desv_norm2(self.ptr, C.byref(v)))
        return v.value

    def scale(self, z: complex) -> None:
        check(lib.desv_scale(self.ptr, float(np.real(z)), float(np.imag(z))))

    def qubit_probs(self, qubit: int) -> Tuple[float, float]:
        """(sum |amp|^2 over qubit = 0, over qubit = 1), FP64, not normalised."""
        p0, p1 = C.c_double(), C.c_double()
        check(lib.desv_qubit_probs(self.ptr, int(qubit), C.byref(p0), C.byref(p1)))
        return p0.value, p1.value

    def apply_1q_probs(self, mat: np.ndarray, qubit: int, next_qubit: Optional[int] = None):
        """2x2 `mat` on `qubit` in place (one streaming pass); with
        `next_qubit`, the same pass's (p0, p1) populations of that qubit in
        the updated state (FP64, not normalised)."""
        m = np.ascontiguousarray(np.asarray(mat, dtype=np.complex128).reshape(2, 2))
        p0, p1 = C.c_double(), C.c_double()
        nq = -1 if next_qubit is None else int(next_qubit)
        check(lib.desv_apply1q_probs(self.ptr, int(qubit), m.view(np.float64).ctypes.data_as(_capi._pd), nq,
                                     C.byref(p0), C.byref(p1)))
        return None if next_qubit is None else (p0.value, p1.value)

    def apply_layer(self, ops, pop_qubits=None):
        """Apply {qubit: 2x2 matrix} in place in one streaming pass (every
        matrix with at most one nonzero entry per row: Paulis, diagonal and
        amplitude-damping forms, their products); with `pop_qubits` (at most
        12), also return the FP64 populations of the result as an array of
        2^len(pop_qubits) values, index bit j <-> pop_qubits[j] (an empty list:
        the squared norm)."""
        qs = [int(q) for q in ops]
        mats = np.ascontiguousarray(np.array([np.asarray(ops[q], dtype=np.complex128).reshape(2, 2) for q in ops],
                                             dtype=np.complex128).reshape(-1))
        qa = np.ascontiguousarray(np.array(qs, dtype=np.intc))
        if pop_qubits is None:
            check(lib.desv_apply_monomial_layer(self.ptr, len(qs), qa.ctypes.data_as(_capi._pi),
                                                mats.view(np.float64).ctypes.data_as(_capi._pd), 0, None, None))
            return None
        pa = np.ascontiguousarray(np.array([int(q) for q in pop_qubits], dtype=np.intc))
        out = np.empty(1 << len(pa), dtype=np.float64)
        check(lib.desv_apply_monomial_layer(self.ptr, len(qs), qa.ctypes.data_as(_capi._pi),
                                            mats.view(np.float64).ctypes.data_as(_capi._pd), len(pa),
                                            pa.ctypes.data_as(_capi._pi), out.ctypes.data_as(_capi._pd)))
        return out

    def reset(self) -> None:
        check(lib.desv_reset(self.ptr))

    def set_compensated(self, enable: bool) -> None:
        """complex64: apply gate tables and matrices with compensated (FP32
        hi + lo) matrices, lo part accumulated first (no-op for complex128)."""
        check(lib.desv_set_compensated(self.ptr, int(bool(enable))))

    def set_double_buffer(self, enable: bool) -> None:
        """Out-of-place streaming passes into a second state-sized buffer
        (default on); off frees it and updates in place."""
        check(lib.desv_set_double_buffer(self.ptr, int(bool(enable))))

    def copy_from(self, other: "StateHandle") -> None:
        """self <- other on the device (converting complex64 <-> complex128)."""
        check(lib.desv_copy(self.ptr, other.ptr))

    def inner(self, other: "StateHandle") -> complex:
        """<self|other> with FP64 accumulation (same dtype)."""
        re, im = C.c_double(), C.c_double()
        check(lib.desv_inner(self.ptr, other.ptr, C.byref(re), C.byref(im)))
        return complex(re.value, im.value)

    def run_ops(self, ops: np.ndarray) -> None:
        ops = np.ascontiguousarray(ops, dtype=np.float64).reshape(-1, 4)
        if len(ops):
            check(lib.desv_run_ops(self.ptr, _capi.dptr(ops), len(ops)))

    def apply_matrix(self, mat: np.ndarray, qubits) -> None:
        k = len(qubits)
        m = np.ascontiguousarray(np.asarray(mat, dtype=np.complex128).reshape(1 << k, 1 << k))
        q = np.ascontiguousarray(qubits, dtype=np.int32)
        check(lib.desv_apply_matrix(self.ptr, m.view(np.float64).ctypes.data_as(_capi._pd), _capi.iptr(q), k))

    def synchronize(self) -> None:
        check(lib.desv_synchronize(self.ptr))

    def device_ptr(self) -> int:
        """Device address of the amplitudes (changes when a streaming pass
        swaps the double buffer)."""
        p = C.c_void_p()
        check(lib.desv_state_device_ptr(self.ptr, C.byref(p)))
        return int(p.value or 0)


class DeviceState:
    """Array-like view of a device-resident state vector.

    ``np.asarray(view)`` copies the amplitudes to the host in the state's
    dtype; indexing and arithmetic go through that host copy.  Shape, dtype
    and length are answered without a copy.
    """

    __array_priority__ = 100

    def __init__(self, handle: StateHandle):
        self._h = handle

    @property
    def handle(self) -> StateHandle:
        return self._h

    @property
    def shape(self):
        return (self._h.dim,)

    @property
    def dtype(self):
        return self._h.dtype

    @property
    def ndim(self):
        return 1

    @property
    def size(self):
        return self._h.dim

    def __len__(self):
        return self._h.dim

    def __array__(self, dtype=None, copy=None):
        a = self._h.download()
        return a if dtype is None else a.astype(dtype)

    def get(self) -> np.ndarray:
        return self._h.download()

    @property
    def __cuda_array_interface__(self):
        """Zero-copy device view (version 3 of the array interface, which
        PyTorch on ROCm reads as well): the amplitudes as a (2^n, 2) float
        array of (re, im) pairs.  Pending work on the state is finished first.
        The view aliases the current buffer: a later gate application may move
        the state to the other half of its double buffer."""
        if hasattr(self._h, "synchronize"):
            self._h.synchronize()
        ptr = self._h.device_ptr()
        real = np.float32 if self.dtype == np.complex64 else np.float64
        return {"shape": (self._h.dim, 2), "typestr": np.dtype(real).str, "data": (ptr, False),
                "version": 3, "strides": None, "stream": None}

    def to_torch(self):
        """A complex torch tensor on the state's device that aliases the
        amplitudes (no copy; see ``__cuda_array_interface__``)."""
        import torch
        t = torch.as_tensor(self, device=f"cuda:{getattr(self._h, 'device', 0)}")
        return torch.view_as_complex(t)

    def __getitem__(self, key):
        return np.asarray(self)[key]

    def reshape(self, *shape):
        return np.asarray(self).reshape(*shape)

    def ravel(self):
        return np.asarray(self)

    def conj(self):
        return np.asarray(self).conj()

    def copy(self):
        return np.asarray(self).copy()

    def __iter__(self):
        return iter(np.asarray(self))

    def _bin(op):
        def f(self, other):
            return op(np.asarray(self), np.asarray(other) if isinstance(other, DeviceState) else other)
        return f

    def _rbin(op):
        def f(self, other):
            return op(other, np.asarray(self))
        return f

    import operator as _op
    __add__, __sub__, __mul__, __truediv__ = (_bin(_op.add), _bin(_op.sub), _bin(_op.mul), _bin(_op.truediv))
    __radd__, __rsub__, __rmul__, __rtruediv__ = (_rbin(_op.add), _rbin(_op.sub), _rbin(_op.mul),
                                                   _rbin(_op.truediv))
    __matmul__, __rmatmul__ = _bin(_op.matmul), _rbin(_op.matmul)
    __abs__ = lambda self: np.abs(np.asarray(self))  # noqa: E731
    __neg__ = lambda self: -np.asarray(self)  # noqa: E731
    del _op, _bin, _rbin

    def __repr__(self):
        return f"DeviceState(n={self._h.n}, dtype={self._h.dtype})"


def _op_name(cmd) -> str:
    return cmd[0].lower() if isinstance(cmd[0], str) else str(cmd[0]).lower()


class DenseSVSimulator:
    """Dense statevector simulator; qubit 0 is the most significant bit.

    Parameters
    ----------
    n_qubits : int in [1, 34]
    use_float32 : complex64 state if True, complex128 otherwise
    device : HIP device index
    compensated : complex64 only: apply gates with compensated matrices (each
        FP32 matrix carried as hi + lo, lo products accumulated first), so the
        rounding of a repeated matrix does not accumulate coherently over long
        circuits (Trotter evolutions); None takes the process default
        (:func:`desv.config.set_compensation`, on unless changed).
    """

    def __init__(self, n_qubits: int, use_float32: bool = False, device: int = 0,
                 compensated: Optional[bool] = None):
        if n_qubits < 1 or n_qubits > 34:
            raise ValueError(f"n_qubits must be in [1, 34], got {n_qubits}")
        self.n = n_qubits
        self.dim = 1 << n_qubits
        self.use_float32 = use_float32
        self.dtype = np.complex64 if use_float32 else np.complex128
        self.xp = np
        self._h = StateHandle(n_qubits, not use_float32, device)
        if compensated is None:
            compensated = default_compensation()
        self.compensated = bool(compensated) and bool(use_float32)
        self._h.set_compensated(self.compensated)

    # -- state -------------------------------------------------------------
    @property
    def sv(self) -> DeviceState:
        return DeviceState(self._h)

    @sv.setter
    def sv(self, value) -> None:
        if isinstance(value, DeviceState):
            if value.handle is self._h:
                return
            value = value.get()
        self._h.upload(np.asarray(value))

    @property
    def handle(self) -> StateHandle:
        return self._h

    def close(self) -> None:
        self._h.close()

    def _reset_sv(self):
        self._h.reset()

    def set_initial_state(self, state: Optional[np.ndarray] = None):
        """Reset to |0...0> (None) or load `state` normalised in the state dtype."""
        if state is None:
            self._reset_sv()
            return
        state = np.asarray(state, dtype=self.dtype)
        if state.shape != (self.dim,):
            raise ValueError(f"State vector length {len(state)} != 2**{self.n} = {self.dim}")
        norm = np.linalg.norm(state)
        if norm < 1e-12:
            raise ValueError("Cannot set a zero-norm state vector")
        self._h.upload(state / norm)

    def set_state(self, state: np.ndarray):
        self.set_initial_state(state)

    def normalize(self):
        norm = float(np.sqrt(self._h.norm2()))
        if norm > 1e-12:
            self._h.scale(1.0 / norm)

    def _check_qubit_range(self, qubit: float, context: str) -> None:
        qi = int(qubit)
        if not 0 <= qi < self.n:
            raise ValueError(f"Qubit index {qi} out of range [0, {self.n}) in {context}")

    # -- gates -------------------------------------------------------------
    def apply_gate_1q(self, gate: np.ndarray, qubit: int):
        if not 0 <= qubit < self.n:
            raise ValueError(f"Qubit index {qubit} out of range [0, {self.n})")
        self._h.apply_matrix(np.asarray(gate, dtype=self.dtype), [qubit])

    def apply_gate_2q(self, gate: np.ndarray, q1: int, q2: int):
        if q1 == q2:
            raise ValueError("Control and target qubits must differ")
        if not (0 <= q1 < self.n and 0 <= q2 < self.n):
            raise ValueError(f"Qubit indices ({q1},{q2}) out of range [0, {self.n})")
        self._h.apply_matrix(np.asarray(gate, dtype=self.dtype).reshape(4, 4), [q1, q2])

    def apply_gate(self, gate: np.ndarray, qubits) -> None:
        """k-qubit unitary (k <= 5), first qubit = most significant index bit."""
        qubits = [int(q) for q in qubits]
        for q in qubits:
            if not 0 <= q < self.n:
                raise ValueError(f"Qubit index {q} out of range [0, {self.n})")
        self._h.apply_matrix(np.asarray(gate, dtype=self.dtype), qubits)

    def _two_qubit_checks(self, ctrl, tgt):
        if ctrl == tgt:
            raise ValueError("Control and target qubits must differ")
        if not (0 <= ctrl < self.n and 0 <= tgt < self.n):
            raise ValueError(f"Qubit indices ({ctrl},{tgt}) out of range [0, {self.n})")

    def apply_cx(self, ctrl: int, tgt: int):
        self._two_qubit_checks(ctrl, tgt)
        self._h.run_ops(np.array([[20, ctrl, tgt, 0.0]]))

    def apply_cz(self, ctrl: int, tgt: int):
        self._two_qubit_checks(ctrl, tgt)
        self._h.run_ops(np.array([[21, ctrl, tgt, 0.0]]))

    def apply_rx(self, qubit: int, theta: float):
        self.apply_gate_1q(PARAMETRIC_GATES['rx'](theta), qubit)

    def apply_ry(self, qubit: int, theta: float):
        self.apply_gate_1q(PARAMETRIC_GATES['ry'](theta), qubit)

    def apply_rz(self, qubit: int, theta: float):
        self.apply_gate_1q(PARAMETRIC_GATES['rz'](theta), qubit)

    # -- measurement ---------------------------------------------------------
    def measure(self, qubit_idx: int, jax_key=None, rng: Optional[np.random.Generator] = None) -> int:
        """Projective measurement of one qubit; collapses and renormalises.

        The outcome is drawn with ``rng`` if given, else with NumPy's global
        ``np.random.choice`` (the reference's default, statevector.py:341).
        """
        if not 0 <= qubit_idx < self.n:
            raise ValueError(f"Qubit {qubit_idx} out of range [0, {self.n})")
        if jax_key is not None:
            raise ValueError("measure(jax_key=...) requires JAX to be installed.")
        p0, p1 = C.c_double(), C.c_double()
        check(lib.desv_qubit_probs(self._h.ptr, qubit_idx, C.byref(p0), C.byref(p1)))
        total = p0.value + p1.value
        if total < 1e-12:
            raise RuntimeError("Statevector norm is zero — cannot measure")
        pr0, pr1 = p0.value / total, p1.value / total
        if rng is not None:
            result = int(rng.choice([0, 1], p=[pr0, pr1]))
        else:
            result = int(np.random.choice([0, 1], p=[pr0, pr1]))
        keep = pr0 if result == 0 else pr1
        # zero the other half and divide by the kept half's norm, sqrt(keep *
        # total) = sqrt(sum |a|^2 over the outcome): the reference's mask then
        # sv / norm(sv) (statevector.py:343-358), in one pass
        check(lib.desv_collapse(self._h.ptr, qubit_idx, result, 1.0 / np.sqrt(keep * total)))
        return result

    # -- circuits --------------------------------------------------------------
    def run_circuit(self, circuit: List[Tuple], transpile: bool = True):
        target = QuantumTranspiler.transpile(circuit) if transpile else circuit
        if all(_op_name(cmd) in GATE_IDS for cmd in target):
            self.run_circuit_jit(target)
            return
        for cmd in target:
            name = cmd[0].lower()
            args = cmd[1:]
            if name in GATES:
                mat = GATES[name]
                if mat.shape == (2, 2):
                    self.apply_gate_1q(mat, int(args[0]))
                elif mat.shape == (4, 4):
                    self.apply_gate_2q(mat, int(args[0]), int(args[1]))
                else:
                    self.apply_gate(mat, [int(a) for a in args[:3]])
            elif name in PARAMETRIC_GATES:
                if name in _TWO_QUBIT_PARAMETRIC_GATES:
                    self.apply_gate_2q(PARAMETRIC_GATES[name](*args[2:]), int(args[0]), int(args[1]))
                else:
                    self.apply_gate_1q(PARAMETRIC_GATES[name](*args[1:]), int(args[0]))
            else:
                raise ValueError(
                    f"unknown gate '{cmd[0]}' -- not in GATES or PARAMETRIC_GATES. "
                    f"A typo in a gate name used to be silently dropped from the "
                    f"circuit instead of raising (issue #4).")

    def compile_ops(self, circuit: List) -> np.ndarray:
        """Op table ``[g_id, q1, q2, param]`` of a circuit (after transpile)."""
        return encode_op_table(QuantumTranspiler.transpile(circuit), self.n)

    def run_circuit_jit(self, circuit: List):
        ops = self.compile_ops(circuit)
        if len(ops):
            self._h.run_ops(ops)

    def run_circuit_jit_beast_mode(self, circuit: List):
        warnings.warn(
            "run_circuit_jit_beast_mode is deprecated, use run_circuit_jit instead "
            "(same behavior, shorter name). This alias will be removed in a future release.",
            DeprecationWarning, stacklevel=2)
        return self.run_circuit_jit(circuit)

    def run_circuit_with_chunking(self, circuit: List, chunk_size: int = 500):
        target = QuantumTranspiler.transpile(circuit)
        for i in range(0, len(target), chunk_size):
            self.run_circuit_jit(target[i:i + chunk_size])

    def run_batch_jit(self, base_circuit: List, parameter_batch: np.ndarray) -> np.ndarray:
        """Run `base_circuit` once per parameter row from |0...0>; returns (B, 2^n).

        Every rx/ry/rz/p/u1/phase/cp/crz/cphase gate consumes one column in
        order, even when the circuit gives a literal angle (reference
        statevector.py:516-547).
        """
        template = encode_op_table(QuantumTranspiler.transpile(base_circuit), self.n, param_slots=True)
        slots = np.nonzero(template[:, 3] == -1.0)[0]
        parameter_batch = np.asarray(parameter_batch)
        if parameter_batch.ndim != 2 or parameter_batch.shape[1] != len(slots):
            raise ValueError(
                f"parameter_batch has {parameter_batch.shape[-1] if parameter_batch.ndim else 0} "
                f"column(s) but base_circuit has {len(slots)} parametric gate(s) (rx/ry/rz/p/u1/"
                f"phase/cp/crz/cphase) -- one column per parametric gate, in gate-appearance order. "
                f"A literal float passed for one of these gates is NOT exempt: it still consumes a "
                f"positional column.")
        rows = np.asarray(parameter_batch, dtype=np.float64)
        if rows.shape[0] == 0:
            return np.empty((0, self.dim), dtype=self.dtype)
        # every row an instance of one forward-only device batch: the rows
        # share one compiled plan (same gate structure), their matrices are
        # refreshed from their own angles, each stage runs for all rows in one
        # launch, and the (B, 2^n) result comes back in one copy
        from .batch import Batch
        b = Batch("complex64" if self.use_float32 else "complex128", self._h.device,
                  compensated=self.compensated)
        try:
            b.set_forward_only(True)
            b.configure(use_graph=False)
            pmap = np.full(len(template), -1, dtype=np.int64)
            none_u = np.zeros(0, dtype=np.uint64)
            none_d = np.zeros(0, dtype=np.float64)
            for row in rows:
                ops = template.copy()
                ops[slots, 3] = row
                b.add_ops(self.n, ops, pmap, 0, none_u, none_u, none_d)
            b.run_forward()
            return b.get_states().reshape(len(rows), self.dim)
        finally:
            b.close()

    def run_parametric_batch_jit(self, base_circuit: List, parameter_batch: np.ndarray) -> np.ndarray:
        warnings.warn(
            "run_parametric_batch_jit is deprecated, use run_batch_jit instead "
            "(same behavior, shorter name). This alias will be removed in a future release.",
            DeprecationWarning, stacklevel=2)
        return self.run_batch_jit(base_circuit, parameter_batch)

    # -- readout ---------------------------------------------------------------
    def get_probabilities(self) -> np.ndarray:
        """|a|^2 (FP64 from the device), clipped to [0, 1] and renormalised."""
        p = np.empty(self.dim, dtype=np.float64)
        check(lib.desv_probabilities(self._h.ptr, _capi.dptr(p)))
        p = np.clip(p, 0.0, 1.0)
        total = p.sum()
        if total > 1e-12:
            p /= total
        return p

    def get_statevector(self) -> np.ndarray:
        return self._h.download()

    def memory_mb(self) -> float:
        return self.dim * (8 if self.use_float32 else 16) / 1_000_000


_ONE_Q_PARAM = ('rx', 'ry', 'rz', 'p', 'u1', 'phase', 'gphase')


def encode_op_table(target: List, n: int, param_slots: bool = False) -> np.ndarray:
    """Rows ``[g_id, q1, q2, param]`` for an already transpiled circuit.

    Same row layout, argument positions and qubit checks as the reference's
    run_circuit_jit (statevector.py:422-463).  With ``param_slots`` the angle
    of every rx/ry/rz/p/u1/phase/cp/crz/cphase row is the sentinel -1.0 that
    run_batch_jit fills from the parameter batch (statevector.py:516-535).
    """
    rows = []

    def chk(q, ctx):
        qi = int(q)
        if not 0 <= qi < n:
            raise ValueError(f"Qubit index {qi} out of range [0, {n}) in {ctx}")

    for cmd in target:
        name = _op_name(cmd)
        if name not in GATE_IDS:
            raise ValueError(
                f"unknown gate '{cmd[0]}' -- not in GATE_IDS. A typo in a gate "
                f"name used to be silently dropped from the circuit instead of "
                f"raising (issue #4).")
        gid = float(GATE_IDS[name])
        args = cmd[1:]
        if name in _ONE_Q_PARAM:
            q1 = float(args[0])
            chk(q1, f"gate '{name}'")
            if param_slots and name != 'gphase':
                p = -1.0
            elif param_slots:
                p = 0.0
            else:
                p = float(args[1]) if len(args) > 1 else 0.0
            rows.append([gid, q1, 0.0, p])
        elif name in ('cp', 'crz', 'cphase'):
            c, t = float(args[0]), float(args[1]) if len(args) > 1 else 0.0
            chk(c, f"gate '{name}' (control)")
            chk(t, f"gate '{name}' (target)")
            p = -1.0 if param_slots else (float(args[2]) if len(args) > 2 else 0.0)
            rows.append([gid, c, t, p])
        elif name in ('cx', 'cz', 'swap', 'cy'):
            c, t = float(args[0]), float(args[1]) if len(args) > 1 else 0.0
            chk(c, f"gate '{name}' (control)")
            chk(t, f"gate '{name}' (target)")
            rows.append([gid, c, t, 0.0])
        else:
            q1 = float(args[0]) if args else 0.0
            chk(q1, f"gate '{name}'")
            rows.append([gid, q1, 0.0, 0.0])
    return np.array(rows, dtype=np.float64).reshape(-1, 4)
