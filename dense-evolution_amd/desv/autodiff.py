"""Differentiable VQE energies: ``circuit_to_energy_fn`` with an adjoint backward.

The reference builds a ``[g_id, q1, q2, sentinel]`` template from a
``QASMCircuit`` (``solvers/autodiff.py:33-89``), injects ``theta`` into the
sentinel rows in order (``:150-157``) and returns
``energy_fn(theta, h_matrix, stato_zero=None, noise=None) -> (energy, sv)``
with ``energy = Re <psi|H|psi>`` (``:166``); callers differentiate it with
``jax.value_and_grad``.  Here the same template and injection order feed the
engine's adjoint driver (``desv_adjoint``): one forward sweep, one
``H|psi>`` build and one reverse sweep give the energy and every parameter
gradient.  With a torch tensor ``theta`` the energy is returned as a torch
scalar whose backward is that precomputed gradient
(:class:`AdjointEnergy`); with a NumPy ``theta`` a float is returned and
``energy_fn.value_and_grad`` gives ``(energy, grad)``.

``h_matrix`` is a :class:`PauliSumOperator` (GPU Pauli kernels) or a dense
``(2^n, 2^n)`` array (n <= 14, device matvec).  All reductions are FP64; the
state precision defaults to complex128 like the reference (``:30``) and can
be set to complex64 (``dtype='complex64'``).
"""
from __future__ import annotations

import ctypes as C
from typing import Callable, Optional, Tuple

import numpy as np

from . import _capi
from ._capi import check, lib
from .config import default_float64
from .gates import GATE_IDS
from .observables import PauliSumOperator
from .simulator import DeviceState, StateHandle
from .transpiler import QuantumTranspiler

#: names that take a theta entry (reference autodiff.py:21)
_PARAMETRIC_GATES = ('rx', 'ry', 'rz', 'u1', 'p', 'cp', 'crz')
_TWO_QUBIT_GATES = ('cx', 'cy', 'cz', 'cp', 'crz', 'swap')


def _build_template(circuit, n_qubits: int) -> np.ndarray:
    """Structural template rows; -1.0 marks a theta slot.

    Only names and qubits are read (literal angles are ignored); ops touching
    a qubit >= n_qubits are dropped; names outside GATE_IDS raise.
    """
    tuples = []
    for op in circuit.ops:
        name = str(op['name']).lower().strip()
        qubits = [int(q) for q in op.get('qubits', [])]
        if not qubits or any(q >= n_qubits for q in qubits):
            continue
        tuples.append((name, *qubits))
    rows = []
    for cmd in QuantumTranspiler.transpile(tuples):
        name = cmd[0].lower()
        if name not in GATE_IDS:
            raise ValueError(
                f"circuit_to_energy_fn: gate '{name}' has no GATE_IDS entry, so it can't be "
                f"represented in this template. Multi-parameter gates like 'u2'/'u3' aren't "
                f"supported here (one parameter per row) -- decompose them into rx/ry/rz/cx "
                f"first, or use DenseSVSimulator.run_circuit instead.")
        sentinel = -1.0 if name in _PARAMETRIC_GATES else 0.0
        qs = cmd[1:]
        if name in _TWO_QUBIT_GATES and len(qs) >= 2:
            rows.append([GATE_IDS[name], qs[0], qs[1], sentinel])
        elif qs:
            rows.append([GATE_IDS[name], qs[0], 0.0, sentinel])
    return np.array(rows, dtype=np.float64).reshape(-1, 4)


class AdjointProgram:
    """A parametric op table bound to a Hamiltonian, evaluated by the adjoint kernel."""

    def __init__(self, template: np.ndarray, n_qubits: int, complex128: bool = True, device: int = 0):
        self.template = np.ascontiguousarray(template, dtype=np.float64).reshape(-1, 4)
        self.n = int(n_qubits)
        self.slots = np.nonzero(self.template[:, 3] == -1.0)[0]
        self.param_of_op = np.full(len(self.template), -1, dtype=np.int64)
        self.param_of_op[self.slots] = np.arange(len(self.slots))
        self.n_params = len(self.slots)
        self.handle = StateHandle(self.n, complex128, device)
        from .config import default_compensation
        self.compensated = default_compensation() and not complex128
        self.handle.set_compensated(self.compensated)

    def ops_for(self, theta) -> np.ndarray:
        theta = np.asarray(theta, dtype=np.float64).reshape(-1)
        if theta.shape[0] != self.n_params:
            raise ValueError(f"theta has {theta.shape[0]} entries, the circuit has {self.n_params} "
                             f"parametric gates")
        ops = self.template.copy()
        ops[self.slots, 3] = theta
        ops[self.template[:, 3] == 0.0, 3] = 0.0
        return ops

    # -- Pauli-sum Hamiltonians: a one-instance batch (compiled once, angles
    #    refreshed per call, launches replayed from a captured hipGraph) --------
    def _batch_for(self, hamiltonian):
        from .batch import Batch
        key = id(hamiltonian)
        if getattr(self, "_bkey", None) != key:
            xm, zm, cs = hamiltonian.encoded()
            if np.any(np.imag(cs) != 0):
                raise ValueError("Pauli-sum Hamiltonians must have real coefficients")
            b = Batch("complex128" if self.handle.code == _capi.C128 else "complex64", self.handle.device,
                      compensated=self.compensated)
            b.add_ops(self.n, self.ops_for(np.zeros(self.n_params)), self.param_of_op, self.n_params,
                      xm, zm, np.real(cs))
            self._batch, self._bkey, self._bham = b, key, hamiltonian
        return self._batch

    def autotune(self, theta, hamiltonian, candidates=None, reps: int = 3, cache: bool = True):
        """Choose the launch shape of this program's adjoint evaluation for a
        Pauli-sum Hamiltonian by timing candidates (see desv.tune)."""
        b = self._prep_batch(theta, hamiltonian, None)
        return b.autotune(candidates=candidates, reps=reps, cache=cache)

    def _prep_batch(self, theta, hamiltonian, initial_state):
        b = self._batch_for(hamiltonian)
        b.set_params(0, np.asarray(theta, dtype=np.float64).reshape(-1))
        if initial_state is None:
            b.set_initial_state(0, None)
        else:
            b.set_initial_state(0, initial_state.get() if isinstance(initial_state, DeviceState)
                                else np.asarray(initial_state))
        self._last_state = _BatchState(b, self.n, self.handle.dtype)
        return b

    def state(self):
        """DeviceState view of the state after the last evaluation."""
        return DeviceState(getattr(self, "_last_state", None) or self.handle)

    def value_and_grad(self, theta, hamiltonian, initial_state=None,
                       tail: Optional[np.ndarray] = None) -> Tuple[float, np.ndarray]:
        """Energy and gradient; `tail` (op-table rows without parameters, e.g.
        a Pauli noise trajectory) runs after the circuit."""
        self._token = getattr(self, "_token", 0) + 1
        if isinstance(hamiltonian, PauliSumOperator) and (tail is None or len(tail) == 0):
            self.ops_for(theta)  # validates the length
            b = self._prep_batch(theta, hamiltonian, initial_state)
            e, g = b.run()
            return float(e[0]), g[0]
        self._last_state = None
        ops = self.ops_for(theta)
        param_of_op = self.param_of_op
        if tail is not None and len(tail):
            tail = np.asarray(tail, dtype=np.float64).reshape(-1, 4)
            ops = np.concatenate([ops, tail])
            param_of_op = np.concatenate([param_of_op, np.full(len(tail), -1, dtype=np.int64)])
        ops = np.ascontiguousarray(ops)
        h = self.handle
        if initial_state is None:
            h.reset()
        elif isinstance(initial_state, DeviceState):
            h.upload(initial_state.get())
        else:
            h.upload(np.asarray(initial_state))
        grad = np.zeros(self.n_params, dtype=np.float64)
        energy = C.c_double()
        if isinstance(hamiltonian, PauliSumOperator):
            xm, zm, cs = hamiltonian.encoded()
            if np.any(np.imag(cs) != 0):
                raise ValueError("Pauli-sum Hamiltonians must have real coefficients")
            coeff = np.ascontiguousarray(np.real(cs), dtype=np.float64)
            check(lib.desv_adjoint(h.ptr, _capi.dptr(ops), len(ops), _capi.i64ptr(param_of_op),
                                   self.n_params, _capi.u64ptr(xm), _capi.u64ptr(zm), _capi.dptr(coeff),
                                   len(xm), None, C.byref(energy), _capi.dptr(grad)))
        else:
            H = _dense_h(hamiltonian, self.n)
            check(lib.desv_adjoint(h.ptr, _capi.dptr(ops), len(ops), _capi.i64ptr(param_of_op),
                                   self.n_params, None, None, None, 0,
                                   H.view(np.float64).ctypes.data_as(_capi._pd), C.byref(energy),
                                   _capi.dptr(grad)))
        return energy.value, grad

    def energy_for_backward(self, theta, hamiltonian, initial_state=None) -> Tuple[float, int]:
        """Energy of a Pauli-sum Hamiltonian from an energies-only run that
        leaves the reverse sweep's inputs on the device; returns (energy,
        token).  :meth:`grad_for_backward` with the token then runs only the
        reverse sweep, if no other evaluation came in between."""
        self.ops_for(theta)
        b = self._prep_batch(theta, hamiltonian, initial_state)
        e = float(b.energies()[0])
        self._token = getattr(self, "_token", 0) + 1
        return e, self._token

    def grad_for_backward(self, token, theta, hamiltonian, initial_state=None) -> np.ndarray:
        """Gradient for the evaluation `token` named: the reverse sweep alone
        when that run is still the last one on the device, else a full
        energy + gradient run at `theta`."""
        if token == getattr(self, "_token", None) and getattr(self, "_bham", None) is hamiltonian:
            return self._batch.run_reverse()[0]
        self._token = getattr(self, "_token", 0) + 1
        return self.value_and_grad(theta, hamiltonian, initial_state)[1]

    def energy(self, theta, hamiltonian, initial_state=None, noise=None) -> float:
        """Forward only (no gradient sweep); `noise` (a NoiseSpec) applies one
        trajectory to the final state before the energy."""
        if isinstance(hamiltonian, PauliSumOperator) and noise is None:
            self.ops_for(theta)
            b = self._prep_batch(theta, hamiltonian, initial_state)
            return float(b.energies()[0])
        self._last_state = None
        ops = self.ops_for(theta)
        h = self.handle
        if initial_state is None:
            h.reset()
        else:
            h.upload(np.asarray(initial_state))
        h.run_ops(ops)
        if noise is not None:
            from .noise import NoiseModel
            NoiseModel.apply_to_sv(h, self.n, noise.model, noise.p, rng=noise.generator(),
                                   qubits=list(noise.qubits) if noise.qubits is not None else None)
        return _energy_of(h, hamiltonian, self.n)

    def close(self):
        self.handle.close()
        if getattr(self, "_batch", None) is not None:
            self._batch.close()
            self._batch = None


class _BatchState:
    """A batch instance's final state seen as a DeviceState handle (host
    transfers and zero-copy device views)."""

    def __init__(self, batch, n, dtype):
        self.batch, self.n, self.dtype = batch, n, np.dtype(dtype)
        self.device = batch.device

    @property
    def dim(self):
        return 1 << self.n

    def download(self):
        return self.batch.get_state(0, self.n)

    def device_ptr(self) -> int:
        return self.batch.state_device_ptr(0)


def _dense_h(hamiltonian, n):
    try:
        import torch
        if isinstance(hamiltonian, torch.Tensor):
            hamiltonian = hamiltonian.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    H = np.ascontiguousarray(np.asarray(hamiltonian), dtype=np.complex128)
    if H.shape != (1 << n, 1 << n):
        raise ValueError(f"h_matrix must be ({1 << n}, {1 << n}), got {H.shape}")
    return H


def _energy_of(h: StateHandle, hamiltonian, n: int) -> float:
    """Re <psi|H|psi> on the device (FP64)."""
    from .observables import expectations
    if isinstance(hamiltonian, PauliSumOperator):
        xm, zm, cs = hamiltonian.encoded()
        vals = expectations(h, xm, zm)
        return float(np.real(np.sum(cs * vals)))
    H = _dense_h(hamiltonian, n)
    tmp = StateHandle(n, h.code == _capi.C128, h.device)
    try:
        check(lib.desv_dense_matvec(tmp.ptr, h.ptr, H.view(np.float64).ctypes.data_as(_capi._pd)))
        re = C.c_double()
        check(lib.desv_inner(h.ptr, tmp.ptr, C.byref(re), None))
        return re.value
    finally:
        tmp.close()


def _make_autograd():
    import torch

    class AdjointEnergy(torch.autograd.Function):
        """E(theta) whose backward runs the adjoint reverse sweep.

        For a Pauli-sum Hamiltonian, forward runs the circuit and the energy
        only (the run keeps the state copy and lambda = H psi on the device)
        and backward runs just the reverse sweep over them, so an evaluation
        whose backward never runs costs no gradient work.  Dense Hamiltonians
        and noise tails take the one-call adjoint in forward."""

        @staticmethod
        def forward(ctx, theta, program, hamiltonian, initial_state, tail):
            th = theta.detach().cpu().numpy()
            ctx.program, ctx.hamiltonian, ctx.initial_state, ctx.theta = program, hamiltonian, initial_state, th
            ctx.dtype, ctx.device = theta.dtype, theta.device
            if isinstance(hamiltonian, PauliSumOperator) and (tail is None or len(tail) == 0):
                e, ctx.token = program.energy_for_backward(th, hamiltonian, initial_state)
                ctx.grad = None
            else:
                e, ctx.grad = program.value_and_grad(th, hamiltonian, initial_state, tail)
            return torch.as_tensor(e, dtype=theta.dtype, device=theta.device)

        @staticmethod
        def backward(ctx, grad_out):
            g = ctx.grad
            if g is None:
                g = ctx.program.grad_for_backward(ctx.token, ctx.theta, ctx.hamiltonian, ctx.initial_state)
            g = torch.as_tensor(g, dtype=ctx.dtype, device=ctx.device)
            return grad_out * g, None, None, None, None

    return AdjointEnergy


_AUTOGRAD = None


def adjoint_energy(theta, program: AdjointProgram, hamiltonian, initial_state=None, tail=None):
    """Torch entry point: differentiable energy of `program` at `theta`
    (`tail`: parameter-free op rows run after the circuit)."""
    global _AUTOGRAD
    if _AUTOGRAD is None:
        _AUTOGRAD = _make_autograd()
    return _AUTOGRAD.apply(theta, program, hamiltonian, initial_state, tail)


def circuit_to_energy_fn(circuit, n_qubits: int, dtype=None, device: int = 0) -> Tuple[Callable, int]:
    """(energy_fn, n_params) for a QASMCircuit.

    ``energy_fn(theta, h_matrix, stato_zero=None, noise=None)`` returns
    ``(energy, sv)``: a float (NumPy theta) or a differentiable torch scalar
    (torch theta), and a :class:`DeviceState` view of the final state (it
    aliases the function's device buffer and changes on the next call).
    ``energy_fn.value_and_grad(theta, h_matrix, stato_zero=None, noise=None)``
    returns ``(energy, grad)`` as NumPy values.  ``noise`` (a
    :class:`desv.NoiseSpec`) applies one trajectory of its channel to the
    final state before the energy (reference ``solvers/autodiff.py:160-164``);
    gradients are available for the Pauli channels (depolarizing, bitflip,
    phaseflip).
    """
    template = _build_template(circuit, n_qubits)
    n_params = sum(1 for op in circuit.ops if str(op['name']).lower().strip() in _PARAMETRIC_GATES)
    if dtype is None:
        complex128 = default_float64()
    else:
        complex128 = np.dtype(dtype) == np.complex128
    prog = AdjointProgram(template, n_qubits, complex128, device)
    if prog.n_params != n_params:
        # ops dropped for qubit >= n_qubits still count in n_params (reference
        # autodiff.py:133-134); theta entries beyond the template are unused.
        pass

    def _theta_for_prog(theta):
        return theta[:prog.n_params] if prog.n_params != n_params else theta

    def energy_fn(theta, h_matrix, stato_zero: Optional[np.ndarray] = None, noise=None):
        try:
            import torch
            is_torch = isinstance(theta, torch.Tensor)
        except ImportError:  # pragma: no cover
            is_torch = False
        if is_torch:
            th = _theta_for_prog(theta)
            if th.requires_grad or theta.requires_grad:
                e = adjoint_energy(th, prog, h_matrix, stato_zero, _noise_tail(noise, n_qubits))
            else:
                e = torch.as_tensor(prog.energy(th.detach().cpu().numpy(), h_matrix, stato_zero, noise),
                                    dtype=theta.dtype, device=theta.device)
            return e, prog.state()
        th = _theta_for_prog(np.asarray(theta, dtype=np.float64).reshape(-1))
        e = prog.energy(th, h_matrix, stato_zero, noise)
        return e, prog.state()

    def value_and_grad(theta, h_matrix, stato_zero=None, noise=None):
        th = np.asarray(theta, dtype=np.float64).reshape(-1)
        e, g = prog.value_and_grad(_theta_for_prog(th), h_matrix, stato_zero, _noise_tail(noise, n_qubits))
        full = np.zeros(len(th))
        full[:prog.n_params] = g
        return e, full

    def autotune(theta, h_matrix, **kw):
        """Time launch shapes of the adjoint evaluation for a Pauli-sum H."""
        th = np.asarray(theta, dtype=np.float64).reshape(-1)
        return prog.autotune(_theta_for_prog(th), h_matrix, **kw)

    energy_fn.value_and_grad = value_and_grad
    energy_fn.autotune = autotune
    energy_fn.program = prog
    return energy_fn, n_params


def _noise_tail(noise, n_qubits: int):
    """A Pauli-channel trajectory as op rows run after the circuit (the same
    draws as ``NoiseModel.apply_to_sv``); the gradient then follows from the
    adjoint sweep.  The trajectory's X/Y/Z are unitary, so the trajectory's
    renormalisation only divides by (1 + 1e-15).  Amplitude damping and
    'combined' rescale by state-dependent jump probabilities and are not
    differentiated here."""
    if noise is None or noise.model == 'ideal' or noise.p <= 0:
        return None
    from .noise import pauli_trajectory
    return pauli_trajectory(n_qubits, noise.model, float(noise.p), noise.generator(),
                            list(noise.qubits) if noise.qubits is not None else None)


def parameter_shift_grad(program: AdjointProgram, theta, hamiltonian) -> np.ndarray:
    """dE/dtheta_k = (E(theta + pi/2 e_k) - E(theta - pi/2 e_k)) / 2 on the plain forward path.

    The reference's rule (``backends/variational_numpy.py:149-156``); exact
    for rx/ry/rz.  Used as the baseline the adjoint sweep is measured against.
    """
    theta = np.asarray(theta, dtype=np.float64)
    g = np.zeros_like(theta)
    for k in range(len(theta)):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += np.pi / 2
        tm[k] -= np.pi / 2
        g[k] = 0.5 * (program.energy(tp, hamiltonian) - program.energy(tm, hamiltonian))
    return g
