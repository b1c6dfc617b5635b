"""Stochastic single-qubit Kraus channels on a device-resident state.

Drop-in for the reference's ``NoiseModel`` (``noise/kraus_channels.py:60-285``
with the channels of ``noise/kraus/*.py``) and ``NoiseSpec``
(``noise/differentiable.py:8-66``).  One trajectory = per target qubit one
fire/no-fire decision drawn from the caller's NumPy generator, in the
reference's order (depolarizing: ``r`` then ``ch``; amplitude damping: ``r``
against the whole-state jump probability, from populations the GPU returns),
the chosen 2x2 operators are applied to the whole state on the GPU (one pass
per group of up to 10 qubits for damping, one pass in all for the Pauli
channels), and the state is divided by (norm + 1e-15).  Noise qubits are indexed like the reference: qubit q is
basis-index bit q (LSB-first, ``kraus_channels.py:38-58``), i.e. engine qubit
n-1-q.

Given the same ``numpy.random.Generator``, the decisions and the result equal
the reference's NumPy path.  The reference's JAX path draws from a JAX key
instead; JAX is not part of this engine, so a ``NoiseSpec``'s ``jax_key`` is
used only to seed a NumPy generator (reproducible per key, not the same
stream as JAX).
"""
from __future__ import annotations

import os
import time
from typing import Dict, List, Optional

import numpy as np

from .simulator import DeviceState, StateHandle

__all__ = ["NoiseModel", "NoiseSpec"]

_THIRD = 1.0 / 3.0
_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)


def _fresh_rng() -> np.random.Generator:
    """OS entropy mixed with a nanosecond counter (``kraus_channels.py:25-35``)."""
    seed = (int.from_bytes(os.urandom(8), 'big') ^ (time.perf_counter_ns() & 0xFFFF_FFFF_FFFF_FFFF))
    return np.random.default_rng(seed & 0xFFFF_FFFF_FFFF_FFFF)


def _rng_from_key(key) -> np.random.Generator:
    return np.random.default_rng(np.asarray(key, dtype=np.uint64).ravel().tolist())


class NoiseSpec:
    """Noise configuration for ``energy_fn(..., noise=spec)``.

    ``model`` and ``qubits`` as in the reference; ``rng`` (a NumPy
    Generator) supplies the decisions; without it ``jax_key`` (any integer
    array) seeds one, and without either the decisions use fresh OS entropy.
    """

    def __init__(self, model: str, p, jax_key=None, qubits: Optional[List[int]] = None,
                 rng: Optional[np.random.Generator] = None):
        self.model = model
        self.p = p
        self.jax_key = jax_key
        self.qubits = tuple(qubits) if qubits is not None else None
        self.rng = rng

    def __repr__(self):
        return f"NoiseSpec(model={self.model!r}, p={self.p!r}, qubits={self.qubits!r})"

    def generator(self) -> np.random.Generator:
        if self.rng is not None:
            return self.rng
        if self.jax_key is not None:
            return _rng_from_key(self.jax_key)
        return _fresh_rng()


class NoiseModel:
    """Stochastic single-qubit Kraus channels (``kraus_channels.py:60-101``)."""

    MODELS = ['ideal', 'depolarizing', 'bitflip', 'phaseflip', 'amplitude_damping', 'combined']

    @staticmethod
    def apply_to_sv(sv, n: int, model: str, p: float, rng: Optional[np.random.Generator] = None,
                    qubits: Optional[List[int]] = None, jax_key=None):
        """One trajectory of `model` with strength `p` on `sv`.

        A host array is copied to the GPU, processed and returned as a new
        normalised host array (the caller's array is not modified); a
        :class:`DeviceState` (e.g. ``sim.sv``) or :class:`StateHandle` is
        updated in place and returned.  ``'ideal'`` or ``p <= 0`` returns
        `sv` unchanged.
        """
        if model not in NoiseModel.MODELS:
            raise ValueError(f"unknown noise model {model!r}, expected one of {NoiseModel.MODELS}")
        if model == 'ideal' or p <= 0.0:
            return sv
        if rng is None:
            rng = _rng_from_key(jax_key) if jax_key is not None else _fresh_rng()
        if isinstance(sv, DeviceState):
            handle, owned = sv.handle, False
        elif isinstance(sv, StateHandle):
            handle, owned = sv, False
        else:
            a = np.asarray(sv)
            dim = a.shape[0]
            if (1 << n) != dim:
                raise ValueError(f"statevector length {dim} != 2**{n}")
            handle = StateHandle(n, a.dtype != np.complex64)
            handle.upload(a)
            owned = True
        try:
            if handle.n != n:
                raise ValueError(f"statevector has {handle.n} qubits, n = {n}")
            qs = list(qubits if qubits is not None else range(n))
            for q in qs:
                if not 0 <= q < n:
                    raise ValueError(f"noise qubit {q} out of range for {n} qubits")
            norm2 = _run_trajectory(handle, [n - 1 - q for q in qs], model, float(p), rng)
            handle.scale(1.0 / (np.sqrt(norm2) + 1e-15))
            if owned:
                return handle.download()
            return sv
        finally:
            if owned:
                handle.close()

    @staticmethod
    def kraus_description(model: str) -> Dict:
        """Kraus operators and physical meaning (``kraus_channels.py:237-285``)."""
        desc = {
            'ideal': {'kraus': 1, 'formula': 'K₀ = I', 'physical': 'No noise'},
            'depolarizing': {'kraus': 4, 'formula': 'K₀=√(1-p)I  K₁=√(p/3)X  K₂=√(p/3)Y  K₃=√(p/3)Z',
                             'physical': 'Isotropic Pauli error — equiprobable X, Y, Z'},
            'bitflip': {'kraus': 2, 'formula': 'K₀=√(1-p)I  K₁=√p·X',
                        'physical': 'Bit flip σ_x with probability p'},
            'phaseflip': {'kraus': 2, 'formula': 'K₀=√(1-p)I  K₁=√p·Z',
                          'physical': 'Pure dephasing σ_z with probability p'},
            'amplitude_damping': {'kraus': 2, 'formula': 'K₀=diag(1,√(1-γ))  K₁=[[0,√γ],[0,0]]',
                                  'physical': 'T₁ energy relaxation |1⟩→|0⟩ with rate γ'},
            'combined': {'kraus': 6, 'formula': 'Depolarizing(p/2) ∘ AmplitudeDamping(p/3)',
                         'physical': 'Worst-case NISQ: dephasing + relaxation'},
        }
        return desc.get(model, desc['ideal'])


# ---- channels: one decision per qubit, one device pass per group ------------

# Distinct qubits per damping layer: the device returns 2^_GROUP populations.
_GROUP = 10


class _Layer:
    """The operators drawn for one group of qubits, composed per qubit in
    draw order, and the FP64 populations over the group's qubits that each
    draw sees.  Every operator of these channels has at most one nonzero entry
    per row, so the populations after it are |m_rc|^2-weighted sums of the
    populations before it (no interference): the host updates them draw by
    draw, exactly as a pass over the state would, and the whole group is
    applied in one device pass (``StateHandle.apply_layer``)."""

    def __init__(self, qubits, pops=None):
        self.pos = {q: j for j, q in enumerate(qubits)}
        self.h = None if pops is None else np.array(pops, dtype=np.float64)
        self.ops = {}

    def _axis(self, q):
        return self.h.reshape(-1, 2, 1 << self.pos[q])

    def probs(self, q):
        h = self._axis(q)
        return float(h[:, 0, :].sum()), float(h[:, 1, :].sum())

    def push(self, m, q):
        m = np.asarray(m, dtype=np.complex128)
        self.ops[q] = m @ self.ops[q] if q in self.ops else m
        if self.h is not None:
            h = self._axis(q)
            w = np.abs(m) ** 2
            h0, h1 = h[:, 0, :].copy(), h[:, 1, :].copy()
            h[:, 0, :] = w[0, 0] * h0 + w[0, 1] * h1
            h[:, 1, :] = w[1, 0] * h0 + w[1, 1] * h1

    # the channels' two kinds of step
    def pauli(self, m, q, swaps=True):
        self.push(m, q)

    def kraus(self, m, q):
        self.push(m, q)


def _distinct(qs):
    return list(dict.fromkeys(qs))


def _run_trajectory(handle: StateHandle, eqs, model: str, p: float, rng) -> float:
    """Draw and apply one trajectory over engine qubits `eqs` (in the
    caller's order); returns the squared norm of the result.  Pauli channels
    draw independently of the state: one layer over all qubits.  Damping
    draws against populations: groups of up to _GROUP distinct qubits, each
    one device pass that also returns the next group's populations."""
    chan = _CHANNELS[model]
    if model in ('depolarizing', 'bitflip', 'phaseflip') or not eqs:
        layer = _Layer(_distinct(eqs))
        for q in eqs:
            chan(layer, q, p, rng)
        return float(handle.apply_layer(layer.ops, [])[0])
    groups, cur = [], []
    for q in eqs:
        if q not in cur and len(_distinct(cur)) == _GROUP:
            groups.append(cur)
            cur = []
        cur.append(q)
    groups.append(cur)
    pops = handle.apply_layer({}, _distinct(groups[0]))
    for gi, g in enumerate(groups):
        layer = _Layer(_distinct(g), pops)
        for q in g:
            chan(layer, q, p, rng)
        nxt = _distinct(groups[gi + 1]) if gi + 1 < len(groups) else []
        pops = handle.apply_layer(layer.ops, nxt)
    return float(pops[0])


def _depolarizing(st: _Layer, q: int, p: float, rng):
    r, ch = rng.random(), rng.random()
    if r < p:
        if ch < _THIRD:
            st.pauli(_X, q, True)
        elif ch < 2.0 * _THIRD:
            st.pauli(_Y, q, True)
        else:
            st.pauli(_Z, q, False)


def _bitflip(st, q, p, rng):
    if rng.random() < p:
        st.pauli(_X, q, True)


def _phaseflip(st, q, p, rng):
    if rng.random() < p:
        st.pauli(_Z, q, False)


def _amplitude_damping(st, q, p, rng):
    gamma = float(np.clip(p, 0.0, 1.0))
    r = rng.random()
    p0, p1 = st.probs(q)
    p1 = float(np.clip(gamma * p1, 0.0, 1.0))
    if r < p1:
        m = np.array([[0, np.sqrt(gamma) / np.sqrt(max(p1, 1e-15))], [0, 0]], dtype=np.complex128)
    else:
        nn = np.sqrt(max(1.0 - p1, 1e-15))
        m = np.array([[1.0 / nn, 0], [0, np.sqrt(1.0 - gamma) / nn]], dtype=np.complex128)
    st.kraus(m, q)


def _combined(st, q, p, rng):
    _depolarizing(st, q, p * 0.5, rng)
    _amplitude_damping(st, q, p * 0.333333, rng)


_CHANNELS = {'depolarizing': _depolarizing, 'bitflip': _bitflip, 'phaseflip': _phaseflip,
             'amplitude_damping': _amplitude_damping, 'combined': _combined}


def pauli_trajectory(n: int, model: str, p: float, rng, qubits=None):
    """For the Pauli channels (depolarizing, bitflip, phaseflip): draw one
    trajectory's decisions exactly as :meth:`NoiseModel.apply_to_sv` does and
    return them as op-table rows (x = 2, y = 3, z = 4 on engine qubits), so
    a noisy circuit is the circuit followed by these gates (unitary: the norm
    is unchanged and gradients follow by the adjoint sweep)."""
    if model not in NoiseModel.MODELS:
        raise ValueError(f"unknown noise model {model!r}, expected one of {NoiseModel.MODELS}")
    if model not in ('depolarizing', 'bitflip', 'phaseflip'):
        raise NotImplementedError(f"gradients through {model!r} trajectories are not supported "
                                  f"(the jump normalisation depends on the state)")
    rows = []
    for q in (qubits if qubits is not None else range(n)):
        if not 0 <= q < n:
            raise ValueError(f"noise qubit {q} out of range for {n} qubits")
        eq = n - 1 - q
        if model == 'depolarizing':
            r, ch = rng.random(), rng.random()
            if r < p:
                rows.append([2 if ch < _THIRD else (3 if ch < 2.0 * _THIRD else 4), eq, 0, 0.0])
        elif rng.random() < p:
            rows.append([2 if model == 'bitflip' else 4, eq, 0, 0.0])
    return np.array(rows, dtype=np.float64).reshape(-1, 4)
