"""Seeded synthetic workloads (the benchmark configurations).

All inputs come from ``numpy.random.default_rng`` with the draw order written
in each docstring, so the GPU engine and the CPU oracle see identical
circuits.  Nothing here touches the device.
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np

from .circuits import trotter_evolve_ops
from .qasm import QASMCircuit


def random_layers_circuit(n: int = 20, layers: int = 20, seed: int = 0) -> List[Tuple]:
    """Plain-method circuit: per layer draw U[0, 2pi)^(n, 3), apply rx, ry, rz
    to every qubit (qubit order, rx->ry->rz), then cx(i, i+1) for i < n-1."""
    rng = np.random.default_rng(seed)
    ops: List[Tuple] = []
    for _ in range(layers):
        a = rng.uniform(0.0, 2.0 * np.pi, (n, 3))
        for q in range(n):
            ops += [('rx', q, float(a[q, 0])), ('ry', q, float(a[q, 1])), ('rz', q, float(a[q, 2]))]
        ops += [('cx', i, i + 1) for i in range(n - 1)]
    return ops


def haar_su4(rng: np.random.Generator) -> np.ndarray:
    """Haar-random SU(4): QR of a complex Gaussian, R-phase fixed, det = 1.

    Draw order: the real 4x4 block, then the imaginary 4x4 block."""
    z = (rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))) / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    q = q * (d / np.abs(d))
    det = np.linalg.det(q)
    return q / det ** 0.25


def brickwork_su4(n: int = 30, depth: int = 24, seed: int = 0) -> List[Tuple[np.ndarray, int, int]]:
    """Brickwork of Haar SU(4) gates: even layers pair (0,1),(2,3),..., odd
    layers (1,2),(3,4),...; each entry is (U, q_low, q_low + 1) with q_low the
    most significant index bit of U (apply_gate_2q convention)."""
    rng = np.random.default_rng(seed)
    out = []
    for layer in range(depth):
        for q in range(layer % 2, n - 1, 2):
            out.append((haar_su4(rng), q, q + 1))
    return out


def hea_circuit(n: int, layers: int) -> QASMCircuit:
    """Hardware-efficient ansatz: per layer ry(theta) and rz(theta) on every
    qubit (all ry first, then all rz), then cz(i, i+1) for i < n-1."""
    ops = []
    for _ in range(layers):
        ops += [{'type': 'gate', 'name': 'ry', 'qubits': [q], 'params': [0.0]} for q in range(n)]
        ops += [{'type': 'gate', 'name': 'rz', 'qubits': [q], 'params': [0.0]} for q in range(n)]
        ops += [{'type': 'gate', 'name': 'cz', 'qubits': [i, i + 1], 'params': []} for i in range(n - 1)]
    return QASMCircuit(n, 0, ops)


def random_pauli_hamiltonian(rng: np.random.Generator, n: int, n_terms: int, max_weight: int = 4):
    """Terms (c, {q: letter}): weight w ~ U{1..max_weight}, distinct qubits
    ``rng.choice(n, w, replace=False)``, letters U{X,Y,Z}, c ~ N(0, 1); per
    term the draw order is w, qubits, letters, c."""
    terms = []
    for _ in range(n_terms):
        w = int(rng.integers(1, max_weight + 1))
        qs = rng.choice(n, size=min(w, n), replace=False)
        letters = rng.integers(0, 3, size=len(qs))
        c = float(rng.normal())
        terms.append((c, {int(q): 'XYZ'[int(l)] for q, l in zip(qs, letters)}))
    return terms


def vqe_instance(n: int = 20, layers: int = 8, n_terms: int = 200, seed=0):
    """(circuit, theta, terms): theta ~ U[-pi, pi) in gate order is drawn
    first, then the Hamiltonian."""
    rng = np.random.default_rng(seed)
    circ = hea_circuit(n, layers)
    theta = rng.uniform(-np.pi, np.pi, 2 * n * layers)
    terms = random_pauli_hamiltonian(rng, n, n_terms)
    return circ, theta, terms


def batch_instances(count: int = 1024, n_lo: int = 10, n_hi: int = 22, l_lo: int = 4, l_hi: int = 8,
                    n_terms: int = 50):
    """Heterogeneous VQE batch: instance i uses default_rng([0, i]) and draws
    n ~ U{n_lo..n_hi}, L ~ U{l_lo..l_hi}, then theta, then its Hamiltonian."""
    out = []
    for i in range(count):
        rng = np.random.default_rng([0, i])
        n = int(rng.integers(n_lo, n_hi + 1))
        L = int(rng.integers(l_lo, l_hi + 1))
        circ = hea_circuit(n, L)
        theta = rng.uniform(-np.pi, np.pi, 2 * n * L)
        terms = random_pauli_hamiltonian(rng, n, n_terms)
        out.append((n, circ, theta, terms))
    return out


def xxz_terms(n: int, J: float = 1.0, delta: float = 0.5):
    """Open XXZ chain: per bond (J, XX), (J, YY), (J*delta, ZZ) in bond order."""
    terms = []
    for i in range(n - 1):
        terms += [(J, {i: 'X', i + 1: 'X'}), (J, {i: 'Y', i + 1: 'Y'}), (J * delta, {i: 'Z', i + 1: 'Z'})]
    return terms


def tfim_terms(n: int, seed: int = 0):
    """Open transverse-field Ising chain H = -sum J_i Z_i Z_{i+1} - sum h_i X_i
    with J_i ~ U[0.5, 1.5), h_i ~ U[0.2, 1.2) (J drawn first)."""
    rng = np.random.default_rng(seed)
    J = rng.uniform(0.5, 1.5, n - 1)
    h = rng.uniform(0.2, 1.2, n)
    return ([(-float(J[i]), {i: 'Z', i + 1: 'Z'}) for i in range(n - 1)]
            + [(-float(h[i]), {i: 'X'}) for i in range(n)]), J, h


def xxz_trotter_circuit(n: int = 28, dt: float = 0.05, steps: int = 200, seed: int = 0):
    """Random product state (ry(theta_q) then rz(phi_q), theta ~ U[0, pi),
    phi ~ U[0, 2pi), drawn as the theta vector then the phi vector) followed
    by first-order Trotter evolution of the XXZ chain for `steps` steps."""
    rng = np.random.default_rng(seed)
    th = rng.uniform(0.0, np.pi, n)
    ph = rng.uniform(0.0, 2.0 * np.pi, n)
    prep = []
    for q in range(n):
        prep += [('ry', q, float(th[q])), ('rz', q, float(ph[q]))]
    return prep + trotter_evolve_ops(xxz_terms(n), dt * steps, steps, order=1)
