"""Circuit generators used by the workloads: Pauli rotations, Trotter
products and entangling layers.

Same gate sequences as the reference (``circuits/trotter.py:52-149``,
``circuits/topology.py:17-79``): exp(-i a P) is a basis change (H for X,
Sdg+H for Y), a CNOT staircase onto the last qubit, rz(2a), and the mirror
image; a first-order Trotter step applies every term once in list order,
second order a forward half step then a reversed half step.
"""
from typing import Dict, List, Tuple

VALID_PATTERNS = ('linear', 'circular', 'full', 'star', 'brick')


def pauli_rotation_ops(pauli_dict: Dict[int, str], angle: float) -> List[Tuple]:
    qubits = sorted(pauli_dict)
    if not qubits:
        return []
    pre, post = [], []
    for q in qubits:
        letter = pauli_dict[q]
        if letter == 'X':
            pre.append(('h', q))
            post.append(('h', q))
        elif letter == 'Y':
            pre += [('sdg', q), ('h', q)]
            post += [('h', q), ('s', q)]
    ladder = [('cx', qubits[i], qubits[i + 1]) for i in range(len(qubits) - 1)]
    return pre + ladder + [('rz', qubits[-1], 2 * angle)] + ladder[::-1] + post


def trotter_evolve_ops(terms, t: float, n_steps: int, order: int = 1) -> List[Tuple]:
    if order not in (1, 2):
        raise ValueError(f"order must be 1 or 2, got {order}")
    dt = t / n_steps
    step: List[Tuple] = []
    if order == 1:
        for c, d in terms:
            step += pauli_rotation_ops(d, c * dt)
    else:
        for c, d in terms:
            step += pauli_rotation_ops(d, c * dt / 2)
        for c, d in reversed(terms):
            step += pauli_rotation_ops(d, c * dt / 2)
    return step * n_steps


def entangling_layer(n_qubits: int, pattern: str = 'linear', gate: str = 'cx', reverse: bool = False,
                     hub: int = 0) -> List[Tuple]:
    if n_qubits < 2:
        raise ValueError(f"entangling_layer needs at least 2 qubits, got {n_qubits}")
    if pattern not in VALID_PATTERNS:
        raise ValueError(f"unknown pattern {pattern!r}, expected one of {VALID_PATTERNS}")
    chain = [(q, q + 1) for q in range(n_qubits - 1)]
    if pattern == 'linear':
        edges = chain
    elif pattern == 'circular':
        edges = chain + ([(n_qubits - 1, 0)] if n_qubits > 2 else [])
    elif pattern == 'full':
        edges = [(i, j) for i in range(n_qubits) for j in range(i + 1, n_qubits)]
    elif pattern == 'star':
        if not 0 <= hub < n_qubits:
            raise ValueError(f"hub={hub} out of range for n_qubits={n_qubits}")
        edges = [(hub, q) for q in range(n_qubits) if q != hub]
    else:
        edges = chain[0::2] + chain[1::2]
    if reverse:
        edges = [(b, a) for a, b in edges]
    return [(gate, a, b) for a, b in edges]


_RC_1Q_FIXED = ('h', 'x', 'y', 'z', 's', 'sdg', 't', 'tdg', 'sx')
_RC_1Q_ANGLE = ('rx', 'ry', 'rz')
_RC_2Q = ('cx', 'cz', 'cy', 'swap')


def random_circuit(n_qubits: int, n_gates: int, seed=None, gate_set=None, two_qubit_prob: float = 0.4):
    """Random circuit of `n_gates` tuples for benchmarks and fuzzing
    (reference ``circuits/random_circuit.py:15-85``).

    Per gate, from one generator (a seed or an existing
    ``numpy.random.Generator``) and in this draw order: one uniform for the
    2-qubit choice (only when both pools are non-empty), then the gate index
    and its qubit(s) (2-qubit: two distinct qubits in one ``choice``), then
    the angle ~ U[0, 2 pi) of rx/ry/rz.  The same seed therefore gives the
    reference's circuit.  ``gate_set`` restricts the names (unknown names
    raise ValueError); 2-qubit gates are never drawn for one qubit.
    """
    if n_qubits < 1:
        raise ValueError(f"random_circuit needs at least 1 qubit, got {n_qubits}")
    if n_gates < 0:
        raise ValueError(f"n_gates must be >= 0, got {n_gates}")
    if not 0.0 <= two_qubit_prob <= 1.0:
        raise ValueError(f"two_qubit_prob must be in [0, 1], got {two_qubit_prob}")
    import numpy as np
    rng = seed if isinstance(seed, np.random.Generator) else np.random.default_rng(seed)
    known = set(_RC_1Q_FIXED) | set(_RC_1Q_ANGLE) | set(_RC_2Q)
    if gate_set is None:
        single = list(_RC_1Q_FIXED) + list(_RC_1Q_ANGLE)
        double = list(_RC_2Q) if n_qubits >= 2 else []
    else:
        names = list(gate_set)
        bad = sorted(set(names) - known)
        if bad:
            raise ValueError(f"unknown gate name(s) in gate_set: {bad}")
        single = [g for g in names if g in _RC_1Q_FIXED] + [g for g in names if g in _RC_1Q_ANGLE]
        double = [g for g in names if g in _RC_2Q]
    if not single and not double:
        raise ValueError("gate_set leaves no usable gates for this n_qubits")
    out: List[Tuple] = []
    for _ in range(n_gates):
        two = bool(double) and n_qubits >= 2 and (not single or rng.random() < two_qubit_prob)
        if two:
            name = double[rng.integers(len(double))]
            a, b = rng.choice(n_qubits, size=2, replace=False)
            out.append((name, int(a), int(b)))
            continue
        name = single[rng.integers(len(single))]
        q = int(rng.integers(n_qubits))
        out.append((name, q, float(rng.uniform(0, 2 * np.pi))) if name in _RC_1Q_ANGLE else (name, q))
    return out
