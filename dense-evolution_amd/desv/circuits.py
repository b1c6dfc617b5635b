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
