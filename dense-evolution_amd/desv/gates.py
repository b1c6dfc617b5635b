"""Gate tables of the drop-in surface.

Same names, matrices and ids as the reference's ``circuits/gates.py:33-91``
(``GATES`` fixed matrices in complex128, ``GATE_IDS`` for the op-table
encoding, ``PARAMETRIC_GATES`` as functions of the angle(s)).  Matrices are
built with NumPy in complex128; the engine rounds them to the state precision
when a program is uploaded.
"""
import numpy as np

INV2 = 1.0 / np.sqrt(2.0)


def _m(rows):
    return np.array(rows, dtype=np.complex128)


GATES = {
    'h': INV2 * _m([[1, 1], [1, -1]]),
    'x': _m([[0, 1], [1, 0]]),
    'y': _m([[0, -1j], [1j, 0]]),
    'z': _m([[1, 0], [0, -1]]),
    's': _m([[1, 0], [0, 1j]]),
    'sdg': _m([[1, 0], [0, -1j]]),
    't': _m([[1, 0], [0, np.exp(1j * np.pi / 4)]]),
    'tdg': _m([[1, 0], [0, np.exp(-1j * np.pi / 4)]]),
    'sx': 0.5 * _m([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]]),
    'id': np.eye(2, dtype=np.complex128),
    'cx': _m([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]]),
    'cz': _m([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, -1]]),
    'cy': _m([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, -1j], [0, 0, 1j, 0]]),
    'swap': _m([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]]),
    'iswap': _m([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]]),
    'ecr': INV2 * _m([[0, 0, 1, 1j], [0, 0, 1j, 1], [1, -1j, 0, 0], [-1j, 1, 0, 0]]),
    'ccx': np.eye(8, dtype=np.complex128)[[0, 1, 2, 3, 4, 5, 7, 6]],
}

#: op-table ids (reference compiler.py:24-48); 23 (swap) is accepted by the
#: engine but never produced by the transpiler.
GATE_IDS = {
    'id': 0, 'h': 1, 'x': 2, 'y': 3, 'z': 4, 's': 5, 'sdg': 6, 't': 7, 'tdg': 8,
    'rx': 9, 'ry': 10, 'rz': 11,
    'p': 12, 'u1': 12, 'phase': 12,
    'sx': 13,
    'gphase': 14,
    'cx': 20, 'cz': 21, 'cp': 22, 'cphase': 22,
    'cy': 24, 'crz': 25,
}

#: parametric gates acting on two qubits (cp/crz take (ctrl, tgt, angle)).
_TWO_QUBIT_PARAMETRIC_GATES = frozenset(('cp', 'cphase', 'crz'))


def _rx(theta):
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    return _m([[c, -1j * s], [-1j * s, c]])


def _ry(theta):
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    return _m([[c, -s], [s, c]])


def _rz(theta):
    return _m([[np.exp(-0.5j * theta), 0], [0, np.exp(0.5j * theta)]])


def _phase(lam):
    return _m([[1, 0], [0, np.exp(1j * lam)]])


def _u3(theta, phi, lam):
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    return _m([[c, -np.exp(1j * lam) * s], [np.exp(1j * phi) * s, np.exp(1j * (phi + lam)) * c]])


def _u2(phi, lam):
    return INV2 * _m([[1, -np.exp(1j * lam)], [np.exp(1j * phi), np.exp(1j * (phi + lam))]])


def _cp(lam):
    return np.diag([1, 1, 1, np.exp(1j * lam)]).astype(np.complex128)


def _crz(theta):
    return np.diag([1, 1, np.exp(-0.5j * theta), np.exp(0.5j * theta)]).astype(np.complex128)


#: the reference's key set exactly (no 'phase' / 'cphase' entries: those
#: names exist only in GATE_IDS, so its eager loop rejects them).
PARAMETRIC_GATES = {
    'rx': _rx, 'ry': _ry, 'rz': _rz,
    'cp': _cp, 'crz': _crz,
    'u3': _u3, 'u2': _u2, 'u1': _phase, 'p': _phase,
    'gphase': lambda alpha: np.exp(1j * alpha) * np.eye(2, dtype=np.complex128),
}
