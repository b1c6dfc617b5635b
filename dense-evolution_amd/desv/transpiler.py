"""QuantumTranspiler: exact decompositions of non-native gates.

Same decompositions and pass-through rules as the reference
(``circuits/compiler.py:362-488``): ccx -> 15 gates, swap -> 3 cx,
iswap -> {s, h, cx}, ecr -> {s, sx, cx, x, gphase(-pi/4)},
u3/u2 -> {rz, ry, rz, gphase}.  The gate sequences are fixed by the
mathematics (they must reproduce the literal matrices exactly), so they match
the reference's term for term.
"""
from typing import List, Tuple

import numpy as np


class QuantumTranspiler:
    """Decomposes non-native gates; everything else passes through."""

    @staticmethod
    def decompose_toffoli(c1: int, c2: int, t: int) -> List[Tuple]:
        """CCX as 6 CX + 9 single-qubit gates (T/Tdg/H construction)."""
        return [('h', t), ('cx', c2, t), ('tdg', t), ('cx', c1, t), ('t', t),
                ('cx', c2, t), ('tdg', t), ('cx', c1, t), ('t', c2), ('t', t),
                ('cx', c1, c2), ('h', t), ('t', c1), ('tdg', c2), ('cx', c1, c2)]

    @staticmethod
    def decompose_swap(q1: int, q2: int) -> List[Tuple]:
        return [('cx', q1, q2), ('cx', q2, q1), ('cx', q1, q2)]

    @staticmethod
    def decompose_iswap(q1: int, q2: int) -> List[Tuple]:
        return [('s', q1), ('h', q1), ('s', q2), ('cx', q1, q2), ('cx', q2, q1), ('h', q2)]

    @staticmethod
    def decompose_ecr(q1: int, q2: int) -> List[Tuple]:
        return [('s', q1), ('sx', q2), ('cx', q1, q2), ('x', q1), ('gphase', q1, -np.pi / 4.0)]

    @staticmethod
    def decompose_u3(q, theta, phi, lam) -> List[Tuple]:
        """U3 = e^{i(phi+lam)/2} Rz(phi) Ry(theta) Rz(lam), applied right to left."""
        return [('rz', q, lam), ('ry', q, theta), ('rz', q, phi), ('gphase', q, (phi + lam) / 2.0)]

    @staticmethod
    def decompose_u2(q, phi, lam) -> List[Tuple]:
        return QuantumTranspiler.decompose_u3(q, np.pi / 2.0, phi, lam)

    @classmethod
    def transpile(cls, circuit) -> List[Tuple]:
        out: List[Tuple] = []
        for cmd in circuit:
            name = cmd[0].lower()
            if name == 'ccx':
                out.extend(cls.decompose_toffoli(*cmd[1:4]))
            elif name == 'swap':
                out.extend(cls.decompose_swap(*cmd[1:3]))
            elif name == 'iswap':
                out.extend(cls.decompose_iswap(*cmd[1:3]))
            elif name == 'ecr':
                out.extend(cls.decompose_ecr(*cmd[1:3]))
            elif name == 'u3' and len(cmd) >= 5:
                # structural (name, q) tuples pass through (reference :482-485)
                out.extend(cls.decompose_u3(*cmd[1:5]))
            elif name == 'u2' and len(cmd) >= 4:
                out.extend(cls.decompose_u2(*cmd[1:4]))
            else:
                out.append(cmd)
        return out
