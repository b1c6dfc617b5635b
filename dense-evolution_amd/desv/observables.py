"""Pauli-string observables computed on the GPU.

Same public functions and term formats as the reference's
``dense_evolution/physics/observables.py`` (string ``'XIZ'`` with
``pauli_terms[q]`` acting on qubit q, dict ``{q: 'X'}``, or ``(q, 'X')``
pairs; qubit q is basis-index bit n-1-q).  Every term is encoded as a pair of
bit masks (xmask: X or Y, zmask: Z or Y) and evaluated by the engine's Pauli
kernels with FP64 accumulation:

* ``pauli_expectation`` / ``pauli_sum_expectation``: one state sweep per
  distinct flip mask, all terms sharing it evaluated in the same sweep;
* ``pauli_sum_matvec`` / ``PauliSumOperator @ v``: H v without the matrix.

Host arrays are uploaded to a device state first; a :class:`DeviceState`
(``sim.sv``) is used in place.  ``pauli_hamiltonian_to_matrix`` and
``multiply_pauli_terms`` are host-side algebra (a dense matrix builder and
symbolic Pauli products), as in the reference.
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, List, Tuple

import numpy as np

from . import _capi
from ._capi import check, lib
from .simulator import DeviceState, StateHandle

__all__ = [
    'pauli_expectation', 'pauli_sum_expectation', 'pauli_hamiltonian_to_matrix',
    'pauli_sum_matvec', 'multiply_pauli_terms', 'pauli_sum_matvec_jax',
    'pauli_sum_expectation_jax', 'PauliSumOperator', 'encode_pauli_terms',
]

def _normalize_terms(pauli_terms, n_qubits=None):
    """str | dict | pairs -> {qubit: 'X'|'Y'|'Z'} with identities dropped."""
    if isinstance(pauli_terms, str):
        if n_qubits is not None and len(pauli_terms) != n_qubits:
            raise ValueError(f"pauli_terms string has length {len(pauli_terms)}, "
                             f"but n_qubits={n_qubits}")
        terms = {q: p.upper() for q, p in enumerate(pauli_terms) if p.upper() != 'I'}
    elif isinstance(pauli_terms, dict):
        terms = {int(q): str(p).upper() for q, p in pauli_terms.items() if str(p).upper() != 'I'}
    else:
        terms = {int(q): str(p).upper() for q, p in pauli_terms if str(p).upper() != 'I'}
    for q, p in terms.items():
        if p not in ('X', 'Y', 'Z'):
            raise ValueError(f"unknown Pauli operator {p!r} for qubit {q}, expected one of X, Y, Z, I")
        if q < 0:
            raise ValueError(f"qubit index {q} must be >= 0")
    return terms


def _masks(terms: dict, n: int) -> Tuple[int, int]:
    x = z = 0
    for q, p in terms.items():
        b = 1 << (n - 1 - q)
        if p in ('X', 'Y'):
            x |= b
        if p in ('Z', 'Y'):
            z |= b
    return x, z


def encode_pauli_terms(terms: Iterable, n: int, n_qubits=None, what='statevector'):
    """(coeff, pauli) pairs -> (xmask u64[T], zmask u64[T], coeff complex128[T])."""
    xs, zs, cs = [], [], []
    for coeff, p in terms:
        d = _normalize_terms(p, n_qubits)
        if d and max(d) >= n:
            raise ValueError(f"term references qubit {max(d)}, but {what} only spans {n} qubits")
        x, z = _masks(d, n)
        xs.append(x)
        zs.append(z)
        cs.append(coeff)
    return (np.array(xs, dtype=np.uint64), np.array(zs, dtype=np.uint64),
            np.array(cs, dtype=np.complex128))


def _as_device(statevector):
    """(handle, owned) for a DeviceState, simulator or host array."""
    if isinstance(statevector, DeviceState) and isinstance(statevector.handle, StateHandle):
        return statevector.handle, False
    h = getattr(statevector, 'handle', None)
    if isinstance(h, StateHandle):
        return h, False
    a = np.asarray(statevector)
    dim = a.shape[0]
    n = dim.bit_length() - 1
    if 1 << n != dim:
        raise ValueError(f"statevector length {dim} is not a power of 2")
    handle = StateHandle(n, a.dtype != np.complex64)
    handle.upload(a)
    return handle, True


def _state_qubits(statevector) -> int:
    dim = len(statevector) if not hasattr(statevector, 'shape') else statevector.shape[0]
    n = int(dim).bit_length() - 1
    if 1 << n != dim:
        raise ValueError(f"statevector length {dim} is not a power of 2")
    return n


def expectations(handle: StateHandle, xmask: np.ndarray, zmask: np.ndarray) -> np.ndarray:
    """<P_t> for every encoded term on a device state (FP64)."""
    out = np.zeros(len(xmask), dtype=np.float64)
    if len(xmask):
        xm = np.ascontiguousarray(xmask, dtype=np.uint64)
        zm = np.ascontiguousarray(zmask, dtype=np.uint64)
        check(lib.desv_pauli_expect(handle.ptr, _capi.u64ptr(xm), _capi.u64ptr(zm), len(xm),
                                    _capi.dptr(out)))
    return out


def pauli_expectation(statevector, pauli_terms, n_qubits=None) -> float:
    """<psi|P|psi> of one Pauli string (real), FP64 accumulation on the GPU."""
    n = _state_qubits(statevector)
    terms = _normalize_terms(pauli_terms, n_qubits)
    if terms and max(terms) >= n:
        raise ValueError(f"pauli_terms references qubit {max(terms)}, but the statevector "
                         f"only spans {n} qubits")
    h, owned = _as_device(statevector)
    try:
        if not terms:
            return float(h.norm2())
        x, z = _masks(terms, n)
        return float(expectations(h, np.array([x], np.uint64), np.array([z], np.uint64))[0])
    finally:
        if owned:
            h.close()


def pauli_sum_expectation(statevector, terms, n_qubits=None):
    """sum_i coeff_i <psi|P_i|psi>; every term in one kernel launch."""
    terms = list(terms)
    n = _state_qubits(statevector)
    xm, zm, cs = encode_pauli_terms(terms, n, n_qubits)
    h, owned = _as_device(statevector)
    try:
        if len(xm) and not np.any(cs.imag):
            # real coefficients: Re<psi|H|psi> in one tiled pass
            c = np.ascontiguousarray(cs.real, dtype=np.float64)
            out = np.zeros(1, dtype=np.float64)
            check(lib.desv_pauli_sum_expect(h.ptr, _capi.u64ptr(np.ascontiguousarray(xm)),
                                            _capi.u64ptr(np.ascontiguousarray(zm)), _capi.dptr(c), len(xm),
                                            _capi.dptr(out)))
            return float(out[0])
        vals = expectations(h, xm, zm)
    finally:
        if owned:
            h.close()
    total = 0.0
    for (coeff, _), v in zip(terms, vals):
        total += coeff * float(v)
    return total


def pauli_hamiltonian_to_matrix(terms, n_qubits):
    """Dense H = sum_i c_i P_i as a (2^n, 2^n) complex128 host matrix.

    Same matrix as the reference's Kronecker-product builder
    (``physics/observables.py:189-247``, qubit 0 = most significant index
    bit), filled column by column from the flip-mask rule instead:
    P|i> = i^(#Y) (-1)^popcount(i & zmask) |i ^ xmask>, one scatter per term.
    """
    if n_qubits < 1:
        raise ValueError(f"n_qubits must be >= 1, got {n_qubits}")
    dim = 1 << n_qubits
    H = np.zeros((dim, dim), dtype=np.complex128)
    cols = np.arange(dim, dtype=np.int64)
    for coeff, p in terms:
        d = _normalize_terms(p, n_qubits)
        if d and max(d) >= n_qubits:
            raise ValueError(f"term references qubit {max(d)}, but n_qubits={n_qubits}")
        x, z = _masks(d, n_qubits)
        ny = bin(x & z).count("1")
        par = np.zeros(dim, dtype=np.int64)
        zz = cols & z
        while np.any(zz):
            par ^= zz & 1
            zz >>= 1
        H[cols ^ x, cols] += (coeff * (1j ** ny)) * (1 - 2 * par)
    return H


def _matvec(vector, terms, n_qubits, out_dtype=np.complex128):
    n = _state_qubits(vector)
    xm, zm, cs = encode_pauli_terms(terms, n, n_qubits, what='vector')
    src, owned = _as_device(vector) if isinstance(vector, DeviceState) else (None, True)
    if src is None:
        src = StateHandle(n, True)
        src.upload(np.asarray(vector, dtype=np.complex128))
    dst = StateHandle(n, src.code == _capi.C128, src.device)
    try:
        matvec_into(dst, src, xm, zm, cs)
        return dst.download().astype(out_dtype)
    finally:
        dst.close()
        if owned:
            src.close()


def matvec_into(dst: StateHandle, src: StateHandle, xm, zm, cs) -> None:
    """dst <- H src on the device."""
    cre = np.ascontiguousarray(np.real(cs), dtype=np.float64)
    cim = np.ascontiguousarray(np.imag(cs), dtype=np.float64)
    xm = np.ascontiguousarray(xm, dtype=np.uint64)
    zm = np.ascontiguousarray(zm, dtype=np.uint64)
    check(lib.desv_pauli_matvec(dst.ptr, src.ptr, _capi.u64ptr(xm), _capi.u64ptr(zm),
                                _capi.dptr(cre), _capi.dptr(cim), len(xm)))


def pauli_sum_matvec(vector, terms, n_qubits=None):
    """H @ vector for a Pauli-sum H, complex128 result, matrix never built."""
    return _matvec(vector, list(terms), n_qubits)


def pauli_sum_matvec_jax(vector, terms, n_qubits=None):
    """Same as :func:`pauli_sum_matvec` (kept for name compatibility)."""
    return _matvec(vector, list(terms), n_qubits)


def pauli_sum_expectation_jax(statevector, terms, n_qubits=None):
    """Same as :func:`pauli_sum_expectation` (kept for name compatibility)."""
    return pauli_sum_expectation(statevector, terms, n_qubits)


class PauliSumOperator:
    """Matrix-free Hamiltonian: ``op @ v`` is the GPU Pauli-sum matvec.

    ``circuit_to_energy_fn``'s energy function recognises it and evaluates
    the energy and gradient with the Pauli kernels directly.
    """

    def __init__(self, terms, n_qubits):
        self.terms = list(terms)
        self.n_qubits = n_qubits
        self._enc = None

    def encoded(self):
        if self._enc is None:
            xm, zm, cs = encode_pauli_terms(self.terms, self.n_qubits, self.n_qubits)
            self._enc = (xm, zm, cs)
        return self._enc

    def __matmul__(self, vector):
        return pauli_sum_matvec(vector, self.terms, n_qubits=self.n_qubits)

    def as_linear_operator(self):
        """``scipy.sparse.linalg.LinearOperator`` whose matvec is the GPU
        kernel (host vectors in and out), the operator the reference hands to
        ``eigsh`` (``tools/dashboard/core/hamiltonians.py:569-576``).
        ``desv.lanczos.eigsh`` solves the same problem without host round
        trips."""
        import scipy.sparse.linalg as sla
        dim = 1 << self.n_qubits
        return sla.LinearOperator(shape=(dim, dim), matvec=lambda v: self @ np.ravel(v),
                                  dtype=np.complex128)


def _letter_bits(letter):
    return {'X': (1, 0), 'Y': (1, 1), 'Z': (0, 1)}[letter]


def multiply_pauli_terms(factors):
    """Operator product of (coeff, pauli) factors, left to right, as
    (coeff, {qubit: letter}) with identities dropped (the reference's
    ``physics/observables.py:461-523``).

    Letters are multiplied in the symplectic form sigma(x, z) =
    i^(x z) X^x Z^z: sigma(x1, z1) sigma(x2, z2) = i^(x1 z1 + x2 z2 - x3 z3 + 2 z1 x2)
    sigma(x3, z3) with (x3, z3) = (x1 ^ x2, z1 ^ z2), so X Y = i Z, Y X = -i Z
    and P P = I.
    """
    merged = {}
    power = 0          # accumulated phase i^power
    total = 1.0 + 0j
    for coeff, p in factors:
        total *= coeff
        for q, letter in _normalize_terms(p).items():
            x2, z2 = _letter_bits(letter)
            if q not in merged:
                merged[q] = (x2, z2)
                continue
            x1, z1 = merged[q]
            x3, z3 = x1 ^ x2, z1 ^ z2
            power += x1 * z1 + x2 * z2 - x3 * z3 + 2 * z1 * x2
            if x3 or z3:
                merged[q] = (x3, z3)
            else:
                del merged[q]
    total *= 1j ** (power % 4)
    letters = {(1, 0): 'X', (1, 1): 'Y', (0, 1): 'Z'}
    return total, {q: letters[b] for q, b in merged.items()}
