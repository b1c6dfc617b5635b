"""Matrix-free Lanczos ground states with the Krylov basis resident on the GPU.

The reference finds ground states of Pauli-sum Hamiltonians with
``scipy.sparse.linalg.eigsh(LinearOperator(matvec=pauli_sum_matvec), k=1,
which='SA', ncv=min(dim - 1, max(2k + 1, 20)))``
(``tools/dashboard/core/hamiltonians.py:512-577``): ARPACK's implicitly
restarted Lanczos, with every product a host -> device -> host round trip.
Here the same problem is solved by a thick-restart (Krylov-Schur) Lanczos
whose basis vectors never leave the device:

* products ``H v_j`` are the Pauli-sum matvec kernel;
* each new vector is orthogonalised against the whole basis twice (classical
  Gram-Schmidt with one re-orthogonalisation) by two kernels: all inner
  products ``<v_i|w>`` in one pass (FP64), then ``w -= sum c_i v_i`` in one
  pass;
* at a restart the basis is rotated in place onto the ``p`` best Ritz vectors
  (one pass reading ``m`` vectors and writing ``p``) and the residual vector
  becomes vector ``p``;
* only the small projected matrix (at most ncv x ncv) is diagonalised on the
  host (``numpy.linalg.eigh``).

Convergence: Ritz pair i is accepted when ``|beta_m s_{m,i}| <= tol *
max(|theta_i|, 1)`` (the residual norm of the pair).  ``tol = 0`` means the
state precision's floor (1e-13 for complex128, 1e-5 for complex64).  As with
ARPACK without shift-invert, interior targets ('SM') converge slowly.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional

import numpy as np

from ._capi import check, lib
from .observables import PauliSumOperator, matvec_into
from .simulator import DeviceState, StateHandle

__all__ = ["eigsh", "ground_state_energy"]

_WHICH = ("SA", "LA", "SM", "LM")


def _arr(handles: List[StateHandle]):
    return (C.c_void_p * len(handles))(*[h.ptr.value for h in handles])


def inner_many(w: StateHandle, vs: List[StateHandle]) -> np.ndarray:
    """[<v_j|w>] (complex, FP64 accumulation) in one pass over w."""
    out = np.empty(2 * len(vs), dtype=np.float64)
    check(lib.desv_inner_many(w.ptr, _arr(vs), len(vs), out.ctypes.data_as(C.POINTER(C.c_double))))
    return out[0::2] + 1j * out[1::2]


def combine(outs: List[StateHandle], ins: List[StateHandle], coef: np.ndarray) -> None:
    """outs[l] <- sum_j coef[j, l] ins[j] element by element (outs may be among ins)."""
    c = np.ascontiguousarray(np.asarray(coef, dtype=np.complex128).reshape(len(ins), len(outs)))
    check(lib.desv_combine(_arr(outs), len(outs), _arr(ins), len(ins),
                           c.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))))


def _order(theta: np.ndarray, which: str) -> np.ndarray:
    if which == "SA":
        return np.argsort(theta, kind="stable")
    if which == "LA":
        return np.argsort(-theta, kind="stable")
    if which == "SM":
        return np.argsort(np.abs(theta), kind="stable")
    return np.argsort(-np.abs(theta), kind="stable")


class _Operator:
    def __init__(self, hamiltonian, n_qubits):
        if isinstance(hamiltonian, PauliSumOperator):
            op = hamiltonian
        else:
            if n_qubits is None:
                raise ValueError("n_qubits is required with a list of Pauli terms")
            op = PauliSumOperator(list(hamiltonian), n_qubits)
        self.n = int(op.n_qubits)
        self.xm, self.zm, self.cs = op.encoded()

    def __call__(self, dst: StateHandle, src: StateHandle) -> None:
        matvec_into(dst, src, self.xm, self.zm, self.cs)


def _orthonormalise(w: StateHandle, basis: List[StateHandle]) -> (np.ndarray, float):
    """Two passes of classical Gram-Schmidt against `basis`; returns the
    summed coefficients and the norm of what is left (w is not scaled)."""
    total = np.zeros(len(basis), dtype=np.complex128)
    for _ in range(2):
        for s in range(0, len(basis), 63):
            chunk = basis[s:s + 63]
            c = inner_many(w, chunk)
            combine([w], [w] + chunk, np.concatenate([[1.0], -c]))
            total[s:s + len(chunk)] += c
    return total, float(np.sqrt(w.norm2()))


def eigsh(hamiltonian, k: int = 1, which: str = "SA", ncv: Optional[int] = None, tol: float = 0.0,
          maxiter: Optional[int] = None, v0=None, return_eigenvectors: bool = False,
          n_qubits: Optional[int] = None, dtype: str = "complex128", device: int = 0, seed: int = 0,
          return_states: bool = False, info: Optional[dict] = None):
    """k eigenvalues (and vectors) of a Pauli-sum Hamiltonian, ``scipy.sparse.
    linalg.eigsh``-style: ``which`` in SA/LA/SM/LM, ``ncv`` basis vectors
    (default ``min(dim, max(2k + 1, 20))``), eigenvalues returned in ascending
    order.  ``hamiltonian`` is a :class:`PauliSumOperator` or a list of
    ``(coeff, term)`` pairs with ``n_qubits``.  ``v0`` (host array or device
    state) is the start vector, else a seeded pseudo-random one made on the
    device.  Eigenvectors come back as a host ``(dim, k)`` array, or as
    :class:`StateHandle` objects with ``return_states=True``.  A dict passed
    as ``info`` receives the counts of products and restarts and the final
    residual norms.

    Device memory: ``ncv + 1`` state vectors.
    """
    if which not in _WHICH:
        raise ValueError(f"which must be one of {_WHICH}, got {which!r}")
    op = _Operator(hamiltonian, n_qubits)
    n = op.n
    dim = 1 << n
    if not 1 <= k < dim:
        raise ValueError(f"k must satisfy 1 <= k < dim = {dim}, got {k}")
    m = ncv if ncv is not None else min(dim, max(2 * k + 1, 20))
    if not k < m <= min(dim, 64):
        raise ValueError(f"ncv must satisfy k < ncv <= min(dim, 64), got ncv={m}, k={k}, dim={dim}")
    c128 = np.dtype(dtype) == np.complex128
    floor = 1e-13 if c128 else 1e-5
    tol_eff = max(float(tol), floor)
    eps = np.finfo(np.float64 if c128 else np.float32).eps
    maxiter = maxiter if maxiter is not None else max(300, 10 * dim // m)

    V = [StateHandle(n, c128, device) for _ in range(m + 1)]
    rng_seed = int(seed)
    try:
        # start vector
        if v0 is None:
            check(lib.desv_fill_random(V[0].ptr, C.c_uint64(rng_seed & 0xFFFF_FFFF_FFFF_FFFF)))
        elif isinstance(v0, (StateHandle, DeviceState)):
            V[0].copy_from(v0 if isinstance(v0, StateHandle) else v0.handle)
        else:
            V[0].upload(np.asarray(v0))
        nrm = np.sqrt(V[0].norm2())
        if not nrm > 0:
            raise ValueError("start vector is zero")
        V[0].scale(1.0 / nrm)

        T = np.zeros((m + 1, m), dtype=np.complex128)
        start, restarts, m_eff, beta, products = 0, 0, m, 0.0, 0
        while True:
            m_eff = m
            for j in range(start, m):
                op(V[j + 1], V[j])
                products += 1
                c, beta = _orthonormalise(V[j + 1], V[:j + 1])
                T[:j + 1, j] = c
                scale = max(1.0, float(np.max(np.abs(T[:j + 2, :j + 1]))))
                if beta <= 100 * eps * scale:
                    # invariant subspace: exact Ritz pairs so far
                    T[j + 1, j] = 0.0
                    beta = 0.0
                    if j + 1 >= k or j + 1 == dim:
                        m_eff = j + 1
                        break
                    # too small to hold k pairs: continue from a fresh direction
                    rng_seed += 0x9E3779B97F4A7C15
                    check(lib.desv_fill_random(V[j + 1].ptr, C.c_uint64(rng_seed & 0xFFFF_FFFF_FFFF_FFFF)))
                    _, b2 = _orthonormalise(V[j + 1], V[:j + 1])
                    V[j + 1].scale(1.0 / b2)
                    continue
                T[j + 1, j] = beta
                V[j + 1].scale(1.0 / beta)
            Tm = T[:m_eff, :m_eff]
            Tm = 0.5 * (Tm + Tm.conj().T)
            theta, S = np.linalg.eigh(Tm)
            order = _order(theta, which)
            theta, S = theta[order], S[:, order]
            resid = np.abs(beta * S[m_eff - 1, :])
            ok = resid[:k] <= tol_eff * np.maximum(np.abs(theta[:k]), 1.0)
            if np.all(ok) or m_eff < m:
                break
            restarts += 1
            if restarts > maxiter:
                raise RuntimeError(f"Lanczos did not converge in {maxiter} restarts "
                                   f"(residuals {resid[:k]}, tol {tol_eff})")
            # thick restart: keep p Ritz vectors, the residual vector follows them
            p = min(m - 1, k + max(1, (m - k) // 2))
            combine(V[:p], V[:m], S[:, :p])
            V[p], V[m] = V[m], V[p]
            b = beta * S[m - 1, :p]
            T = np.zeros((m + 1, m), dtype=np.complex128)
            T[:p, :p] = np.diag(theta[:p])
            T[p, :p] = b
            start = p
        if info is not None:
            info.update(matvecs=products, restarts=restarts, ncv=m, residuals=resid[:k].tolist())
        vals = theta[:k]
        asc = np.argsort(vals, kind="stable")
        vals = vals[asc]
        if not (return_eigenvectors or return_states):
            return vals
        kk = min(k, m_eff)
        combine(V[:kk], V[:m_eff], S[:, :kk][:, asc])
        if return_states:
            keep, V = V[:kk], V[kk:]
            return vals, keep
        vecs = np.stack([V[i].download().astype(np.complex128) for i in range(kk)], axis=1)
        return vals, vecs
    finally:
        for h in V:
            h.close()


def ground_state_energy(hamiltonian, n_qubits: Optional[int] = None, **kw) -> float:
    """Lowest eigenvalue (``eigsh(k=1, which='SA')``) of a Pauli-sum
    Hamiltonian, the quantity the reference's sparse ground-state path
    returns (``hamiltonians.py:512-577``)."""
    return float(eigsh(hamiltonian, k=1, which="SA", n_qubits=n_qubits, **kw)[0])
