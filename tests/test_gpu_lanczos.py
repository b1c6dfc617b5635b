"""Krylov-basis kernels and the device Lanczos solver.

Kernels: multi-vector inner products and in-place combinations against
NumPy on the same vectors.  Solver: against dense diagonalisation of the
oracle's Kronecker-product Hamiltonian (n <= 10), ARPACK (scipy eigsh over the
oracle's matvec, n = 12), the free-fermion closed form of the open
transverse-field Ising chain (E0 = -sum of the singular values of the
bidiagonal matrix diag(h) + superdiag(J); n = 20), and brute force over the
diagonal of a Z-only Hamiltonian (n = 18)."""
import numpy as np
import pytest

import desv
from desv import lanczos, workloads
from oracle import sv_oracle as O

pytestmark = pytest.mark.gpu


def _handles(n, count, seed, c128=True):
    rng = np.random.default_rng(seed)
    hs, arrs = [], []
    for _ in range(count):
        a = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        h = desv.StateHandle(n, c128)
        h.upload(a)
        hs.append(h)
        arrs.append(h.download().astype(np.complex128))
    return hs, arrs


@pytest.mark.parametrize("m", [1, 7, 8, 9, 33, 64])
@pytest.mark.parametrize("c128", [True, False])
def test_inner_many(m, c128):
    n = 11
    hs, arrs = _handles(n, m + 1, seed=m, c128=c128)
    got = lanczos.inner_many(hs[0], hs[1:])
    want = np.array([np.vdot(a, arrs[0]) for a in arrs[1:]])
    assert np.max(np.abs(got - want)) < 1e-9 * np.max(np.abs(want))
    for h in hs:
        h.close()


@pytest.mark.parametrize("m,p", [(1, 1), (5, 1), (9, 3), (20, 10), (40, 2), (64, 63)])
def test_combine_in_place(m, p):
    n = 10
    hs, arrs = _handles(n, m, seed=100 + m)
    rng = np.random.default_rng(m * 7 + p)
    coef = rng.normal(size=(m, p)) + 1j * rng.normal(size=(m, p))
    lanczos.combine(hs[:p], hs, coef)  # outputs alias the first p inputs
    A = np.stack(arrs, axis=1)
    want = A @ coef
    for l in range(p):
        assert np.max(np.abs(hs[l].download() - want[:, l])) < 1e-11 * np.max(np.abs(want))
    for j in range(p, m):  # inputs that are not outputs are untouched
        assert np.array_equal(hs[j].download(), arrs[j])
    for h in hs:
        h.close()


def test_combine_c64_and_fill_random():
    n = 12
    a = desv.StateHandle(n, False)
    b = desv.StateHandle(n, False)
    assert desv._capi.lib.desv_fill_random(a.ptr, 5) == 0
    assert desv._capi.lib.desv_fill_random(b.ptr, 5) == 0
    va, vb = a.download(), b.download()
    assert np.array_equal(va, vb)
    assert np.all(np.abs(va.real) <= 0.5) and np.all(np.abs(va.imag) <= 0.5)
    assert abs(np.mean(va.real)) < 0.02 and abs(np.std(va.real) - np.sqrt(1 / 12)) < 0.02
    assert desv._capi.lib.desv_fill_random(b.ptr, 6) == 0
    vb = b.download()
    assert not np.array_equal(vb, va)
    lanczos.combine([b], [a, b], np.array([[2.0 - 1j], [0.5j]]))
    want = (2 - 1j) * va.astype(np.complex128) + 0.5j * vb.astype(np.complex128)
    assert np.max(np.abs(b.download() - want)) < 1e-6
    a.close()
    b.close()


def test_basis_calls_reject_bad_arguments():
    a = desv.StateHandle(6, True)
    b = desv.StateHandle(7, True)
    with pytest.raises(ValueError, match="share n"):
        lanczos.inner_many(a, [b])
    with pytest.raises(ValueError, match="1..64"):
        lanczos.inner_many(a, [a] * 65)
    a.close()
    b.close()


@pytest.mark.parametrize("k,which", [(1, "SA"), (3, "SA"), (2, "LA"), (2, "LM"), (1, "LA")])
def test_eigenvalues_against_dense(k, which):
    n = 8
    terms = workloads.random_pauli_hamiltonian(np.random.default_rng(3), n, 40)
    ev = np.linalg.eigvalsh(O.pauli_matrix(terms, n))
    key = {"SA": ev, "LA": -ev, "SM": np.abs(ev), "LM": -np.abs(ev)}[which]
    want = np.sort(ev[np.argsort(key, kind="stable")[:k]])
    got = lanczos.eigsh(terms, k=k, which=which, n_qubits=n)
    assert np.max(np.abs(got - want)) < 1e-9, (got, want)


def test_eigenvectors_and_states():
    n = 9
    terms = workloads.random_pauli_hamiltonian(np.random.default_rng(4), n, 60)
    H = O.pauli_matrix(terms, n)
    ev, evec = np.linalg.eigh(H)
    vals, vecs = lanczos.eigsh(desv.PauliSumOperator(terms, n), k=2, return_eigenvectors=True)
    assert np.max(np.abs(vals - ev[:2])) < 1e-9
    for i in range(2):
        v = vecs[:, i]
        assert abs(np.linalg.norm(v) - 1) < 1e-10
        assert np.linalg.norm(O.pauli_sum_matvec(v, terms) - vals[i] * v) < 1e-6
        assert abs(np.vdot(evec[:, i], v)) ** 2 > 1 - 1e-10
    vals2, states = lanczos.eigsh(terms, k=1, n_qubits=n, return_states=True)
    assert abs(vals2[0] - ev[0]) < 1e-9
    assert abs(np.vdot(evec[:, 0], states[0].download())) ** 2 > 1 - 1e-10
    # a start vector that already is the ground state
    e0 = lanczos.eigsh(terms, k=1, n_qubits=n, v0=evec[:, 0])
    assert abs(e0[0] - ev[0]) < 1e-9
    for s in states:
        s.close()


def test_small_space_breakdown():
    # dim 4 with ncv = dim: the Krylov space is exhausted (exact answer)
    terms = [(0.7, {0: 'Z'}), (-0.3, {1: 'X'}), (0.25, {0: 'Y', 1: 'Y'})]
    ev = np.linalg.eigvalsh(O.pauli_matrix(terms, 2))
    got = lanczos.eigsh(terms, k=2, n_qubits=2, ncv=4)
    assert np.max(np.abs(got - ev[:2])) < 1e-12
    # a start vector inside a 1-dimensional invariant subspace
    diag = [(1.0, {0: 'Z'}), (0.5, {1: 'Z'}), (0.25, {2: 'Z'})]
    v0 = np.zeros(8)
    v0[3] = 1.0
    got = lanczos.eigsh(diag, k=2, n_qubits=3, ncv=6, v0=v0)
    d = np.sort(np.real(np.diag(O.pauli_matrix(diag, 3))))
    assert np.max(np.abs(got - d[:2])) < 1e-12


def test_against_arpack_12q():
    import scipy.sparse.linalg as sla
    n = 12
    terms = workloads.random_pauli_hamiltonian(np.random.default_rng(6), n, 80)
    op = sla.LinearOperator((1 << n, 1 << n), matvec=lambda v: O.pauli_sum_matvec(np.ravel(v), terms),
                            dtype=np.complex128)
    want = np.sort(sla.eigsh(op, k=2, which="SA", ncv=20, return_eigenvectors=False))
    got = lanczos.eigsh(terms, k=2, n_qubits=n)
    assert np.max(np.abs(got - want)) < 1e-9
    # the scipy-facing operator of PauliSumOperator (GPU matvec) gives the same
    got2 = np.sort(sla.eigsh(desv.PauliSumOperator(terms, n).as_linear_operator(), k=2, which="SA",
                             ncv=20, return_eigenvectors=False))
    assert np.max(np.abs(got2 - want)) < 1e-9


@pytest.mark.parametrize("n,dtype,tol", [(20, "complex128", 1e-10), (16, "complex64", 2e-5)])
def test_tfim_closed_form(n, dtype, tol):
    terms, J, h = workloads.tfim_terms(n, seed=n)
    want = -np.sum(np.linalg.svd(np.diag(h) + np.diag(J, 1), compute_uv=False))
    got = desv.ground_state_energy(terms, n, dtype=dtype)
    print(f"TFIM {n}q {dtype}: E0 = {got:.12f}, closed form {want:.12f}")
    assert abs(got - want) <= tol * abs(want)


def test_z_only_brute_force_18q():
    n = 18
    rng = np.random.default_rng(9)
    terms = [(float(rng.normal()), {int(a): 'Z', int(b): 'Z'}) for a, b in
             (rng.choice(n, 2, replace=False) for _ in range(30))]
    terms += [(float(rng.normal()), {q: 'Z'}) for q in range(n)]
    idx = np.arange(1 << n)
    diag = np.zeros(1 << n)
    for c, t in terms:
        s = np.ones(1 << n)
        for q in t:
            s *= 1 - 2 * ((idx >> (n - 1 - q)) & 1)
        diag += c * s
    got = desv.ground_state_energy(terms, n)
    assert abs(got - diag.min()) < 1e-9
