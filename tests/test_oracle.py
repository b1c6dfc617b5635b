"""Pins the CPU oracle before it is trusted as the checker.

Every oracle function is checked against something it does not share code
with: closed forms, brute-force Kronecker products built from bit arithmetic,
a second formulation (eager contraction vs. index gather), known answers
from the reference's own tests, scipy's expm, and the three gradient rules
against each other.
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg

from oracle import sv_oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def rand_state(n, rng, dtype=np.complex128):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return (v / np.linalg.norm(v)).astype(dtype)


def random_ops(n, count, rng):
    rows = []
    for _ in range(count):
        if n > 1 and rng.random() < 0.35:
            g = int(rng.choice([20, 21, 22, 23, 24, 25]))
            a, b = rng.choice(n, 2, replace=False)
            rows.append([g, a, b, rng.uniform(-np.pi, np.pi)])
        else:
            rows.append([int(rng.integers(0, 15)), int(rng.integers(0, n)), 0, rng.uniform(-np.pi, np.pi)])
    return np.array(rows, dtype=np.float64)


# ---------------------------------------------------------------- gate tables
def test_gate_tables_are_unitary_and_closed_forms():
    for g in range(15):
        m = O.gate_1q_matrix(g, 0.731)
        assert np.allclose(m.conj().T @ m, np.eye(2), atol=1e-14)
    for g in (20, 21, 22, 23, 24, 25):
        m = O.gate_2q_matrix(g, -1.3)
        assert np.allclose(m.conj().T @ m, np.eye(4), atol=1e-14)
    # sx^2 = x (reference test_simulator.py:234-239)
    assert np.allclose(O.GATES['sx'] @ O.GATES['sx'], O.GATES['x'])
    # rotations are exp(-i theta sigma / 2)
    th = 0.9
    for g, s in ((9, 'X'), (10, 'Y'), (11, 'Z')):
        assert np.allclose(O.gate_1q_matrix(g, th), scipy.linalg.expm(-0.5j * th * O.PAULI[s]), atol=1e-14)
    # crz != cp (reference test_simulator.py:220-232)
    assert not np.allclose(O.gate_2q_matrix(22, 0.7), O.gate_2q_matrix(25, 0.7))
    # t^2 = s, s^2 = z
    assert np.allclose(O.GATES['t'] @ O.GATES['t'], O.GATES['s'])
    assert np.allclose(O.GATES['s'] @ O.GATES['s'], O.GATES['z'])


# ------------------------------------------------------------ gate application
@pytest.mark.parametrize("n", [1, 2, 3, 5])
def test_gather_kernel_matches_kronecker(n, rng):
    sv = rand_state(n, rng)
    for op in random_ops(n, 30, rng):
        g = int(op[0])
        if g <= 14:
            full = O.kron_operator(O.gate_1q_matrix(g, op[3]), [int(op[1])], n)
        else:
            if n < 2:
                continue
            full = O.kron_operator(O.gate_2q_matrix(g, op[3]), [int(op[1]), int(op[2])], n)
        expect = full @ sv
        sv = O.apply_op_gather(sv, op)
        assert np.allclose(sv, expect, atol=1e-12)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_fast_path_matches_gather(dtype, rng):
    n = 7
    ops = random_ops(n, 120, rng)
    a = O.run_ops(rand_state(n, rng, dtype), ops)
    sv0 = rand_state(n, np.random.default_rng(1234), dtype)
    a = O.run_ops(sv0, ops)
    b = O.run_ops_fast(sv0, ops)
    assert b.dtype == dtype
    assert np.allclose(a, b, atol=1e-12 if dtype == np.complex128 else 1e-5)


def test_eager_matches_kronecker_k_qubits(rng):
    n = 6
    sv = rand_state(n, rng)
    for k in (1, 2, 3, 4):
        q = list(rng.choice(n, k, replace=False))
        z = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
        u, _ = np.linalg.qr(z)
        assert np.allclose(O.apply_matrix(sv, u, q), O.kron_operator(u, q, n) @ sv, atol=1e-12)


def test_two_qubit_index_order_is_q1_high():
    # index = 2*bit(q1) + bit(q2) (reference statevector.py:188-203):
    # cx(0,1) on |10> gives |11>
    sv = O.zero_state(2)
    sv = O.apply_op_gather(sv, [2, 0, 0, 0])
    sv = O.apply_op_gather(sv, [20, 0, 1, 0])
    assert np.argmax(np.abs(sv)) == 3


def test_known_answers_from_reference_tests():
    with open(os.path.join(GOLD, "known_answers.json")) as f:
        cases = json.load(f)
    for case in cases:
        n = case["n"]
        sv = O.run_circuit([tuple(c) for c in case["circuit"]], n)
        p = np.abs(sv) ** 2
        for idx, val in case.get("probs", {}).items():
            assert abs(p[int(idx)] - val) < 1e-9, case["name"]
        for term, val in case.get("expect", {}).items():
            assert abs(O.pauli_expectation(sv, term) - val) < 1e-9, case["name"]


def test_float32_path_tracks_float64(rng):
    ops = random_ops(6, 80, rng)
    a = O.run_ops(O.zero_state(6, np.complex64), ops)
    b = O.run_ops(O.zero_state(6), ops)
    assert a.dtype == np.complex64
    assert np.allclose(a, b, atol=1e-5)


# ------------------------------------------------------------------ transpiler
def _unitary_of(circuit, n):
    cols = []
    for i in range(1 << n):
        e = np.zeros(1 << n, dtype=np.complex128)
        e[i] = 1
        cols.append(O.run_circuit(circuit, n, sv=e))
    return np.array(cols).T


def test_transpiler_decompositions_are_exact(rng):
    assert np.allclose(_unitary_of([('ccx', 0, 1, 2)], 3), O.kron_operator(np.eye(8)[[0, 1, 2, 3, 4, 5, 7, 6]], [0, 1, 2], 3))
    for name in ('swap', 'iswap', 'ecr'):
        assert np.allclose(_unitary_of([(name, 0, 1)], 2), O.GATES[name], atol=1e-12), name
    th, ph, la = rng.uniform(-np.pi, np.pi, 3)
    assert np.allclose(_unitary_of([('u3', 0, th, ph, la)], 1), O.parametric_matrix('u3', th, ph, la), atol=1e-12)
    assert np.allclose(_unitary_of([('u2', 0, ph, la)], 1), O.parametric_matrix('u2', ph, la), atol=1e-12)
    assert O.transpile([('u3', 0)]) == [('u3', 0)]
    assert len(O.transpile([('ccx', 0, 1, 2)])) == 15


def test_encode_ops_errors():
    with pytest.raises(ValueError):
        O.encode_ops([('foo', 0)], 2)
    with pytest.raises(ValueError):
        O.encode_ops([('h', 3)], 2)


# ---------------------------------------------------------------------- Pauli
def test_pauli_expectation_matches_kronecker(rng):
    n = 5
    sv = rand_state(n, rng)
    for _ in range(40):
        w = int(rng.integers(0, 5))
        qs = rng.choice(n, w, replace=False)
        term = {int(q): 'XYZ'[int(rng.integers(0, 3))] for q in qs}
        H = O.pauli_matrix([(1.0, term)], n)
        assert abs(O.pauli_expectation(sv, term) - np.real(np.vdot(sv, H @ sv))) < 1e-12
        assert np.allclose(O.apply_pauli(sv, O.normalize_terms(term), n), H @ sv, atol=1e-12)


def test_pauli_known_values():
    bell = np.zeros(4, dtype=np.complex128)
    bell[0] = bell[3] = 1 / np.sqrt(2)
    assert abs(O.pauli_expectation(bell, 'ZZ') - 1) < 1e-12
    assert abs(O.pauli_expectation(bell, 'XX') - 1) < 1e-12
    assert abs(O.pauli_expectation(bell, 'YY') + 1) < 1e-12
    ghz = np.zeros(8, dtype=np.complex128)
    ghz[0] = ghz[7] = 1 / np.sqrt(2)
    assert abs(O.pauli_expectation(ghz, 'ZZZ')) < 1e-12


# -------------------------------------------------------------- gradients
VQE_QASM_OPS = [  # reference test_autodiff.py:27-31
    {'name': 'ry', 'qubits': [0]}, {'name': 'rx', 'qubits': [1]}, {'name': 'cx', 'qubits': [0, 1]},
    {'name': 'rz', 'qubits': [1]}, {'name': 'cx', 'qubits': [0, 1]}, {'name': 'ry', 'qubits': [0]},
    {'name': 'rx', 'qubits': [1]},
]


def test_three_gradient_rules_agree(rng):
    n = 4
    circ = []
    for layer in range(2):
        circ += [{'name': 'ry', 'qubits': [q]} for q in range(n)]
        circ += [{'name': 'rz', 'qubits': [q]} for q in range(n)]
        circ += [{'name': 'cz', 'qubits': [q, q + 1]} for q in range(n - 1)]
        circ += [{'name': 'rx', 'qubits': [0]}, {'name': 'cx', 'qubits': [1, 2]}]
    tmpl = O.build_template(circ, n)
    theta = rng.uniform(-np.pi, np.pi, int((tmpl[:, 3] == -1).sum()))
    terms = [(0.7, {0: 'X', 2: 'Y'}), (-1.1, {1: 'Z'}), (0.3, {0: 'Z', 3: 'Z'}), (0.5, {2: 'X', 3: 'Y'})]
    e, g_adj = O.adjoint_gradient(tmpl, theta, terms, n)
    g_ps = O.param_shift_gradient(tmpl, theta, terms, n)
    g_fd = O.fd_gradient(tmpl, theta, terms, n, h=1e-5)
    assert np.allclose(g_adj, g_ps, atol=1e-12)
    assert np.allclose(g_adj, g_fd, atol=1e-8)
    H = O.pauli_matrix(terms, n)
    sv = O.run_ops(O.zero_state(n), O.patch_template(tmpl, theta))
    assert abs(e - np.real(np.vdot(sv, H @ sv))) < 1e-12


def test_adjoint_handles_phase_generators(rng):
    n = 3
    circ = [{'name': 'h', 'qubits': [q]} for q in range(n)] + [
        {'name': 'p', 'qubits': [0]}, {'name': 'cp', 'qubits': [0, 2]}, {'name': 'crz', 'qubits': [1, 2]},
        {'name': 'u1', 'qubits': [1]}, {'name': 'ry', 'qubits': [2]}, {'name': 'cx', 'qubits': [2, 0]}]
    tmpl = O.build_template(circ, n)
    theta = rng.uniform(-np.pi, np.pi, int((tmpl[:, 3] == -1).sum()))
    terms = [(1.0, {0: 'X'}), (0.4, {1: 'Y', 2: 'X'}), (0.2, {0: 'Y', 2: 'Z'})]
    _, g = O.adjoint_gradient(tmpl, theta, terms, n)
    assert np.allclose(g, O.fd_gradient(tmpl, theta, terms, n, h=1e-5), atol=1e-8)


def test_vqe_qasm_fixture_matches_dense_hamiltonian():
    """Reference fixture: VQE_QASM, H = diag(sort(U(-2.5, 2.5, 4))) seed 7,
    theta ~ U(-pi, pi) seed 3 (test_autodiff.py:27-69)."""
    tmpl = O.build_template(VQE_QASM_OPS, 2)
    assert int((tmpl[:, 3] == -1).sum()) == 5
    hv = np.sort(np.random.default_rng(7).uniform(-2.5, 2.5, 4))
    theta = np.random.default_rng(3).uniform(-np.pi, np.pi, 5)
    # diagonal H as a Pauli sum: h = sum_P c_P P over {I,Z}^2
    terms = []
    for a in 'IZ':
        for b in 'IZ':
            P = np.kron(O.PAULI[a], O.PAULI[b])
            terms.append((float(np.real(np.trace(P @ np.diag(hv))) / 4), a + b))
    e, g = O.adjoint_gradient(tmpl, theta, terms, 2)
    sv = O.run_ops(O.zero_state(2), O.patch_template(tmpl, theta))
    assert abs(e - np.real(np.vdot(sv, np.diag(hv) @ sv))) < 1e-12
    assert np.allclose(g, O.fd_gradient(tmpl, theta, terms, 2, h=1e-6), atol=1e-6)


# ------------------------------------------------------------------ Trotter
def test_pauli_rotation_is_exact_exponential(rng):
    n = 3
    for d in ({0: 'X', 1: 'Y'}, {0: 'Z', 2: 'Z'}, {1: 'Y'}, {0: 'X', 1: 'Z', 2: 'Y'}):
        a = float(rng.uniform(-1, 1))
        U = _unitary_of(O.pauli_rotation_ops(d, a), n)
        P = O.pauli_matrix([(1.0, d)], n)
        assert np.allclose(U, scipy.linalg.expm(-1j * a * P), atol=1e-12)


def test_trotter_converges_to_expm():
    n = 3
    terms = [(1.0, {0: 'X', 1: 'X'}), (1.0, {0: 'Y', 1: 'Y'}), (0.5, {0: 'Z', 1: 'Z'}),
             (1.0, {1: 'X', 2: 'X'}), (1.0, {1: 'Y', 2: 'Y'}), (0.5, {1: 'Z', 2: 'Z'})]
    H = O.pauli_matrix(terms, n)
    psi0 = np.zeros(8, dtype=np.complex128)
    psi0[1] = 1
    exact = scipy.linalg.expm(-1j * H * 0.5) @ psi0
    e1 = np.linalg.norm(O.run_circuit(O.trotter_evolve_ops(terms, 0.5, 20), n, sv=psi0) - exact)
    e2 = np.linalg.norm(O.run_circuit(O.trotter_evolve_ops(terms, 0.5, 40), n, sv=psi0) - exact)
    assert e2 < e1 < 0.05
    assert 1.6 < e1 / e2 < 2.4  # first order


# ---- stochastic Kraus channels (oracle restatement of the NumPy path) -------

def _kraus_average(sv, n, model, p, shots, seed):
    rng = np.random.default_rng(seed)
    rho = np.zeros((1 << n, 1 << n), dtype=np.complex128)
    for _ in range(shots):
        v = O.noise_apply_to_sv(sv, n, model, p, rng)
        rho += np.outer(v, v.conj())
    return rho / shots


def _lsb_op(m, q, n):
    """2x2 operator on bit q (LSB-first, the noise indexing) of an n-bit index."""
    out = np.array([[1.0 + 0j]])
    for b in reversed(range(n)):
        out = np.kron(out, m if b == q else np.eye(2))
    return out


def test_noise_closed_forms():
    rng = np.random.default_rng(2)
    n = 4
    sv = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    sv /= np.linalg.norm(sv)
    # p = 1: the flip always fires on every qubit
    assert np.allclose(O.noise_apply_to_sv(sv, n, 'bitflip', 1.0, rng), sv[::-1])
    parity = np.array([(-1) ** bin(i).count('1') for i in range(1 << n)])
    assert np.allclose(O.noise_apply_to_sv(sv, n, 'phaseflip', 1.0, rng), parity * sv)
    assert O.noise_apply_to_sv(sv, n, 'ideal', 0.3, rng) is sv
    assert O.noise_apply_to_sv(sv, n, 'depolarizing', 0.0, rng) is sv
    # a full decay of |1...1> on every qubit lands in |0...0>
    one = np.zeros(1 << n, dtype=np.complex128)
    one[-1] = 1
    assert np.allclose(O.noise_apply_to_sv(one, n, 'amplitude_damping', 1.0, rng), np.eye(1 << n)[0])


@pytest.mark.parametrize("model,p", [("depolarizing", 0.3), ("amplitude_damping", 0.4)])
def test_noise_trajectories_average_to_the_kraus_map(model, p):
    """The average of |psi'><psi'| over trajectories equals the Kraus map
    applied qubit by qubit (Monte Carlo, 6000 shots, tolerance 0.03)."""
    n = 2
    sv = np.array([0.6, 0.48j, -0.36, 0.52], dtype=np.complex128)
    sv /= np.linalg.norm(sv)
    rho = np.outer(sv, sv.conj())
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    Y = np.array([[0, -1j], [1j, 0]])
    Z = np.diag([1.0 + 0j, -1.0])
    for q in range(n):
        if model == "depolarizing":
            ks = [np.sqrt(1 - p) * np.eye(2), np.sqrt(p / 3) * X, np.sqrt(p / 3) * Y, np.sqrt(p / 3) * Z]
        else:
            ks = [np.diag([1.0, np.sqrt(1 - p)]) + 0j, np.array([[0, np.sqrt(p)], [0, 0]], dtype=complex)]
        ks = [_lsb_op(k, q, n) for k in ks]
        rho = sum(k @ rho @ k.conj().T for k in ks)
    est = _kraus_average(sv, n, model, p, 6000, seed=5)
    assert np.max(np.abs(est - rho)) < 0.03


def _measure0_freq(sv, model, p, shots, seed):
    """Frequency of outcome 0 when each 1-qubit trajectory is measured."""
    rng = np.random.default_rng(seed)
    zeros = 0
    for _ in range(shots):
        v = O.noise_apply_to_sv(sv.copy(), 1, model, p, rng)
        pr = np.abs(v) ** 2
        zeros += rng.choice(2, p=pr / pr.sum()) == 0
    return zeros / shots


@pytest.mark.parametrize("sv,model,p,want", [
    # the reference's own statistical known answers (tests/unit/test_registry.py:68-236)
    (np.array([0, 1], dtype=complex), "depolarizing", 0.3, 2 * 0.3 / 3),
    (np.array([0, 1], dtype=complex), "combined", 0.6, 2 * 0.3 / 3 + (1 - 2 * 0.3 / 3) * 0.6 * 0.333333),
    (np.array([0, 1], dtype=complex), "amplitude_damping", 0.4, 0.4),
    (np.array([1, 1], dtype=complex) / np.sqrt(2), "amplitude_damping", 0.5, 0.75),
])
def test_noise_reference_known_answers(sv, model, p, want):
    assert abs(_measure0_freq(sv, model, p, 20000, seed=3) - want) < 0.02
