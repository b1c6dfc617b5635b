"""Host-side drop-in helpers and the oracle's gate tables, pinned by
mathematics (no GPU):

* the oracle's 2-qubit tables against identities built from its 1-qubit
  tables (which ``test_oracle.py`` pins to ``expm``): CZ = (I (x) H) CX (I (x) H),
  CY = (I (x) S) CX (I (x) Sdg), CP(l) and CRZ(t) from CX and rotations, SWAP
  = three CXs, and the controlled forms |0><0| (x) I + |1><1| (x) U;
* ``desv.GATES`` / ``desv.PARAMETRIC_GATES``: the reference's key sets
  (``circuits/gates.py:33-91``), every entry equal to the oracle's table and
  to a closed form;
* ``pauli_hamiltonian_to_matrix`` against Kronecker products of Pauli
  matrices and the oracle's dense H; ``multiply_pauli_terms`` against the
  reference's docstring examples (``physics/observables.py:503-506``) and
  matrix products;
* ``entangling_layer`` (the cases of ``tests/unit/test_topology.py:12-60``),
  ``random_circuit`` (``tests/unit/test_random_circuit.py:12-60``) and
  ``set_precision`` / ``ensure_x64`` (``config.py:29-62``).
"""
import itertools

import numpy as np
import pytest
import scipy.linalg

import desv
from desv import config
from oracle import sv_oracle as O

I2 = np.eye(2, dtype=np.complex128)
P0 = np.diag([1, 0]).astype(np.complex128)
P1 = np.diag([0, 1]).astype(np.complex128)
PAULI = {'I': I2, 'X': np.array([[0, 1], [1, 0]], complex), 'Y': np.array([[0, -1j], [1j, 0]], complex),
         'Z': np.array([[1, 0], [0, -1]], complex)}


def ctrl(u):
    """|0><0| (x) I + |1><1| (x) u, control on the more significant index bit."""
    return np.kron(P0, I2) + np.kron(P1, u)


def g1(name, *a):
    return O.gate_1q_matrix(O.GATE_IDS[name], a[0] if a else 0.0)


# ------------------------------------------------------ oracle 2-qubit tables
def test_oracle_two_qubit_tables_from_one_qubit_identities():
    X, H, S, Sdg = g1('x'), g1('h'), g1('s'), g1('sdg')
    CX = O.gate_2q_matrix(20, 0.0)
    assert np.allclose(CX, ctrl(X), atol=1e-15)
    # CZ = (I H) CX (I H); also the controlled form of Z
    assert np.allclose(O.gate_2q_matrix(21, 0.0), np.kron(I2, H) @ CX @ np.kron(I2, H), atol=1e-15)
    assert np.allclose(O.gate_2q_matrix(21, 0.0), ctrl(g1('z')), atol=1e-15)
    # CY = (I S) CX (I Sdg): a swapped +-i in the table fails this
    assert np.allclose(O.gate_2q_matrix(24, 0.0), np.kron(I2, S) @ CX @ np.kron(I2, Sdg), atol=1e-15)
    assert np.allclose(O.gate_2q_matrix(24, 0.0), ctrl(g1('y')), atol=1e-15)
    # SWAP = CX (reversed CX) CX
    CXr = np.kron(I2, P0) + np.kron(X, P1)
    assert np.allclose(O.gate_2q_matrix(23, 0.0), CX @ CXr @ CX, atol=1e-15)
    for th in (0.37, -1.9, 2.8):
        # CRZ(t) = (I Rz(t/2)) CX (I Rz(-t/2)) CX = controlled Rz(t)
        crz = np.kron(I2, g1('rz', th / 2)) @ CX @ np.kron(I2, g1('rz', -th / 2)) @ CX
        assert np.allclose(O.gate_2q_matrix(25, th), crz, atol=1e-14)
        assert np.allclose(O.gate_2q_matrix(25, th), ctrl(g1('rz', th)), atol=1e-14)
        # CP(l) = controlled P(l), and P(l) = e^{i l / 2} Rz(l)
        assert np.allclose(O.gate_2q_matrix(22, th), ctrl(g1('p', th)), atol=1e-14)
        assert np.allclose(g1('p', th), np.exp(0.5j * th) * g1('rz', th), atol=1e-14)
        # CP(l) = (P(l/2) I) CX (I P(-l/2)) CX (I P(l/2))
        cp = np.kron(g1('p', th / 2), I2) @ CX @ np.kron(I2, g1('p', -th / 2)) @ CX @ np.kron(I2, g1('p', th / 2))
        assert np.allclose(O.gate_2q_matrix(22, th), cp, atol=1e-14)


# --------------------------------------------------------- drop-in gate tables
REF_GATES = {'h', 'x', 'y', 'z', 's', 'sdg', 't', 'tdg', 'sx', 'id', 'cx', 'cz', 'cy', 'swap', 'iswap', 'ecr', 'ccx'}
REF_PARAMETRIC = {'rx', 'ry', 'rz', 'cp', 'crz', 'u3', 'u2', 'u1', 'p', 'gphase'}


def test_gates_table_matches_oracle_and_closed_forms():
    assert set(desv.GATES) == REF_GATES
    for name, m in desv.GATES.items():
        assert m.dtype == np.complex128
        d = m.shape[0]
        assert np.allclose(m.conj().T @ m, np.eye(d), atol=1e-15), name
        if name in O.GATES:
            assert np.array_equal(m, O.GATES[name]), name
    G = desv.GATES
    X, Y, Z = PAULI['X'], PAULI['Y'], PAULI['Z']
    assert np.allclose(G['h'], (X + Z) / np.sqrt(2))
    assert np.allclose(G['sx'] @ G['sx'], X)
    assert np.allclose(G['t'] @ G['t'], G['s']) and np.allclose(G['s'] @ G['sdg'], I2)
    assert np.allclose(G['tdg'], G['t'].conj().T)
    # iSWAP = exp(i pi/4 (XX + YY)); ECR = (X (x) I - Y (x) X) / sqrt 2, first factor on
    # the first qubit (the more significant index bit)
    assert np.allclose(G['iswap'], scipy.linalg.expm(0.25j * np.pi * (np.kron(X, X) + np.kron(Y, Y))), atol=1e-14)
    assert np.allclose(G['ecr'], (np.kron(X, I2) - np.kron(Y, X)) / np.sqrt(2), atol=1e-15)
    # CCX: the Toffoli permutation (|110> <-> |111>)
    assert np.array_equal(G['ccx'], np.eye(8)[[0, 1, 2, 3, 4, 5, 7, 6]])
    # the 2-qubit entries are the controlled forms (control = first qubit)
    assert np.allclose(G['cx'], ctrl(X)) and np.allclose(G['cz'], ctrl(Z)) and np.allclose(G['cy'], ctrl(Y))


@pytest.mark.parametrize("a", [0.0, 0.41, -2.3, 3.9])
def test_parametric_gates_match_oracle_and_closed_forms(a):
    P = desv.PARAMETRIC_GATES
    assert set(P) == REF_PARAMETRIC
    for name, sig in (('rx', 'X'), ('ry', 'Y'), ('rz', 'Z')):
        assert np.allclose(P[name](a), scipy.linalg.expm(-0.5j * a * PAULI[sig]), atol=1e-14)
        assert np.allclose(P[name](a), O.parametric_matrix(name, a), atol=1e-15)
    for name in ('p', 'u1'):
        assert np.allclose(P[name](a), np.diag([1, np.exp(1j * a)]), atol=1e-15)
        assert np.allclose(P[name](a), O.parametric_matrix(name, a), atol=1e-15)
    assert np.allclose(P['gphase'](a), np.exp(1j * a) * I2)
    assert np.allclose(P['cp'](a), ctrl(P['p'](a))) and np.allclose(P['crz'](a), ctrl(P['rz'](a)))
    assert np.allclose(P['cp'](a), O.gate_2q_matrix(22, a)) and np.allclose(P['crz'](a), O.gate_2q_matrix(25, a))
    th, ph, la = a, 0.7 - a, 1.3 * a
    # u3(t, p, l) = e^{i (p + l) / 2} Rz(p) Ry(t) Rz(l); u2(p, l) = u3(pi/2, p, l)
    u3 = np.exp(0.5j * (ph + la)) * P['rz'](ph) @ P['ry'](th) @ P['rz'](la)
    assert np.allclose(P['u3'](th, ph, la), u3, atol=1e-14)
    assert np.allclose(P['u2'](ph, la), P['u3'](np.pi / 2, ph, la), atol=1e-14)
    assert np.allclose(P['u3'](th, ph, la), O.parametric_matrix('u3', th, ph, la), atol=1e-15)


# ------------------------------------------------------------ Pauli algebra
def kron_term(d, n):
    m = np.array([[1.0 + 0j]])
    for q in range(n):
        m = np.kron(m, PAULI[d.get(q, 'I')])
    return m


def test_pauli_hamiltonian_to_matrix(rng):
    n = 4
    terms = [(1.0, 'ZZII'), (0.5, {0: 'X'}), (-0.3 + 0.2j, [(1, 'Y'), (3, 'X')]), (0.8, {2: 'Y', 0: 'Y', 3: 'Z'}),
             (1.1, 'IIII')]
    H = desv.pauli_hamiltonian_to_matrix(terms, n)
    want = sum(c * kron_term(O.normalize_terms(p, n), n) for c, p in terms)
    assert H.shape == (16, 16) and H.dtype == np.complex128
    assert np.allclose(H, want, atol=1e-15)
    real_terms = [t for t in terms if np.isreal(t[0])]
    Hr = desv.pauli_hamiltonian_to_matrix(real_terms, n)
    assert np.allclose(Hr, O.pauli_matrix(real_terms, n), atol=1e-15)
    assert np.allclose(Hr, Hr.conj().T)
    # every single-letter string on every qubit, random 6-qubit strings
    for q, letter in itertools.product(range(3), 'XYZ'):
        assert np.allclose(desv.pauli_hamiltonian_to_matrix([(1.0, {q: letter})], 3), kron_term({q: letter}, 3))
    for _ in range(10):
        d = {int(q): str(rng.choice(list('XYZ'))) for q in rng.choice(6, int(rng.integers(1, 6)), replace=False)}
        assert np.allclose(desv.pauli_hamiltonian_to_matrix([(0.7, d)], 6), 0.7 * kron_term(d, 6))
    with pytest.raises(ValueError):
        desv.pauli_hamiltonian_to_matrix([(1.0, {3: 'X'})], 2)
    with pytest.raises(ValueError):
        desv.pauli_hamiltonian_to_matrix([(1.0, 'X')], 0)


def test_multiply_pauli_terms(rng):
    # the reference's docstring examples
    assert desv.multiply_pauli_terms([(1.0, 'X'), (1.0, 'Y')]) == (1j, {0: 'Z'})
    assert desv.multiply_pauli_terms([(2.0, 'X'), (3.0, {1: 'Z'})]) == (6.0, {0: 'X', 1: 'Z'})
    # P P = I on every letter; the cyclic products
    for a in 'XYZ':
        assert desv.multiply_pauli_terms([(1.0, {0: a}), (1.0, {0: a})]) == (1.0, {})
    assert desv.multiply_pauli_terms([(1.0, 'Y'), (1.0, 'X')]) == (-1j, {0: 'Z'})
    assert desv.multiply_pauli_terms([(1.0, 'Z'), (1.0, 'X')]) == (1j, {0: 'Y'})
    # random products of up to four 4-qubit strings against matrix products
    n = 4
    for _ in range(40):
        facs = []
        for _ in range(int(rng.integers(1, 5))):
            d = {int(q): str(rng.choice(list('XYZ'))) for q in rng.choice(n, int(rng.integers(0, n + 1)), replace=False)}
            facs.append((complex(rng.normal(), rng.normal()), d))
        c, d = desv.multiply_pauli_terms(facs)
        want = np.eye(1 << n, dtype=np.complex128)
        for cf, df in facs:
            want = want @ (cf * kron_term(df, n))
        assert np.allclose(c * kron_term(d, n), want, atol=1e-12)


# ---------------------------------------------------------------- topology
def test_entangling_layer_cases():
    el = desv.entangling_layer
    assert el(4, pattern='linear') == [('cx', 0, 1), ('cx', 1, 2), ('cx', 2, 3)]
    assert el(4, pattern='circular') == [('cx', 0, 1), ('cx', 1, 2), ('cx', 2, 3), ('cx', 3, 0)]
    assert el(2, pattern='circular') == el(2, pattern='linear')
    full = el(4, pattern='full')
    assert {(a, b) for _, a, b in full} == {(i, j) for i in range(4) for j in range(i + 1, 4)} and len(full) == 6
    assert {(a, b) for _, a, b in el(5, pattern='star', hub=2)} == {(2, 0), (2, 1), (2, 3), (2, 4)}
    assert [(a, b) for _, a, b in el(5, pattern='brick')] == [(0, 1), (2, 3), (1, 2), (3, 4)]
    assert all(op[0] == 'cz' for op in el(3, pattern='linear', gate='cz'))
    assert el(3, pattern='linear', reverse=True) == [(g, b, a) for g, a, b in el(3, pattern='linear')]
    for bad in (lambda: el(4, pattern='star', hub=9), lambda: el(1), lambda: el(3, pattern='hexagonal')):
        with pytest.raises(ValueError):
            bad()


# ----------------------------------------------------------- random circuits
def test_random_circuit_cases():
    rc = desv.random_circuit
    assert len(rc(4, 25, seed=1)) == 25 and rc(3, 0, seed=1) == []
    assert rc(4, 30, seed=42) == rc(4, 30, seed=42) and rc(4, 30, seed=1) != rc(4, 30, seed=2)
    one = rc(1, 30, seed=3)
    assert all(len(op) == 2 or op[0] in ('rx', 'ry', 'rz') for op in one) and all(op[1] == 0 for op in one)
    assert all(op[0] in ('h', 'cx') for op in rc(4, 40, seed=9, gate_set=['h', 'cx']))
    for bad in (lambda: rc(3, 5, gate_set=['not_a_gate']), lambda: rc(0, 5), lambda: rc(3, -1),
                lambda: rc(3, 5, two_qubit_prob=1.5), lambda: rc(1, 5, gate_set=['cx'])):
        with pytest.raises(ValueError):
            bad()
    # an existing generator is used as is; 2-qubit gates get distinct qubits
    g = np.random.default_rng(5)
    c = rc(6, 200, seed=g, two_qubit_prob=1.0)
    assert all(len(op) == 3 and op[1] != op[2] and op[0] in ('cx', 'cz', 'cy', 'swap') for op in c)
    angles = [op[2] for op in rc(3, 300, seed=4, gate_set=['rx', 'ry', 'rz'])]
    assert len(angles) == 300 and all(0.0 <= a < 2 * np.pi for a in angles)


def test_random_circuit_runs_through_the_oracle():
    c = desv.random_circuit(5, 50, seed=7)
    sv = O.run_circuit(c, 5)
    assert abs(np.vdot(sv, sv).real - 1) < 1e-12


# ----------------------------------------------------------------- precision
def test_set_precision_and_ensure_x64():
    saved = (config._explicit, config._float64)
    try:
        config._explicit, config._float64 = False, True
        desv.ensure_x64()
        assert config.default_float64()
        desv.set_precision(False)
        assert not config.default_float64()
        desv.ensure_x64()        # an explicit choice sticks
        assert not config.default_float64()

        @desv.with_x64
        def f():
            return config.default_float64()
        assert f() is False
        desv.set_precision(True)
        assert f() is True
    finally:
        config._explicit, config._float64 = saved
