"""The C restatement of the oracle (oracle/c) against the NumPy oracle:
op tables with every gate id and arbitrary k-qubit matrices, in complex128
(to 1e-12) and complex64 (to FP32 rounding), with 1 and several OpenMP
threads (identical results: every group of amplitudes is independent)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import sv_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def C():
    from oracle import sv_oracle_c
    if not os.path.exists(sv_oracle_c.LIB_PATH):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle", "c")], check=True)
    return sv_oracle_c


def random_ops(n, count, rng):
    rows = []
    for _ in range(count):
        if n > 1 and rng.random() < 0.4:
            g = int(rng.choice([20, 21, 22, 23, 24, 25]))
            a, b = rng.choice(n, 2, replace=False)
            rows.append([g, a, b, rng.uniform(-np.pi, np.pi)])
        else:
            rows.append([int(rng.integers(0, 15)), int(rng.integers(0, n)), 0, rng.uniform(-np.pi, np.pi)])
    return np.array(rows, dtype=np.float64)


def rand_state(n, rng):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("n", [1, 2, 3, 6, 10])
def test_op_tables(C, n, rng):
    ops = random_ops(n, 120, rng)
    sv0 = rand_state(n, rng)
    want = O.run_ops(sv0.copy(), ops)
    got = C.run_ops(sv0.copy(), ops, threads=1)
    assert np.allclose(got, want, atol=1e-12)
    got4 = C.run_ops(sv0.copy(), ops, threads=4)
    assert np.array_equal(got, got4)
    g32 = C.run_ops(sv0.astype(np.complex64), ops, threads=3)
    assert O.fidelity(g32, want) > 1 - 1e-6


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_k_qubit_matrices(C, k, rng):
    n = 9
    sv = rand_state(n, rng)
    want = sv.copy()
    got = sv.copy()
    for _ in range(4):
        q = [int(x) for x in rng.choice(n, k, replace=False)]
        u, _ = np.linalg.qr(rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k)))
        want = O.apply_matrix(want, u, q)
        C.apply_matrix(got, u, q, threads=2)
    assert np.allclose(got, want, atol=1e-12)


def test_errors(C):
    sv = np.zeros(8, dtype=np.complex128)
    with pytest.raises(ValueError):
        C.run_ops(sv, np.array([[99, 0, 0, 0.0]]))
    with pytest.raises(ValueError):
        C.apply_matrix(sv, np.eye(4), [0, 0])
    with pytest.raises(ValueError):
        C.apply_matrix(np.zeros(6, dtype=np.complex128), np.eye(2), [0])
