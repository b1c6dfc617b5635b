"""Seeded workload generators: shapes, counts and the Haar SU(4) invariants."""
import numpy as np

from desv import workloads
from oracle import sv_oracle as O


def test_haar_su4_is_special_unitary():
    rng = np.random.default_rng(0)
    for _ in range(50):
        u = workloads.haar_su4(rng)
        assert np.allclose(u.conj().T @ u, np.eye(4), atol=1e-12)
        assert abs(np.linalg.det(u) - 1) < 1e-12


def test_brickwork_structure():
    g = workloads.brickwork_su4(30, 24, seed=0)
    assert len(g) == 348
    assert [(a, b) for _, a, b in g[:15]] == [(q, q + 1) for q in range(0, 29, 2)]
    assert [(a, b) for _, a, b in g[15:29]] == [(q, q + 1) for q in range(1, 28, 2)]
    g2 = workloads.brickwork_su4(30, 24, seed=0)
    assert all(np.array_equal(x[0], y[0]) for x, y in zip(g, g2))


def test_plain_circuit_counts():
    c = workloads.random_layers_circuit(20, 20, seed=0)
    assert len(c) == 20 * (3 * 20 + 19) == 1580


def test_vqe_instance_and_batch():
    circ, theta, terms = workloads.vqe_instance(20, 8, 200, seed=0)
    tmpl = O.build_template(circ.ops, 20)
    assert len(tmpl) == 8 * (20 + 20 + 19) == 472
    assert int((tmpl[:, 3] == -1).sum()) == len(theta) == 320
    assert len(terms) == 200 and all(1 <= len(t) <= 4 for _, t in terms)
    b = workloads.batch_instances(count=40)
    assert all(10 <= n <= 22 for n, *_ in b)
    assert len({n for n, *_ in b}) > 5


def test_xxz_trotter_gate_count():
    c = workloads.xxz_trotter_circuit(28, 0.05, 200, seed=0)
    # 21 gates per bond per step (XX: 2h+2cx+rz+2h ... ) plus the 2*28 preparation rotations
    per_step = 27 * (7 + 11 + 3)
    assert len(c) == 2 * 28 + 200 * per_step
