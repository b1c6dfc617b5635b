"""Device noise trajectories against the oracle's restatement of the
reference's NumPy path (``noise/kraus_channels.py:104-235``): the same seeded
generator must give the same state (complex128 to 1e-12, complex64 to 1e-5),
and ``energy_fn(noise=...)`` must give the oracle's noisy energy and, for
Pauli channels, its gradient."""
import numpy as np
import pytest

import desv
from desv import workloads
from oracle import sv_oracle as O

pytestmark = pytest.mark.gpu

MODELS = ["depolarizing", "bitflip", "phaseflip", "amplitude_damping", "combined"]


def _random_state(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("model", MODELS)
@pytest.mark.parametrize("qubits", [None, [4, 0, 7], [2]])
def test_host_array_trajectories_match_oracle(model, qubits):
    n = 9
    sv = _random_state(n, 3)
    for seed in range(4):
        got = desv.NoiseModel.apply_to_sv(sv, n, model, 0.4, rng=np.random.default_rng(seed), qubits=qubits)
        want = O.noise_apply_to_sv(sv, n, model, 0.4, np.random.default_rng(seed), qubits)
        assert got.dtype == np.complex128
        assert np.max(np.abs(got - want)) < 1e-12, (model, qubits, seed)
    # the caller's array is not modified
    assert np.array_equal(sv, _random_state(n, 3))


@pytest.mark.parametrize("model", MODELS)
def test_device_state_in_place_c64(model):
    n = 14
    sim = desv.DenseSVSimulator(n, use_float32=True)
    circ = workloads.random_layers_circuit(n, 3, seed=5)
    sim.run_circuit(circ)
    before = sim.get_statevector().astype(np.complex128)
    view = sim.sv
    out = desv.NoiseModel.apply_to_sv(view, n, model, 0.3, rng=np.random.default_rng(8))
    assert out is view and out.handle is sim.handle
    got = sim.get_statevector()
    want = O.noise_apply_to_sv(before, n, model, 0.3, np.random.default_rng(8))
    assert np.max(np.abs(got - want)) < 1e-5
    sim.close()


@pytest.mark.parametrize("complex128", [True, False])
def test_apply_1q_with_next_populations(complex128):
    """One pass applies a 2x2 operator in place and returns the populations
    of another qubit (or the same one) in the updated state."""
    n = 12
    rng = np.random.default_rng(11)
    h = desv.StateHandle(n, complex128)
    sv = _random_state(n, 4)
    # index bits 8, 11, 0, 6, 3, 1, 2: the complex64 kernel's pair, strided
    # and non-temporal forms
    for q, nq in ((3, 7), (0, 11), (11, 0), (5, 5), (8, None), (10, 1), (9, 10)):
        m = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
        h.upload(sv)
        got = h.apply_1q_probs(m, q, nq)
        want = O.apply_matrix(sv, m, [q])
        tol = 1e-12 if complex128 else 1e-5
        assert np.max(np.abs(h.download() - want)) < tol
        if nq is None:
            assert got is None
        else:
            bits = (np.arange(1 << n) >> (n - 1 - nq)) & 1
            pw = np.abs(want) ** 2
            assert abs(got[0] - pw[bits == 0].sum()) < tol and abs(got[1] - pw[bits == 1].sum()) < tol
    h.close()


def _monomial(rng, kind):
    a, b = rng.normal(size=2) + 1j * rng.normal(size=2)
    return np.array([[[a, 0], [0, b]], [[0, a], [0, 0]], [[0, a], [b, 0]], [[0, 0], [0, a]],
                     [[a, 0], [0, 0]], [[0, 0], [a, 0]]][kind], dtype=np.complex128)


@pytest.mark.parametrize("complex128", [True, False])
@pytest.mark.parametrize("n", [1, 3, 7, 12, 17])
def test_monomial_layer_and_populations(complex128, n):
    """One pass applies operators with one nonzero entry per row on several
    qubits and returns the populations of the result over up to 12 qubits
    (lane-local and run-uniform index bits, the flipped partner's bins)."""
    rng = np.random.default_rng(100 + n)
    h = desv.StateHandle(n, complex128)
    tol = 1e-11 if complex128 else 2e-5
    for trial in range(6):
        sv = _random_state(n, trial)
        k = int(rng.integers(0, n + 1)) if trial else n
        qs = [int(q) for q in rng.permutation(n)[:k]]
        ops = {q: _monomial(rng, int(rng.integers(6))) for q in qs}
        npop = int(rng.integers(0, min(n, 12) + 1)) if trial != 1 else min(n, 12)
        pq = [int(q) for q in rng.permutation(n)[:npop]]
        h.upload(sv)
        pops = h.apply_layer(ops, pq)
        want = sv.copy()
        for q, m in ops.items():
            want = O.apply_matrix(want, m, [q])
        assert np.max(np.abs(h.download() - want)) < tol * max(1.0, np.max(np.abs(want))), (n, trial)
        key = np.zeros(1 << n, dtype=np.int64)
        for j, q in enumerate(pq):
            key |= ((np.arange(1 << n) >> (n - 1 - q)) & 1) << j
        pw = np.bincount(key, weights=np.abs(want) ** 2, minlength=1 << npop)
        assert np.max(np.abs(pops - pw)) < tol * max(1.0, pw.sum()), (n, trial)
        # populations only: the state is not written
        h.upload(sv)
        p0 = h.apply_layer({}, pq)
        assert np.array_equal(h.download(), sv.astype(h.download().dtype))
        assert abs(p0.sum() - 1.0) < tol
    with pytest.raises(ValueError, match="two nonzero entries"):
        h.apply_layer({0: np.ones((2, 2))}, [])
    with pytest.raises(ValueError, match="both rows"):
        h.apply_layer({0: np.array([[1, 0], [1, 0]])}, [])
    with pytest.raises(ValueError, match="at most 12"):
        h.apply_layer({}, list(range(min(n, 13))) + [0] * max(0, 13 - n))
    h.close()


@pytest.mark.parametrize("model", ["amplitude_damping", "combined"])
@pytest.mark.parametrize("complex128", [True, False])
def test_damping_groups_with_repeated_qubits(model, complex128):
    """More than ten distinct qubits (two device passes) and qubits listed
    twice, inside a group and across the group boundary."""
    n = 13
    sv = _random_state(n, 21)
    qubits = [0, 3, 0, 5, 1, 2, 4, 6, 7, 8, 9, 3, 10, 11, 12, 10, 1]
    for seed in range(3):
        h = desv.StateHandle(n, complex128)
        h.upload(sv)
        desv.NoiseModel.apply_to_sv(h, n, model, 0.6, rng=np.random.default_rng(seed), qubits=qubits)
        want = O.noise_apply_to_sv(sv, n, model, 0.6, np.random.default_rng(seed), qubits)
        tol = 1e-12 if complex128 else 1e-5
        assert np.max(np.abs(h.download() - want)) < tol, (model, seed)
        h.close()


def test_sequence_of_trajectories_advances_the_generator():
    n = 5
    sv = _random_state(n, 9)
    g_dev, g_ref = np.random.default_rng(42), np.random.default_rng(42)
    for _ in range(5):
        got = desv.NoiseModel.apply_to_sv(sv, n, 'combined', 0.5, rng=g_dev)
        want = O.noise_apply_to_sv(sv, n, 'combined', 0.5, g_ref)
        assert np.max(np.abs(got - want)) < 1e-12


def _vqe(n=6, seed=0):
    vc, theta, terms = workloads.vqe_instance(n, 2, 12, seed=seed)
    return vc, theta, terms


@pytest.mark.parametrize("model", MODELS)
def test_energy_fn_with_noise_matches_oracle(model):
    n = 6
    vc, theta, terms = _vqe(n)
    fn, _ = desv.circuit_to_energy_fn(vc, n)
    H = desv.PauliSumOperator(terms, n)
    spec = desv.NoiseSpec(model, 0.35, rng=np.random.default_rng(17))
    e, sv = fn(theta, H, noise=spec)
    template = O.build_template(vc.ops, n)
    _, psi = O.energy(O.patch_template(template, theta), terms, n)
    noisy = O.noise_apply_to_sv(psi, n, model, 0.35, np.random.default_rng(17))
    want = float(np.real(np.vdot(noisy, O.pauli_sum_matvec(noisy, terms))))
    assert abs(e - want) < 1e-10
    assert np.max(np.abs(np.asarray(sv) - noisy)) < 1e-12
    # dense Hamiltonian, same trajectory
    Hd = O.pauli_matrix(terms, n)
    e2, _ = fn(theta, Hd, noise=desv.NoiseSpec(model, 0.35, rng=np.random.default_rng(17)))
    assert abs(e2 - want) < 1e-10


@pytest.mark.parametrize("model", ["depolarizing", "bitflip", "phaseflip"])
def test_gradient_through_pauli_trajectory(model):
    """The gradient with a fixed trajectory equals the oracle's adjoint
    gradient of the circuit followed by the drawn Paulis, and a central
    finite difference of the noisy energy with the same draws."""
    n = 6
    vc, theta, terms = _vqe(n, seed=2)
    fn, _ = desv.circuit_to_energy_fn(vc, n)
    H = desv.PauliSumOperator(terms, n)
    e, g = fn.value_and_grad(theta, H, noise=desv.NoiseSpec(model, 0.5, rng=np.random.default_rng(3)))
    e_val, _ = fn(theta, H, noise=desv.NoiseSpec(model, 0.5, rng=np.random.default_rng(3)))
    assert abs(e - e_val) < 1e-12
    from desv.noise import pauli_trajectory
    tail = pauli_trajectory(n, model, 0.5, np.random.default_rng(3))
    assert len(tail) > 0
    template = np.concatenate([O.build_template(vc.ops, n), tail])
    e_ref, g_ref = O.adjoint_gradient(template, theta, terms, n)
    assert abs(e - e_ref) < 1e-12
    assert np.max(np.abs(g - g_ref)) < 1e-10
    g_fd = O.fd_gradient(template, theta, terms, n)
    assert np.max(np.abs(g - g_fd)) < 1e-6


def test_torch_backward_with_noise():
    torch = pytest.importorskip("torch")
    n = 5
    vc, theta, terms = _vqe(n, seed=4)
    fn, _ = desv.circuit_to_energy_fn(vc, n)
    H = desv.PauliSumOperator(terms, n)
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    e, _ = fn(th, H, noise=desv.NoiseSpec('depolarizing', 0.6, rng=np.random.default_rng(1)))
    e.backward()
    _, g = fn.value_and_grad(theta, H, noise=desv.NoiseSpec('depolarizing', 0.6, rng=np.random.default_rng(1)))
    assert np.max(np.abs(th.grad.numpy() - g)) < 1e-12
    with pytest.raises(NotImplementedError):
        fn.value_and_grad(theta, H, noise=desv.NoiseSpec('amplitude_damping', 0.2, rng=np.random.default_rng(1)))
