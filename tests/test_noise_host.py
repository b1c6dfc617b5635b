"""Host logic of the noise trajectories (no GPU): the Pauli-channel
trajectory drawn as op rows consumes the generator exactly like the
oracle's ``noise_apply_to_sv`` and, applied by the oracle, gives the same
state; spec and argument handling."""
import numpy as np
import pytest

import desv
from desv.noise import pauli_trajectory
from oracle import sv_oracle as O


def _random_state(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("model", ["depolarizing", "bitflip", "phaseflip"])
@pytest.mark.parametrize("qubits", [None, [3, 0, 5], [1]])
def test_pauli_trajectory_rows_match_oracle(model, qubits):
    n = 6
    sv = _random_state(n, 1)
    for seed in range(6):
        rows = pauli_trajectory(n, model, 0.45, np.random.default_rng(seed), qubits)
        got = O.run_ops(sv.copy(), rows)
        got = got / (np.linalg.norm(got) + 1e-15)
        want = O.noise_apply_to_sv(sv, n, model, 0.45, np.random.default_rng(seed), qubits)
        assert np.allclose(got, want, atol=1e-14)
        # the generator is left in the same position
        ga, gb = np.random.default_rng(seed), np.random.default_rng(seed)
        pauli_trajectory(n, model, 0.45, ga, qubits)
        O.noise_apply_to_sv(sv, n, model, 0.45, gb, qubits)
        assert ga.random() == gb.random()


def test_pauli_trajectory_rejects_non_pauli_and_bad_args():
    rng = np.random.default_rng(0)
    with pytest.raises(NotImplementedError):
        pauli_trajectory(3, "amplitude_damping", 0.1, rng)
    with pytest.raises(ValueError):
        pauli_trajectory(3, "nonsense", 0.1, rng)
    with pytest.raises(ValueError):
        pauli_trajectory(3, "bitflip", 0.1, rng, [3])


def test_noise_spec_and_description():
    spec = desv.NoiseSpec(model="depolarizing", p=0.1, jax_key=np.array([0, 7], dtype=np.uint32),
                          qubits=[0, 1])
    assert repr(spec) == "NoiseSpec(model='depolarizing', p=0.1, qubits=(0, 1))"
    # a key reproduces its draws; an explicit generator is used as given
    assert spec.generator().random() == spec.generator().random()
    g = np.random.default_rng(4)
    assert desv.NoiseSpec("bitflip", 0.2, rng=g).generator() is g
    for model in desv.NoiseModel.MODELS:
        assert "kraus" in desv.NoiseModel.kraus_description(model)


def test_ideal_and_zero_probability_return_the_input():
    sv = _random_state(3, 2)
    assert desv.NoiseModel.apply_to_sv(sv, 3, 'ideal', 0.5) is sv
    assert desv.NoiseModel.apply_to_sv(sv, 3, 'depolarizing', 0.0) is sv
    with pytest.raises(ValueError):
        desv.NoiseModel.apply_to_sv(sv, 3, 'thermal', 0.1)


class _NumpyLayers:
    """Stand-in for a device handle: applies a layer with the oracle's
    matrix routine and returns populations like ``StateHandle.apply_layer``,
    so the trajectory's host logic (groups, per-draw population updates,
    operator composition) is checked on the CPU."""

    def __init__(self, sv, n):
        self.sv, self.n, self.passes = sv.copy(), n, 0

    def apply_layer(self, ops, pop_qubits):
        self.passes += 1
        for q, m in ops.items():
            self.sv = O.apply_matrix(self.sv, m, [q])
        key = np.zeros(1 << self.n, dtype=np.int64)
        for j, q in enumerate(pop_qubits):
            key |= ((np.arange(1 << self.n) >> (self.n - 1 - q)) & 1) << j
        return np.bincount(key, weights=np.abs(self.sv) ** 2, minlength=1 << len(pop_qubits))


@pytest.mark.parametrize("model", ["amplitude_damping", "combined", "depolarizing"])
@pytest.mark.parametrize("qubits", [None, [0, 3, 0, 5, 1, 2, 4, 6, 7, 8, 9, 3, 10, 11, 12, 10, 1], []])
def test_grouped_trajectory_host_logic_matches_oracle(model, qubits):
    from desv.noise import _run_trajectory
    n = 13
    sv = _random_state(n, 7)
    qs = list(range(n)) if qubits is None else qubits
    for seed in range(3):
        fake = _NumpyLayers(sv, n)
        norm2 = _run_trajectory(fake, [n - 1 - q for q in qs], model, 0.6, np.random.default_rng(seed))
        got = fake.sv / (np.sqrt(norm2) + 1e-15)
        want = O.noise_apply_to_sv(sv, n, model, 0.6, np.random.default_rng(seed), qs)
        assert np.max(np.abs(got - want)) < 1e-12, (model, seed)
        # damping: a populations pass, then one layer per group of at most 10 qubits
        distinct = len(set(qs))
        expect = 1 if (model == "depolarizing" or not qs) else 1 + -(-distinct // 10)
        assert fake.passes == expect
