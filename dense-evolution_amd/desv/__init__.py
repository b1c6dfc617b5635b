"""desv -- MI355X-native dense statevector engine.

Drop-in surface of Dense-Evolution's statevector hot path
(``dense_evolution/__init__.py:78-131`` names): ``DenseSVSimulator``,
``QASMParser``/``QASMCircuit``, ``QuantumTranspiler``, the gate tables, the
Pauli observables, ``circuit_to_energy_fn`` and the statevector noise
trajectories (``NoiseModel``, ``NoiseSpec``).  The compute runs in
hand-written HIP kernels (``libdesv.so``) on the GPU; importing this package
fails if the library is not built.
"""
from ._capi import device_count
from .autodiff import AdjointProgram, circuit_to_energy_fn, parameter_shift_grad
from .circuits import entangling_layer, pauli_rotation_ops, random_circuit, trotter_evolve_ops
from .config import ensure_x64, set_compensation, set_precision, with_x64
from .gates import GATE_IDS, GATES, PARAMETRIC_GATES
from .measurement import sample_counts, statevector_fidelity
from .noise import NoiseModel, NoiseSpec
from . import lanczos
from .lanczos import ground_state_energy
from .observables import (PauliSumOperator, multiply_pauli_terms, pauli_expectation,
                          pauli_hamiltonian_to_matrix, pauli_sum_expectation,
                          pauli_sum_expectation_jax, pauli_sum_matvec, pauli_sum_matvec_jax)
from .program import Program
from . import tune, workloads
from .qasm import QASMCircuit, QASMParser
from .simulator import DenseSVSimulator, DeviceState, StateHandle
from .transpiler import QuantumTranspiler

__version__ = "0.1.0"

__all__ = [
    "__version__", "set_precision", "set_compensation", "ensure_x64", "with_x64",
    "DenseSVSimulator", "DeviceState", "StateHandle", "Program",
    "QASMParser", "QASMCircuit", "QuantumTranspiler",
    "GATES", "PARAMETRIC_GATES", "GATE_IDS",
    "entangling_layer", "pauli_rotation_ops", "trotter_evolve_ops", "random_circuit",
    "pauli_expectation", "pauli_sum_expectation", "pauli_hamiltonian_to_matrix", "pauli_sum_matvec",
    "multiply_pauli_terms", "pauli_sum_matvec_jax", "pauli_sum_expectation_jax", "PauliSumOperator",
    "circuit_to_energy_fn", "AdjointProgram", "parameter_shift_grad",
    "sample_counts", "statevector_fidelity", "device_count", "NoiseModel", "NoiseSpec",
    "lanczos", "ground_state_energy",
]
