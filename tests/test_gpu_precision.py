"""Precision study at 20 qubits against the complex128 oracle: the XXZ chain
(J = 1, Delta = 0.5, open boundary), first-order Trotter, dt = 0.05, 200
steps (79,800 gates after the random product-state preparation).

The oracle evolves the state bond by bond with the exact 4x4 product of each
bond's 21 Trotter gates (built from the oracle's own gate tables, checked
below against gate-by-gate application), which keeps the 20-qubit reference
to about 3,800 two-qubit applications on the CPU.

Bars: complex128 state fidelity >= 1 - 1e-12 and <Z_i Z_i+1> within 1e-10;
complex64 (FP32 gate arithmetic with compensated matrices, the simulator's
default; FP64 norm and expectation accumulation) fidelity |<ref|psi>|^2 (the
reference's statevector_fidelity, unnormalised) >= 1 - 1e-6, |norm - 1| <=
1e-6 and <Z_i Z_i+1> within 1e-6.  Plain complex64 (each fused matrix rounded
once to FP32, as the reference's complex64 path rounds its gate matrices)
loses about 3e-5 of norm over these 79,800 gates, because the same rounded
matrices act every step and their rounding error adds up coherently; it is
kept as a strict expected failure of the same bars.
"""
import numpy as np
import pytest

import desv
from desv import workloads
from oracle import sv_oracle as O

pytestmark = pytest.mark.gpu

DT, STEPS = 0.05, 200


def bond_unitary():
    """4x4 product of one bond's XX, YY, ZZ rotation gates (qubits 0, 1)."""
    ops = []
    for c, d in [(1.0, {0: 'X', 1: 'X'}), (1.0, {0: 'Y', 1: 'Y'}), (0.5, {0: 'Z', 1: 'Z'})]:
        ops += O.pauli_rotation_ops(d, c * DT)
    table = O.encode_ops(ops, 2)
    U = np.zeros((4, 4), dtype=np.complex128)
    for j in range(4):
        e = np.zeros(4, dtype=np.complex128)
        e[j] = 1
        U[:, j] = O.run_ops(e, table)
    return U


def oracle_state(n, steps=STEPS, seed=0):
    circ = workloads.xxz_trotter_circuit(n, DT, steps, seed=seed)
    prep = circ[:2 * n]
    sv = O.run_ops_fast(O.zero_state(n), O.encode_ops(prep, n))
    U = bond_unitary()
    for _ in range(steps):
        for i in range(n - 1):
            sv = O.apply_matrix(sv, U, [i, i + 1])
    return sv


def test_bond_fused_oracle_matches_gate_by_gate():
    n, steps = 6, 7
    circ = workloads.xxz_trotter_circuit(n, DT, steps, seed=3)
    want = O.run_ops(O.zero_state(n), O.encode_ops(circ, n))
    assert O.fidelity(oracle_state(n, steps, seed=3), want) > 1 - 1e-13


def zz(sv_like, n):
    return np.array([desv.pauli_expectation(sv_like, {i: 'Z', i + 1: 'Z'}) for i in range(n - 1)])


_STATES = {}


def run_xxz(n, f32, compensated):
    key = (n, f32, compensated)
    if key not in _STATES:
        circ = workloads.xxz_trotter_circuit(n, DT, STEPS, seed=0)
        sim = desv.DenseSVSimulator(n, use_float32=f32, compensated=compensated)
        sim.run_circuit(circ)
        _STATES[key] = (sim.get_statevector(), sim.handle.norm2(), zz(sim.sv, n))
        sim.close()
    return _STATES[key]


_REF = {}


def reference(n):
    if n not in _REF:
        ref = oracle_state(n)
        _REF[n] = (ref, np.array([O.pauli_expectation(ref, {i: 'Z', i + 1: 'Z'}) for i in range(n - 1)]))
    return _REF[n]


def errors(n, f32, compensated):
    ref, zz_ref = reference(n)
    sv, nrm, zzs = run_xxz(n, f32, compensated)
    fd = O.fidelity(sv.astype(np.complex128), ref)
    out = (fd, fd / nrm, float(np.max(np.abs(zzs - zz_ref))), nrm - 1.0)
    name = 'c128' if not f32 else ('c64 compensated' if compensated else 'c64 plain')
    print(f"20q XXZ {name} vs c128 oracle: 1 - fidelity = {1 - out[0]:.3e}, normalised {1 - out[1]:.3e}, "
          f"max |d<ZZ>| = {out[2]:.3e}, norm - 1 = {out[3]:.3e}")
    return out


def test_xxz_20q_complex128_against_oracle():
    fd, fdn, dzz, dn = errors(20, False, False)
    assert fd >= 1 - 1e-12 and dzz < 1e-10 and abs(dn) < 1e-10


def test_xxz_20q_complex64_compensated_against_oracle():
    fd, fdn, dzz, dn = errors(20, True, True)
    assert fd >= 1 - 1e-6 and abs(dn) <= 1e-6 and dzz <= 1e-6


@pytest.mark.xfail(strict=True, reason="plain complex64 (matrices rounded once to FP32) loses about 3e-5 of norm "
                                       "over 200 Trotter steps: the rounding of the repeated matrices adds up "
                                       "coherently; the compensated default meets the bars")
def test_xxz_20q_complex64_plain_against_oracle():
    fd, fdn, dzz, dn = errors(20, True, False)
    assert fdn >= 1 - 1e-6  # the direction stays accurate; the norm does not
    assert fd >= 1 - 1e-6 and abs(dn) <= 1e-6 and dzz <= 1e-6
