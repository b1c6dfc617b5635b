"""GPU parity of the gate engine and observables against the CPU oracle.

Bars: complex128 results within 1e-10 of the oracle; complex64 results at
state fidelity >= 1 - 1e-6 against the complex128 oracle and expectations
within 1e-6 (the precision policy: FP32 gate arithmetic, FP64 reductions).
"""
import json
import os

import numpy as np
import pytest

import desv
from desv import workloads
from oracle import sv_oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def rand_state(n, rng):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


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


def fid(a, b):
    return O.fidelity(a, b)


@pytest.mark.parametrize("use_float32", [False, True])
def test_known_answers(use_float32):
    with open(os.path.join(GOLD, "known_answers.json")) as f:
        cases = json.load(f)
    for case in cases:
        sim = desv.DenseSVSimulator(case["n"], use_float32=use_float32)
        sim.run_circuit([tuple(c) for c in case["circuit"]])
        p = sim.get_probabilities()
        for idx, val in case.get("probs", {}).items():
            assert abs(p[int(idx)] - val) < 1e-6, case["name"]
        for term, val in case.get("expect", {}).items():
            assert abs(desv.pauli_expectation(sim.sv, term) - val) < 1e-6, case["name"]


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 12, 15])
def test_random_op_tables_c128(n, rng):
    ops = random_ops(n, 200, rng)
    sv0 = rand_state(n, rng)
    sim = desv.DenseSVSimulator(n)
    sim.sv = sv0
    sim.handle.run_ops(ops)
    want = O.run_ops_fast(sv0.astype(np.complex128), ops)
    assert np.allclose(sim.get_statevector(), want, atol=1e-10)


@pytest.mark.parametrize("n", [3, 9, 14])
def test_random_op_tables_c64(n, rng):
    ops = random_ops(n, 200, rng)
    sv0 = rand_state(n, rng)
    sim = desv.DenseSVSimulator(n, use_float32=True)
    sim.sv = sv0.astype(np.complex64)
    sim.handle.run_ops(ops)
    want = O.run_ops_fast(sv0, ops)
    assert fid(sim.get_statevector(), want) >= 1 - 1e-6


@pytest.mark.parametrize("opts", [
    dict(fuse_bits=1, tile_bits=6, low_bits=2),
    dict(fuse_bits=2, tile_bits=7, low_bits=3),
    dict(fuse_bits=3, tile_bits=8, low_bits=4),
    dict(fuse_bits=5, tile_bits=9, low_bits=2),
    dict(fuse_bits=2, tile_bits=8, low_bits=5, per_gate=True),
])
@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_programs_multi_stage(opts, dtype, rng):
    n = 13
    ops = random_ops(n, 300, rng)
    sv0 = rand_state(n, rng)
    prog = desv.Program(n).add_ops(ops).compile(dtype, **opts)
    assert prog.stats()["passes"] > 1
    h = desv.StateHandle(n, dtype == "complex128")
    h.upload(sv0)
    prog.run(h)
    want = O.run_ops_fast(sv0, ops)
    got = h.download()
    if dtype == "complex128":
        assert np.allclose(got, want, atol=1e-10)
    else:
        assert fid(got, want) >= 1 - 1e-6


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_streaming_passes_double_buffer(dtype, rng):
    """Per-gate programs run every gate on the streaming kernel, out of place
    into the second buffer (controlled gates copy the untouched half) or in
    place; both match the oracle, also after toggling between the modes and
    with an odd number of buffer swaps."""
    n = 14
    ops = random_ops(n, 121, rng)
    sv0 = rand_state(n, rng)
    prog = desv.Program(n).add_ops(ops).compile(dtype, fuse_bits=1, per_gate=True)
    want = O.run_ops_fast(sv0, ops)
    want2 = O.run_ops_fast(want, ops)
    for mode in (True, False):
        h = desv.StateHandle(n, dtype == "complex128")
        h.set_double_buffer(mode)
        h.upload(sv0)
        prog.run(h)
        got = h.download()
        h.set_double_buffer(not mode)
        prog.run(h)
        got2 = h.download()
        nrm = h.norm2()
        h.close()
        if dtype == "complex128":
            assert np.allclose(got, want, atol=1e-10)
            assert np.allclose(got2, want2, atol=1e-10)
        else:
            assert fid(got, want) >= 1 - 1e-6
            assert fid(got2, want2) >= 1 - 1e-6
        assert abs(nrm - 1) < 1e-5


@pytest.mark.parametrize("n", [9, 12])
@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_streaming_low_index_bits(n, dtype, rng):
    """Every gate id on the lowest index bits on the streaming kernels: bit 0
    (the last qubit) as target, as control and as one of two targets
    (complex64 keeps the amplitude pair (x, x | 1) in one 16-byte lane), and
    targets in bits 1 and 2 (complex64 exchanges group members between lanes),
    in place and out of place, plus dense 4x4 unitaries on such pairs;
    amplitudes within 1e-10 (complex128) or 2e-6 (complex64) of the oracle."""
    last = n - 1
    rows = [[g, q, 0, rng.uniform(-np.pi, np.pi)] for q in (last, n - 2, n - 3) for g in range(15)]
    pairs = ((last, 3), (2, last), (last, 0), (0, last), (last, n - 2), (n - 2, last), (n - 2, n - 3),
             (n - 3, n - 2), (n - 2, n - 5), (n - 7, n - 3), (n - 3, 0), (n - 2, 1))
    for g in (20, 21, 22, 23, 24, 25):
        for a, b in pairs:
            rows.append([g, a, b, rng.uniform(-np.pi, np.pi)])
    ops = np.array(rows, dtype=np.float64)
    sv0 = rand_state(n, rng)
    prog = desv.Program(n).add_ops(ops).compile(dtype, fuse_bits=1, per_gate=True)
    want = O.run_ops_fast(sv0, ops)
    tol = 1e-10 if dtype == "complex128" else 2e-6
    for mode in (True, False):
        h = desv.StateHandle(n, dtype == "complex128")
        h.set_double_buffer(mode)
        h.upload(sv0)
        prog.run(h)
        got = h.download()
        h.close()
        assert np.max(np.abs(got - want)) < tol, mode
    sim = desv.DenseSVSimulator(n, use_float32=dtype == "complex64")
    sim.sv = sv0.astype(sim.dtype)
    want = sv0.copy()
    for a, b in ((last, 3), (1, last), (last, n - 2), (0, last), (n - 2, n - 3), (n - 5, n - 2)):
        z = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
        u, _ = np.linalg.qr(z)
        sim.apply_gate_2q(u, a, b)
        want = O.apply_matrix(want, u, [a, b])
        v, _ = np.linalg.qr(rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2)))
        sim.apply_gate_1q(v, b)
        want = O.apply_matrix(want, v, [b])
    assert np.max(np.abs(sim.get_statevector() - want)) < tol
    sim.close()


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 9, 10])
def test_streaming_small_widths_c64(n, rng):
    """One streaming pass per gate on complex64 states of 1..10 qubits (grids
    of a single partly used workgroup), in place and out of place: every
    gate id on every qubit and qubit pair against the oracle."""
    rows = [[g, q, 0, rng.uniform(-np.pi, np.pi)] for q in range(n) for g in range(1, 15)]
    for a in range(n):
        for b in range(n):
            if a != b:
                rows.append([int(rng.choice([20, 21, 22, 23, 24, 25])), a, b, rng.uniform(-np.pi, np.pi)])
    ops = np.array(rows, dtype=np.float64)
    sv0 = rand_state(n, rng)
    prog = desv.Program(n).add_ops(ops).compile("complex64", fuse_bits=1, per_gate=True)
    want = O.run_ops_fast(sv0, ops)
    for mode in (True, False):
        h = desv.StateHandle(n, False)
        h.set_double_buffer(mode)
        h.upload(sv0)
        prog.run(h)
        got = h.download()
        h.close()
        assert np.max(np.abs(got - want)) < 5e-6, mode


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("use_float32", [False, True])
def test_apply_arbitrary_unitaries(k, use_float32, rng):
    n = 11
    sv = rand_state(n, rng)
    sim = desv.DenseSVSimulator(n, use_float32=use_float32)
    sim.sv = sv.astype(sim.dtype)
    want = sv.copy()
    for _ in range(6):
        q = [int(x) for x in rng.choice(n, k, replace=False)]
        z = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
        u, _ = np.linalg.qr(z)
        if k == 1:
            sim.apply_gate_1q(u, q[0])
        elif k == 2:
            sim.apply_gate_2q(u, q[0], q[1])
        else:
            sim.apply_gate(u, q)
        want = O.apply_matrix(want, u, q)
    got = sim.get_statevector()
    if use_float32:
        assert fid(got, want) >= 1 - 1e-6
    else:
        assert np.allclose(got, want, atol=1e-10)


@pytest.mark.parametrize("fuse,mfma", [(2, False), (4, False), (5, False), (4, True), (5, True)])
@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_brickwork_su4_small(fuse, mfma, dtype):
    n, depth = 12, 8
    gates = workloads.brickwork_su4(n, depth, seed=5)
    prog = desv.Program(n)
    want = O.zero_state(n)
    for u, a, b in gates:
        prog.add_matrix(u, [a, b])
        want = O.apply_matrix(want, u, [a, b])
    if mfma and dtype == "complex128":
        # MFMA blocks are complex64 only; asking for them in complex128 is an error
        with pytest.raises(ValueError, match="complex64"):
            prog.compile(dtype, fuse_bits=fuse, use_mfma=True)
        return
    prog.compile(dtype, fuse_bits=fuse, use_mfma=mfma)
    if mfma:
        assert prog.stats()["mfma_blocks"] > 0
    h = desv.StateHandle(n, dtype == "complex128")
    prog.run(h)
    got = h.download()
    if dtype == "complex128":
        assert np.allclose(got, want, atol=1e-10)
    else:
        assert fid(got, want) >= 1 - 1e-6


def test_config_plain_20q_c64_parity():
    """20-qubit random circuit, 20 layers of rx/ry/rz + cx ladder, complex64,
    vs the complex128 oracle: fidelity and every <Z_i>."""
    n = 20
    circ = workloads.random_layers_circuit(n, 20, seed=0)
    sim = desv.DenseSVSimulator(n, use_float32=True)
    sim.run_circuit(circ)
    want = O.run_ops_fast(O.zero_state(n), O.encode_ops(circ, n))
    got = sim.get_statevector()
    assert got.dtype == np.complex64
    assert fid(got, want) >= 1 - 1e-6
    zs = [desv.pauli_expectation(sim.sv, {q: 'Z'}) for q in range(n)]
    want_z = [O.pauli_expectation(want, {q: 'Z'}) for q in range(n)]
    assert np.max(np.abs(np.array(zs) - want_z)) < 1e-6


@pytest.mark.parametrize("use_float32", [False, True])
def test_pauli_expectations_and_matvec(use_float32, rng):
    n = 10
    sv = rand_state(n, rng)
    terms = workloads.random_pauli_hamiltonian(rng, n, 60)
    terms += [(0.3, {}), (1.2, {0: 'Y'}), (-0.4, {n - 1: 'X', 0: 'Y'})]
    dt = np.complex64 if use_float32 else np.complex128
    sim = desv.DenseSVSimulator(n, use_float32=use_float32)
    sim.sv = sv.astype(dt)
    ref_sv = sv.astype(dt).astype(np.complex128)
    tol = 1e-6 if use_float32 else 1e-11
    for c, t in terms[:20]:
        assert abs(desv.pauli_expectation(sim.sv, t) - O.pauli_expectation(ref_sv, t)) < tol
    e = desv.pauli_sum_expectation(sim.sv, terms)
    assert abs(e - O.pauli_sum_expectation(ref_sv, terms)) < 10 * tol
    mv = desv.pauli_sum_matvec(sim.sv, terms)
    assert np.allclose(mv, O.pauli_sum_matvec(ref_sv, terms), atol=10 * tol)
    op = desv.PauliSumOperator(terms, n)
    assert np.allclose(op @ ref_sv, O.pauli_sum_matvec(ref_sv, terms), atol=1e-10)
    # host array input (uploaded) gives the same value
    assert abs(desv.pauli_expectation(ref_sv, 'XY' + 'I' * (n - 2)) -
               O.pauli_expectation(ref_sv, 'XY' + 'I' * (n - 2))) < 1e-11


@pytest.mark.parametrize("n", [12, 14, 17])
@pytest.mark.parametrize("use_float32", [False, True])
def test_pauli_matvec_across_tiles(n, use_float32, rng):
    """H psi on states wider than one 2^11-amplitude matvec tile: groups whose
    flip masks share their bits above the tile read one staged input tile,
    masks straddling the tile boundary and high-bit signs included;
    complex128 within 1e-10, complex64 within 1e-5 of the oracle."""
    sv = rand_state(n, rng)
    terms = workloads.random_pauli_hamiltonian(rng, n, 80)
    terms += [(0.7, {0: 'X'}), (0.2 - 0.1j, {0: 'Y', 1: 'X', n - 1: 'Z'}), (-0.5, {n - 12: 'Y', n - 11: 'Y'}),
              (0.9, {2: 'Z', n - 3: 'X'}), (0.3, {}), (1.1, {0: 'Z', n - 1: 'Z'})]
    dt = np.complex64 if use_float32 else np.complex128
    sim = desv.DenseSVSimulator(n, use_float32=use_float32)
    sim.sv = sv.astype(dt)
    ref_sv = sv.astype(dt).astype(np.complex128)
    want = O.pauli_sum_matvec(ref_sv, terms)
    got = desv.pauli_sum_matvec(sim.sv, terms)
    assert np.max(np.abs(got - want)) < (1e-5 if use_float32 else 1e-10)
    # the real-coefficient sum takes the one-pass energy path
    real_terms = [(float(np.real(c)), t) for c, t in terms]
    e = desv.pauli_sum_expectation(sim.sv, real_terms)
    assert abs(e - O.pauli_sum_expectation(ref_sv, real_terms)) < (1e-5 if use_float32 else 1e-10)
    sim.close()


def test_observable_validation():
    sim = desv.DenseSVSimulator(3)
    with pytest.raises(ValueError):
        desv.pauli_expectation(sim.sv, 'XZ', n_qubits=3)
    with pytest.raises(ValueError):
        desv.pauli_expectation(sim.sv, {5: 'X'})
    with pytest.raises(ValueError):
        desv.pauli_expectation(sim.sv, {0: 'Q'})


def test_readout_measure_and_normalisation(rng):
    n = 6
    sim = desv.DenseSVSimulator(n)
    sv = rand_state(n, rng)
    sim.set_initial_state(3.0 * sv)
    assert np.allclose(sim.get_statevector(), sv, atol=1e-12)
    assert np.allclose(sim.get_probabilities(), O.probabilities(sv), atol=1e-12)
    r = sim.measure(2, rng=np.random.default_rng(0))
    after = sim.get_statevector()
    assert abs(np.linalg.norm(after) - 1) < 1e-12
    bit = n - 1 - 2
    idx = np.arange(1 << n)
    assert np.all(np.abs(after[((idx >> bit) & 1) != r]) == 0)
    with pytest.raises(ValueError):
        sim.set_initial_state(np.zeros(1 << n))
    with pytest.raises(ValueError):
        sim.set_initial_state(np.ones(3))
    assert sim.memory_mb() == (1 << n) * 16 / 1e6


@pytest.mark.parametrize("n", [1, 2, 7, 13, 20])
@pytest.mark.parametrize("complex128", [False, True])
def test_norm_and_qubit_populations(n, complex128, rng):
    """FP64 norm and per-qubit populations of a state (complex64: 16-byte
    loads, a partial last wave below 2^9 amplitudes) against NumPy sums of the
    same amplitudes."""
    sv = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(np.complex128 if complex128 else np.complex64)
    h = desv.StateHandle(n, complex128)
    h.upload(sv)
    p = np.abs(sv.astype(np.complex128)) ** 2
    assert abs(h.norm2() - p.sum()) < 1e-12 * p.sum()
    idx = np.arange(1 << n)
    for q in sorted({0, n // 2, n - 1}):
        bits = (idx >> (n - 1 - q)) & 1
        p0, p1 = h.qubit_probs(q)
        assert abs(p0 - p[bits == 0].sum()) < 1e-12 * p.sum() and abs(p1 - p[bits == 1].sum()) < 1e-12 * p.sum()
    # <sv|other> in FP64 from the stored amplitudes
    other = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(sv.dtype)
    g = desv.StateHandle(n, complex128)
    g.upload(other)
    want = np.vdot(sv.astype(np.complex128), other.astype(np.complex128))
    assert abs(h.inner(g) - want) < 1e-12 * np.sqrt(p.sum() * (np.abs(other.astype(np.complex128)) ** 2).sum())
    g.close()
    h.close()


def test_errors_match_reference_messages():
    sim = desv.DenseSVSimulator(3)
    with pytest.raises(ValueError, match="unknown gate"):
        sim.run_circuit([('foo', 0)])
    with pytest.raises(ValueError, match="out of range"):
        sim.run_circuit_jit([('h', 5)])
    with pytest.raises(ValueError, match="must differ"):
        sim.apply_gate_2q(np.eye(4), 1, 1)
    with pytest.raises(ValueError, match="parameter_batch"):
        sim.run_batch_jit([('rx', 0, 0.1), ('h', 1)], np.zeros((2, 3)))


def test_eager_and_table_paths_agree(rng):
    n = 5
    circ = [('h', 0), ('u3', 1, 0.3, -0.7, 1.1), ('iswap', 0, 2), ('ecr', 3, 4), ('ccx', 0, 1, 4),
            ('u2', 2, 0.5, 0.25), ('swap', 1, 3), ('crz', 2, 0, 0.8), ('cp', 4, 1, -0.4), ('cy', 0, 3)]
    a = desv.DenseSVSimulator(n)
    a.run_circuit(circ)
    want = O.run_circuit(circ, n)
    assert np.allclose(a.get_statevector(), want, atol=1e-10)
    b = desv.DenseSVSimulator(n)
    b.run_circuit([('iswap', 0, 2)], transpile=False)  # eager matrix path
    assert np.allclose(b.get_statevector(), O.apply_matrix(O.zero_state(n), O.GATES['iswap'], [0, 2]))


def test_run_batch_jit(rng):
    n = 4
    circ = [('ry', q, 0.0) for q in range(n)] + [('cx', 0, 1), ('rz', 2, 0.0), ('cp', 1, 3, 0.0)]
    P = rng.uniform(-np.pi, np.pi, (5, n + 2))
    sim = desv.DenseSVSimulator(n)
    out = sim.run_batch_jit(circ, P)
    for b in range(5):
        c = [('ry', q, P[b, q]) for q in range(n)] + [('cx', 0, 1), ('rz', 2, P[b, n]), ('cp', 1, 3, P[b, n + 1])]
        assert np.allclose(out[b], O.run_circuit(c, n), atol=1e-10)


@pytest.mark.parametrize("use_float32", [False, True])
def test_run_batch_jit_device_batch(use_float32, rng):
    """Many rows in one device batch (reference statevector.py:496-584): every
    parametric gate consumes a column even with a literal angle, gphase from a
    transpiled u3 gets angle 0 (the reference's template), other literals are
    kept; each row against the oracle; column-count errors keep the reference's
    message."""
    n, B = 10, 24
    circ = ([('h', q) for q in range(n)] + [('rx', q, 1.5) for q in range(n)] + desv.entangling_layer(n) +
            [('u3', 3, 0.3, 0.2, 0.1), ('crz', 2, 7, 0.9), ('t', 4), ('p', 5, 2.0), ('swap', 1, 8)] +
            [('ry', q, 0.0) for q in range(n)])
    target = desv.QuantumTranspiler.transpile(circ)
    cols = sum(1 for c in target if c[0] in ('rx', 'ry', 'rz', 'p', 'u1', 'phase', 'cp', 'crz', 'cphase'))
    P = rng.uniform(-np.pi, np.pi, (B, cols))
    sim = desv.DenseSVSimulator(n, use_float32=use_float32)
    out = sim.run_batch_jit(circ, P)
    assert out.shape == (B, 1 << n) and out.dtype == (np.complex64 if use_float32 else np.complex128)
    for b in range(B):
        k, rows = 0, []
        for c in target:
            if c[0] in ('rx', 'ry', 'rz', 'p', 'u1', 'phase'):
                rows.append((c[0], c[1], P[b, k]))
                k += 1
            elif c[0] in ('cp', 'crz', 'cphase'):
                rows.append((c[0], c[1], c[2], P[b, k]))
                k += 1
            elif c[0] == 'gphase':
                rows.append(('gphase', c[1], 0.0))
            else:
                rows.append(c)
        want = O.run_ops_fast(O.zero_state(n), O.encode_ops(rows, n))
        if use_float32:
            assert O.fidelity(out[b], want) >= 1 - 1e-6
        else:
            assert np.allclose(out[b], want, atol=1e-10)
    with pytest.raises(ValueError, match="parameter_batch"):
        sim.run_batch_jit(circ, P[:, :-1])
    assert sim.run_batch_jit(circ, P[:0]).shape == (0, 1 << n)


def test_qasm_to_state_integration():
    qasm = ('OPENQASM 2.0; include "qelib1.inc"; qreg q[6]; h q[0]; cx q[0],q[1]; rx(pi/3) q[2]; '
            'ry(0.4) q[3]; cz q[2],q[3]; crz(0.7) q[1],q[4]; ccx q[0],q[4],q[5]; u3(0.1,0.2,0.3) q[5]; '
            'swap q[2],q[5];')
    circ = desv.QASMParser().parse(qasm)
    sim = desv.DenseSVSimulator(circ.n_qubits)
    sim.run_circuit(circ)
    assert np.allclose(sim.get_statevector(), O.run_circuit(list(circ), circ.n_qubits), atol=1e-10)


def test_fidelity_and_sampling():
    sim = desv.DenseSVSimulator(3)
    sim.run_circuit([('h', 0), ('cx', 0, 1), ('cx', 1, 2)])
    ghz = np.zeros(8)
    ghz[[0, 7]] = 1 / np.sqrt(2)
    assert abs(desv.statevector_fidelity(sim.sv, ghz) - 1) < 1e-12
    counts = desv.sample_counts(sim.sv, 2000, rng=np.random.default_rng(1))
    assert set(counts) == {'000', '111'} and sum(counts.values()) == 2000


def test_program_autotune_keeps_parity(rng, tmp_path, monkeypatch):
    monkeypatch.setenv("DESV_TUNE_CACHE", str(tmp_path / "tune.json"))
    n = 14
    prog = desv.Program(n)
    want = rand_state(n, rng)
    sv0 = want.copy()
    for u, a, b in workloads.brickwork_su4(n, 8, seed=5):
        prog.add_matrix(u, [a, b])
        want = O.apply_matrix(want, u, [a, b])
    cands = [dict(fuse_bits=2, tile_bits=11, use_mfma=False), dict(fuse_bits=4, tile_bits=12, use_mfma=True),
             dict(fuse_bits=5, tile_bits=12, use_mfma=True), dict(fuse_bits=3, tile_bits=10, use_mfma=False)]
    res = desv.tune.autotune_program(prog, "complex64", candidates=cands, signature="brick14")
    assert not res["cached"] and all(c["ms"] is not None for c in res["candidates"])
    h = desv.StateHandle(n, False)
    h.upload(sv0)
    prog.run(h)
    assert fid(h.download(), want) >= 1 - 1e-6
    h.close()
    assert desv.tune.autotune_program(prog, "complex64", candidates=cands, signature="brick14")["cached"]


@pytest.mark.parametrize("n,dtype", [(12, np.complex128), (22, np.complex64), (3, np.complex128), (9, np.complex64),
                                     (16, np.complex128)])
def test_device_sampling_matches_rng_choice(n, dtype, rng):
    """Shots drawn on the GPU from the caller's generator equal the
    reference's rng.choice(dim, n_shots, p=probs) with the same seed.  The
    device CDF is an exact integer prefix sum of the probabilities quantised
    to 2^-62 of the total, NumPy's a float64 cumsum: a uniform within their
    rounding of a bin edge may land in the neighbouring bin, so at most a
    couple of shots move (integer output, parity unpinned at bin edges)."""
    sv = rand_state(n, rng).astype(dtype)
    shots = 20000
    got = desv.sample_counts(sv, shots, rng=np.random.default_rng(7))
    p = np.abs(sv.astype(np.complex128)) ** 2
    p /= p.sum()
    vals, freqs = np.unique(np.random.default_rng(7).choice(1 << n, size=shots, p=p), return_counts=True)
    want = {format(int(v), f'0{n}b'): int(f) for v, f in zip(vals, freqs)}
    moved = sum(abs(got.get(k, 0) - want.get(k, 0)) for k in set(got) | set(want))
    assert sum(got.values()) == shots and moved <= 4


@pytest.mark.parametrize("scale", [0.05, 1.0, 4.0])
def test_device_sampling_of_unnormalised_states(scale, rng):
    """Shots are drawn from |a|^2 / total whatever the total: states within
    [1/16, 3.5) of unit norm take the two-level path (chunk totals, then the
    chunks that hold shots), others the full-CDF path whose fixed-point scale
    follows the total; both against rng.choice with the same uniforms."""
    n = 13
    sv = rand_state(n, rng) * np.sqrt(scale)
    shots = 20000
    got = desv.sample_counts(sv.astype(np.complex64), shots, rng=np.random.default_rng(11))
    p = np.abs(sv.astype(np.complex64).astype(np.complex128)) ** 2
    p /= p.sum()
    vals, freqs = np.unique(np.random.default_rng(11).choice(1 << n, size=shots, p=p), return_counts=True)
    want = {format(int(v), f'0{n}b'): int(f) for v, f in zip(vals, freqs)}
    moved = sum(abs(got.get(k, 0) - want.get(k, 0)) for k in set(got) | set(want))
    assert sum(got.values()) == shots and moved <= 4


def test_device_sampling_on_device_state_and_errors():
    sim = desv.DenseSVSimulator(3)
    sim.run_circuit([('h', 0), ('cx', 0, 1), ('cx', 1, 2)])
    counts = desv.sample_counts(sim.sv, 4000, rng=np.random.default_rng(3))
    assert set(counts) == {'000', '111'} and abs(counts['000'] - 2000) < 200
    with pytest.raises(ValueError, match="zero or non-finite"):
        desv.sample_counts(np.zeros(8, dtype=np.complex128), 10)
    with pytest.raises(ValueError, match="n_shots"):
        desv.sample_counts(np.ones(8, dtype=np.complex128), 0)


@pytest.mark.parametrize("use_float32", [False, True])
def test_named_gate_helpers(use_float32, rng):
    """apply_cx / apply_cz / apply_rx / apply_ry / apply_rz against the oracle
    (reference statevector.py:207-268) and the cases of
    tests/unit/test_simulator.py:536-588."""
    tol = 1e-6 if use_float32 else 1e-12
    sim2 = desv.DenseSVSimulator(2, use_float32=use_float32)
    sim2.apply_cx(0, 1)
    assert sim2.get_probabilities()[0] > 0.99
    sim2.apply_gate_1q(desv.GATES['x'], 0)
    sim2.apply_cx(0, 1)
    assert sim2.get_probabilities()[3] > 0.99          # |10> -> |11>
    sim2.apply_cx(0, 1)
    assert sim2.get_probabilities()[2] > 0.99          # CX twice = identity
    sim2.set_initial_state(None)
    sim2.apply_cz(0, 1)
    assert sim2.get_probabilities()[0] > 0.99
    sim2.apply_gate_1q(desv.GATES['h'], 0)
    sim2.apply_cx(0, 1)
    assert abs(np.sum(sim2.get_probabilities()) - 1) < tol
    for bad in (lambda: sim2.apply_cx(0, 0), lambda: sim2.apply_cx(0, 5), lambda: sim2.apply_cz(1, 1),
                lambda: sim2.apply_cz(0, 5), lambda: sim2.apply_gate_2q(np.eye(4), 0, 0),
                lambda: sim2.apply_gate_2q(np.eye(4), 0, 5), lambda: sim2.apply_rx(2, 0.3)):
        with pytest.raises(ValueError):
            bad()
    # random sequences of the helpers on 7 qubits against the oracle's table path
    n = 7
    sv0 = rand_state(n, rng)
    sim = desv.DenseSVSimulator(n, use_float32=use_float32)
    sim.sv = sv0
    want = sv0.copy()
    for _ in range(60):
        k = int(rng.integers(5))
        if k < 2:
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            (sim.apply_cx if k == 0 else sim.apply_cz)(a, b)
            row = [20 if k == 0 else 21, a, b, 0.0]
        else:
            q, th = int(rng.integers(n)), float(rng.uniform(-np.pi, np.pi))
            (sim.apply_rx, sim.apply_ry, sim.apply_rz)[k - 2](q, th)
            row = [(9, 10, 11)[k - 2], q, 0, th]
        want = O.run_ops(want, np.array([row], dtype=np.float64))
    assert fid(sim.get_statevector(), want) >= 1 - tol


@pytest.mark.parametrize("use_float32", [False, True])
def test_normalize(use_float32, rng):
    n = 9
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    sim = desv.DenseSVSimulator(n, use_float32=use_float32)
    sim.sv = 3.7 * v / np.linalg.norm(v)           # assignment does not normalise
    assert abs(sim.handle.norm2() - 3.7 ** 2) < 1e-4
    sim.normalize()
    got = sim.get_statevector()
    assert abs(np.vdot(got, got).real - 1) < (1e-6 if use_float32 else 1e-12)
    assert np.allclose(got, (v / np.linalg.norm(v)).astype(got.dtype), atol=1e-6 if use_float32 else 1e-12)
    z = desv.DenseSVSimulator(3, use_float32=use_float32)
    z.sv = np.zeros(8)
    z.normalize()                                   # a zero state stays zero (norm below 1e-12)
    assert np.all(z.get_statevector() == 0)


@pytest.mark.parametrize("pattern", ['linear', 'circular', 'full', 'star', 'brick'])
def test_entangling_layers_and_random_circuits_on_device(pattern):
    """Every entangling pattern (tests/unit/test_topology.py:56-64) and a
    generated random circuit (tests/unit/test_random_circuit.py:27-32) run
    through run_circuit and match the oracle."""
    circ = [('h', 0)] + desv.entangling_layer(5, pattern=pattern) + desv.random_circuit(5, 50, seed=7)
    sim = desv.DenseSVSimulator(5)
    sim.run_circuit(circ)
    want = O.run_circuit(O.transpile(circ), 5)
    assert np.allclose(sim.get_statevector(), want, atol=1e-12)
    assert abs(sim.get_probabilities().sum() - 1.0) < 1e-9


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_sampling_never_draws_zero_probability_outcomes(dtype):
    """Amplitudes that are exactly zero (here at both ends of every
    4096-amplitude CDF chunk, and after a measurement collapse) get zero
    width in the device CDF: uniforms placed on and around every cumulative
    edge never select them, and the draws are ordered with the uniforms."""
    import ctypes as C
    from desv import _capi
    n = 14
    rng = np.random.default_rng(21)
    v = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(dtype)
    v[rng.random(1 << n) < 0.3] = 0
    v[::4096] = 0
    v[4095::4096] = 0
    h = desv.StateHandle(n, dtype == np.complex128)
    h.upload(v)
    p = np.abs(v.astype(np.complex128)) ** 2
    edges = np.cumsum(p) / p.sum()
    u = np.concatenate([edges, np.nextafter(edges, 0), np.nextafter(edges, 1), rng.random(20000)])
    u = np.clip(u, 0.0, np.nextafter(1.0, 0))
    u.sort()
    out = np.empty(len(u), dtype=np.int64)
    _capi.check(_capi.lib.desv_sample(h.ptr, _capi.dptr(u), len(u), _capi.i64ptr(out)))
    assert np.all(p[out] > 0)
    assert np.all(np.diff(out) >= 0)          # a nondecreasing CDF: sorted uniforms, sorted outcomes
    # after a collapse half the amplitudes are zero
    sim = desv.DenseSVSimulator(n, use_float32=dtype == np.complex64)
    sim.sv = v
    sim.measure(3, rng=np.random.default_rng(2))
    pc = np.abs(sim.get_statevector().astype(np.complex128)) ** 2
    counts = desv.sample_counts(sim.sv, 200000, rng=np.random.default_rng(4))
    assert all(pc[int(k, 2)] > 0 for k in counts)
    h.close()
