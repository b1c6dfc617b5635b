"""GPU adjoint gradients against the oracle (c128 adjoint, parameter shift
and central finite differences) and the torch autograd binding.

Bars: complex128 energy and gradient within 1e-9 of the oracle; complex64
(FP32 state, FP64 reductions) within 1e-6 relative to the complex128 oracle:
|dE| <= 1e-6 max(1, |E|) and max |dgrad| <= 1e-6 max(1, max |grad|).
"""
import numpy as np
import pytest

import desv
from desv import workloads
from oracle import sv_oracle as O

pytestmark = pytest.mark.gpu


def close_rel(got, ref, rtol=1e-6):
    """max |got - ref| <= rtol max(1, max |ref|)."""
    got, ref = np.asarray(got, dtype=np.float64), np.asarray(ref, dtype=np.float64)
    return float(np.max(np.abs(got - ref), initial=0.0)) <= rtol * max(1.0, float(np.max(np.abs(ref), initial=0.0)))

VQE_QASM = ('OPENQASM 2.0; include "qelib1.inc"; qreg q[2]; creg c[2]; '
            'ry(0.5) q[0]; rx(0.5) q[1]; cx q[0],q[1]; rz(0.2) q[1]; cx q[0],q[1]; '
            'ry(0.5) q[0]; rx(0.5) q[1]; measure q -> c;')


def test_vqe_qasm_dense_hamiltonian_fixture():
    """Reference fixture test_autodiff.py:27-69 (dense diag H seed 7, theta seed 3)."""
    circ = desv.QASMParser().parse(VQE_QASM)
    energy_fn, n_params = desv.circuit_to_energy_fn(circ, circ.n_qubits)
    assert n_params == 5
    H = np.diag(np.sort(np.random.default_rng(7).uniform(-2.5, 2.5, 4))).astype(np.complex128)
    theta = np.random.default_rng(3).uniform(-np.pi, np.pi, n_params)
    e, g = energy_fn.value_and_grad(theta, H)
    tmpl = O.build_template(circ.ops, 2)
    ops = O.patch_template(tmpl, theta)
    sv = O.run_ops(O.zero_state(2), ops)
    assert abs(e - np.real(np.vdot(sv, H @ sv))) < 1e-10
    eps = 1e-6
    fd = np.zeros(n_params)
    for i in range(n_params):
        tp, tm = theta.copy(), theta.copy()
        tp[i] += eps
        tm[i] -= eps
        fd[i] = (energy_fn(tp, H)[0] - energy_fn(tm, H)[0]) / (2 * eps)
    assert np.allclose(g, fd, atol=1e-6)
    # empty circuit -> |00>, energy = H[0, 0] (test_autodiff.py:71-79)
    empty = desv.QASMParser().parse('OPENQASM 2.0; qreg q[2]; creg c[2]; measure q -> c;')
    f0, n0 = desv.circuit_to_energy_fn(empty, 2)
    e0, sv0 = f0(np.array([]), np.diag([1.0, 2.0, 3.0, 4.0]))
    assert n0 == 0 and abs(e0 - 1.0) < 1e-9 and abs(abs(np.asarray(sv0)[0]) - 1) < 1e-9


@pytest.mark.parametrize("n,layers", [(3, 2), (8, 3), (12, 4)])
def test_hea_gradients_c128(n, layers):
    circ, theta, terms = workloads.vqe_instance(n, layers, 30, seed=n)
    tmpl = O.build_template(circ.ops, n)
    e_ref, g_ref = O.adjoint_gradient(tmpl, theta, terms, n)
    energy_fn, n_params = desv.circuit_to_energy_fn(circ, n)
    e, g = energy_fn.value_and_grad(theta, desv.PauliSumOperator(terms, n))
    assert abs(e - e_ref) < 1e-9
    assert np.allclose(g, g_ref, atol=1e-9)
    if n <= 8:
        assert np.allclose(g, O.param_shift_gradient(tmpl, theta, terms, n), atol=1e-9)


def test_hea_gradients_c64_error_bound():
    n, layers = 12, 4
    circ, theta, terms = workloads.vqe_instance(n, layers, 40, seed=2)
    tmpl = O.build_template(circ.ops, n)
    e_ref, g_ref = O.adjoint_gradient(tmpl, theta, terms, n)
    energy_fn, _ = desv.circuit_to_energy_fn(circ, n, dtype="complex64")
    e, g = energy_fn.value_and_grad(theta, desv.PauliSumOperator(terms, n))
    assert close_rel(e, e_ref) and close_rel(g, g_ref)


def test_phase_type_generators():
    n = 4
    ops = [{'name': 'h', 'qubits': [q]} for q in range(n)] + [
        {'name': 'p', 'qubits': [0]}, {'name': 'cp', 'qubits': [0, 2]}, {'name': 'crz', 'qubits': [1, 3]},
        {'name': 'u1', 'qubits': [1]}, {'name': 'rx', 'qubits': [2]}, {'name': 'cx', 'qubits': [2, 0]},
        {'name': 'ry', 'qubits': [3]}, {'name': 'crz', 'qubits': [3, 0]}]
    circ = desv.QASMCircuit(n, 0, ops)
    rng = np.random.default_rng(4)
    tmpl = O.build_template(ops, n)
    theta = rng.uniform(-np.pi, np.pi, int((tmpl[:, 3] == -1).sum()))
    terms = [(1.0, {0: 'X'}), (0.4, {1: 'Y', 2: 'X'}), (0.2, {0: 'Y', 3: 'Z'}), (-0.7, {2: 'Z', 1: 'X'})]
    e_ref, g_ref = O.adjoint_gradient(tmpl, theta, terms, n)
    fn, _ = desv.circuit_to_energy_fn(circ, n)
    e, g = fn.value_and_grad(theta, desv.PauliSumOperator(terms, n))
    assert abs(e - e_ref) < 1e-10
    assert np.allclose(g, g_ref, atol=1e-10)


def test_torch_autograd_binding():
    """torch.autograd.grad through the energy function equals the oracle's
    adjoint gradient; forward runs the energy only and backward the reverse
    sweep (a second evaluation in between forces a full rerun, still exact);
    the returned state is a zero-copy device tensor view."""
    torch = pytest.importorskip("torch")
    n = 8
    circ, theta, terms = workloads.vqe_instance(n, 3, 20, seed=9)
    fn, _ = desv.circuit_to_energy_fn(circ, n)
    H = desv.PauliSumOperator(terms, n)
    tmpl = O.build_template(circ.ops, n)
    e_ref, g_ref = O.adjoint_gradient(tmpl, theta, terms, n)
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    e, sv = fn(th, H)
    assert abs(e.item() - e_ref) < 1e-10
    (g,) = torch.autograd.grad(e, th)
    assert np.allclose(g.numpy(), g_ref, atol=1e-10)
    # interleaved evaluations: backward of the first after a second forward
    theta2 = theta + 0.3
    e1, _ = fn(th, H)
    e2, _ = fn(torch.tensor(theta2, requires_grad=True), H)
    (g1,) = torch.autograd.grad(e1, th)
    assert np.allclose(g1.numpy(), g_ref, atol=1e-10)
    # the state as a device tensor aliasing the function's buffer
    e3, sv = fn(th, H)
    t = sv.to_torch()
    assert t.is_cuda and t.dtype == torch.complex128 and t.shape == (1 << n,)
    assert t.data_ptr() == sv.handle.device_ptr()
    assert np.allclose(t.cpu().numpy(), np.asarray(sv), atol=0)
    want = O.run_ops_fast(O.zero_state(n), O.patch_template(tmpl, theta))
    assert np.allclose(t.cpu().numpy(), want, atol=1e-10)
    # a simulator's state: writes through the view are seen by the engine
    sim = desv.DenseSVSimulator(n, use_float32=True)
    sim.run_circuit([('h', q) for q in range(n)])
    v = sim.sv.to_torch()
    assert v.dtype == torch.complex64 and v.data_ptr() == sim.handle.device_ptr()
    v.mul_(2.0)
    torch.cuda.synchronize()
    assert abs(sim.handle.norm2() - 4.0) < 1e-5
    assert abs(float(e2.detach()) - O.energy(O.patch_template(tmpl, theta2), terms, n)[0]) < 1e-10


def test_parameter_shift_baseline_on_gpu():
    n = 6
    circ, theta, terms = workloads.vqe_instance(n, 2, 15, seed=1)
    fn, _ = desv.circuit_to_energy_fn(circ, n)
    H = desv.PauliSumOperator(terms, n)
    e, g = fn.value_and_grad(theta, H)
    gps = desv.parameter_shift_grad(fn.program, theta, H)
    assert np.allclose(g, gps, atol=1e-10)


def test_batch_matches_single_instances():
    from desv.batch import Batch
    insts = workloads.batch_instances(count=6, n_lo=4, n_hi=9, l_lo=1, l_hi=3, n_terms=10)
    b = Batch("complex128")
    for n, circ, theta, terms in insts:
        b.add(circ, n, theta, terms)
    es, gs = b.run()
    for (n, circ, theta, terms), e, g in zip(insts, es, gs):
        e_ref, g_ref = O.adjoint_gradient(O.build_template(circ.ops, n), theta, terms, n)
        assert abs(e - e_ref) < 1e-9
        assert np.allclose(g, g_ref, atol=1e-9)


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-9), ("complex64", 1e-6)])
def test_batch_refresh_graph_and_states(dtype, tol):
    """A compiled batch re-run with new angles (matrices refreshed, graph
    replayed) must match the oracle at the new angles; graph and direct
    launches agree; initial states and energies-only runs work."""
    from desv.batch import Batch
    insts = workloads.batch_instances(count=5, n_lo=3, n_hi=11, l_lo=1, l_hi=3, n_terms=12)
    rng = np.random.default_rng(5)
    for use_graph in (True, False):
        b = Batch(dtype)
        b.configure(tile_bits=7, use_graph=use_graph)   # small tiles: several stages per instance
        for n, circ, theta, terms in insts:
            b.add(circ, n, theta, terms)
        b.run()
        for rep in range(2):
            thetas = [rng.uniform(-np.pi, np.pi, len(t)) for _, _, t, _ in insts]
            for i, th in enumerate(thetas):
                b.set_params(i, th)
            es, gs = b.run()
            for (n, circ, _, terms), th, e, g in zip(insts, thetas, es, gs):
                e_ref, g_ref = O.adjoint_gradient(O.build_template(circ.ops, n), th, terms, n)
                assert close_rel(e, e_ref, tol) and close_rel(g, g_ref, tol)
        e_only = b.energies()
        assert close_rel(e_only, es, tol)
        # initial state for instance 0
        n0, c0, _, t0 = insts[0]
        psi0 = rng.normal(size=1 << n0) + 1j * rng.normal(size=1 << n0)
        psi0 /= np.linalg.norm(psi0)
        b.set_initial_state(0, psi0)
        es2, gs2 = b.run()
        tmpl = O.build_template(c0.ops, n0)
        ops = O.patch_template(tmpl, thetas[0])
        sv = O.run_ops_fast(psi0.astype(np.complex128), ops)
        assert close_rel(es2[0], np.real(np.vdot(sv, O.pauli_sum_matvec(sv, t0))), tol)
        assert O.fidelity(b.get_state(0, n0), sv) > 1 - max(tol, 1e-9)
        b.close()


@pytest.mark.parametrize("dtype,tol,shape", [
    ("complex128", 1e-9, dict(tile_bits=8, fuse_bits=3)),
    ("complex64", 1e-6, dict(tile_bits=8, fuse_bits=4, use_mfma=True)),
    ("complex64", 1e-6, dict(tile_bits=10, fuse_bits=5, use_mfma=True)),
])
def test_batch_launch_shapes(dtype, tol, shape):
    """Wider fusion (VALU k = 3 blocks, forward MFMA blocks) in the batched
    launches keeps energies and gradients on the oracle."""
    from desv.batch import Batch
    insts = workloads.batch_instances(count=5, n_lo=6, n_hi=11, l_lo=2, l_hi=3, n_terms=12)
    b = Batch(dtype)
    b.configure(**shape)
    for n, circ, theta, terms in insts:
        b.add(circ, n, theta, terms)
    es, gs = b.run()
    for (n, circ, theta, terms), e, g in zip(insts, es, gs):
        e_ref, g_ref = O.adjoint_gradient(O.build_template(circ.ops, n), theta, terms, n)
        assert close_rel(e, e_ref, tol) and close_rel(g, g_ref, tol)
    b.close()


def test_batch_autotune_keeps_parity(tmp_path, monkeypatch):
    from desv.batch import Batch
    monkeypatch.setenv("DESV_TUNE_CACHE", str(tmp_path / "tune.json"))
    insts = workloads.batch_instances(count=6, n_lo=5, n_hi=12, l_lo=1, l_hi=3, n_terms=10)
    b = Batch("complex64")
    for n, circ, theta, terms in insts:
        b.add(circ, n, theta, terms)
    res = b.autotune(reps=1)
    assert res["best"] in [{k: c[k] for k in ("tile_bits", "fuse_bits", "use_mfma")}
                           for c in res["candidates"] if c["ms"] is not None]
    es, gs = b.run()
    for (n, circ, theta, terms), e, g in zip(insts, es, gs):
        e_ref, g_ref = O.adjoint_gradient(O.build_template(circ.ops, n), theta, terms, n)
        assert close_rel(e, e_ref) and close_rel(g, g_ref)
    assert b.autotune()["cached"]
    b.close()


@pytest.mark.parametrize("compensated", [False, True])
def test_batch_compensation_modes_and_completion_times(compensated):
    """Both complex64 modes of a batch match the oracle; the device completion
    stamps are positive, within the run, and ordered by the instances' number
    of reverse stages (an instance's sums run right after its last one)."""
    from desv.batch import Batch
    insts = workloads.batch_instances(count=12, n_lo=5, n_hi=12, l_lo=1, l_hi=5, n_terms=10)
    for n, layers, seed in ((5, 1, 1), (12, 5, 2)):
        insts.append((n, *workloads.vqe_instance(n, layers, 10, seed=seed)))
    b = Batch("complex64", compensated=compensated)
    b.configure(tile_bits=8)
    for n, circ, theta, terms in insts:
        b.add(circ, n, theta, terms)
    es, gs = b.run()
    for (n, circ, theta, terms), e, g in zip(insts, es, gs):
        e_ref, g_ref = O.adjoint_gradient(O.build_template(circ.ops, n), theta, terms, n)
        assert close_rel(e, e_ref) and close_rel(g, g_ref)
    done = b.completion_us()
    assert np.all(done > 0) and done.max() <= b.last_time_us() + 1.0
    # a 5-qubit, 1-layer instance (one reverse stage at 2^8-amplitude tiles)
    # finishes before a 12-qubit, 5-layer one (several reverse stages): its
    # sums are launched earlier on the same stream
    assert done[-2] < done[-1], done
    e_only = b.energies()
    assert close_rel(e_only, es) and np.all(b.completion_us() > 0)
    b.close()
