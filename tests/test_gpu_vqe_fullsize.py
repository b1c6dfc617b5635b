"""The differentiation and batching workloads at their full sizes.

* 20-qubit hardware-efficient ansatz, 8 layers of RY + RZ on every qubit and
  a CZ ladder (320 parameters), 200 random Pauli terms of weight 1-4 with
  N(0, 1) coefficients: energy and all 320 gradient entries from the adjoint
  kernel against the oracle's complex128 adjoint (reference energy function
  ``solvers/autodiff.py:92-169``), in complex128 (1e-9) and complex64 (1e-6
  relative), and against central finite differences (h = 1e-4, complex128)
  on a sample of parameters (the pattern of ``tests/unit/test_autodiff.py:
  48-69``).
* 1024 independent instances, n ~ U{10..22}, 4-8 layers, 50 terms each, one
  complex64 batch in the autotuned launch shape: every instance against its
  own single-instance run, and one instance per qubit count against the
  oracle's adjoint (1e-6 relative).

Relative bar: max |got - ref| <= 1e-6 max(1, max |ref|).
"""
import numpy as np
import pytest

import desv
from desv import workloads
from oracle import sv_oracle as O

pytestmark = pytest.mark.gpu


def rel_err(got, ref):
    got, ref = np.asarray(got, dtype=np.float64), np.asarray(ref, dtype=np.float64)
    return float(np.max(np.abs(got - ref), initial=0.0)) / max(1.0, float(np.max(np.abs(ref), initial=0.0)))


@pytest.fixture(scope="module")
def vqe20():
    n = 20
    circ, theta, terms = workloads.vqe_instance(n, 8, 200, seed=0)
    e_ref, g_ref = O.adjoint_gradient(O.build_template(circ.ops, n), theta, terms, n)
    return n, circ, theta, terms, e_ref, g_ref


def test_vqe20_complex128_against_oracle(vqe20):
    n, circ, theta, terms, e_ref, g_ref = vqe20
    fn, n_params = desv.circuit_to_energy_fn(circ, n, dtype="complex128")
    assert n_params == 320 and len(g_ref) == 320
    e, g = fn.value_and_grad(theta, desv.PauliSumOperator(terms, n))
    print(f"20q VQE c128: |dE| = {abs(e - e_ref):.2e}, max |dgrad| = {np.max(np.abs(g - g_ref)):.2e}")
    assert abs(e - e_ref) <= 1e-9
    assert np.max(np.abs(g - g_ref)) <= 1e-9


def test_vqe20_complex64_against_oracle(vqe20):
    n, circ, theta, terms, e_ref, g_ref = vqe20
    fn, _ = desv.circuit_to_energy_fn(circ, n, dtype="complex64")
    e, g = fn.value_and_grad(theta, desv.PauliSumOperator(terms, n))
    print(f"20q VQE c64: |dE| = {abs(e - e_ref):.2e}, max |dgrad| = {np.max(np.abs(g - g_ref)):.2e} "
          f"(relative {rel_err(e, e_ref):.2e}, {rel_err(g, g_ref):.2e})")
    assert rel_err(e, e_ref) <= 1e-6
    assert rel_err(g, g_ref) <= 1e-6


def test_vqe20_finite_differences(vqe20):
    n, circ, theta, terms, e_ref, g_ref = vqe20
    fn, _ = desv.circuit_to_energy_fn(circ, n, dtype="complex128")
    H = desv.PauliSumOperator(terms, n)
    h = 1e-4
    for i in np.random.default_rng(1).choice(len(theta), 12, replace=False):
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (fn(tp, H)[0] - fn(tm, H)[0]) / (2 * h)
        # central differences: truncation error h^2 |E'''| / 6 (about 1e-9 here)
        assert abs(fd - g_ref[i]) <= 1e-6, (i, fd, g_ref[i])


@pytest.fixture(scope="module")
def batch1024(tmp_path_factory):
    from desv.batch import Batch
    insts = workloads.batch_instances(count=1024)
    b = Batch("complex64")
    for n, circ, theta, terms in insts:
        b.add(circ, n, theta, terms)
    import os
    old = os.environ.get("DESV_TUNE_CACHE")
    os.environ["DESV_TUNE_CACHE"] = str(tmp_path_factory.mktemp("tune") / "tune.json")
    try:
        tuned = b.autotune(reps=1, cache=False)
    finally:
        if old is None:
            os.environ.pop("DESV_TUNE_CACHE", None)
        else:
            os.environ["DESV_TUNE_CACHE"] = old
    es, gs = b.run()
    b.close()
    return insts, tuned["best"], es, gs


def test_batch1024_matches_single_instances(batch1024):
    """Every instance of the complex64 batch against its own single-instance
    complex128 run (the FP64 result, itself pinned to the oracle above)."""
    insts, shape, es, gs = batch1024
    assert sorted({n for n, *_ in insts}) == list(range(10, 23))
    errs = []
    for (n, circ, theta, terms), e, g in zip(insts, es, gs):
        fn, _ = desv.circuit_to_energy_fn(circ, n, dtype="complex128")
        e1, g1 = fn.value_and_grad(theta, desv.PauliSumOperator(terms, n))
        fn.program.close()
        errs.append((rel_err(e, e1), rel_err(g, g1)))
    errs = np.array(errs)
    print(f"1024-instance c64 batch {shape} vs single-instance c128 runs: relative dE max {errs[:, 0].max():.2e} "
          f"p50 {np.median(errs[:, 0]):.2e}, dgrad max {errs[:, 1].max():.2e} p50 {np.median(errs[:, 1]):.2e}")
    assert errs.max() <= 1e-6


def test_batch1024_against_oracle_one_per_width(batch1024):
    insts, shape, es, gs = batch1024
    chosen = {}
    for i, (n, circ, theta, terms) in enumerate(insts):
        L = len(theta) // (2 * n)
        if n not in chosen or L < chosen[n][1]:
            chosen[n] = (i, L)
    for n in sorted(chosen):
        i, L = chosen[n]
        _, circ, theta, terms = insts[i]
        e_ref, g_ref = O.adjoint_gradient(O.build_template(circ.ops, n), theta, terms, n)
        de, dg = rel_err(es[i], e_ref), rel_err(gs[i], g_ref)
        print(f"instance {i} (n={n}, {L} layers): relative dE {de:.2e}, dgrad {dg:.2e}")
        assert de <= 1e-6 and dg <= 1e-6
