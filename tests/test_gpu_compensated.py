"""Compensated complex64 programs (FP32 matrices carried as hi + lo pairs, the
lo products accumulated first) against the complex128 oracle.

Two checks per op kind: (1) parity, fidelity >= 1 - 1e-6 on random programs
that exercise every device op (dense 1-/2-qubit, VALU 3-5-qubit blocks, MFMA
3-, 4- and 5-qubit blocks, diagonal and scalar ops); (2) the property the
compensation exists for: one layer applied many times drifts in norm in plain
complex64 (the FP32 rounding of each matrix repeats coherently), and the
compensated run must remove that drift, not merely match the plain run.  A
compensated kernel that dropped or mis-signed its lo part would pass (1) and
fail (2).
"""
import numpy as np
import pytest

import desv
from desv import workloads
from oracle import sv_oracle as O

pytestmark = pytest.mark.gpu


def rand_state(n, rng):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


def random_ops(n, count, rng):
    rows = []
    for _ in range(count):
        if rng.random() < 0.4:
            g = int(rng.choice([20, 21, 22, 23, 24, 25]))
            a, b = rng.choice(n, 2, replace=False)
            rows.append([g, a, b, rng.uniform(-np.pi, np.pi)])
        else:
            rows.append([int(rng.integers(0, 15)), int(rng.integers(0, n)), 0, rng.uniform(-np.pi, np.pi)])
    return np.array(rows, dtype=np.float64)


OPTS = [
    dict(fuse_bits=1, tile_bits=8, low_bits=3),
    dict(fuse_bits=2, tile_bits=9, low_bits=4),
    dict(fuse_bits=3, tile_bits=9, low_bits=3),
    dict(fuse_bits=5, tile_bits=10, low_bits=2),
    dict(fuse_bits=3, tile_bits=10, use_mfma=True),
    dict(fuse_bits=4, tile_bits=11, use_mfma=True),
    dict(fuse_bits=5, tile_bits=12, use_mfma=True),
    dict(fuse_bits=2, tile_bits=8, per_gate=True),
]


@pytest.mark.parametrize("opts", OPTS)
def test_compensated_programs_match_oracle(opts, rng):
    n = 13
    ops = random_ops(n, 300, rng)
    sv0 = rand_state(n, rng)
    prog = desv.Program(n).add_ops(ops).compile("complex64", compensated=True, **opts)
    h = desv.StateHandle(n, complex128=False)
    h.upload(sv0)
    prog.run(h)
    f = O.fidelity(h.download(), O.run_ops_fast(sv0, ops))
    h.close()
    assert f >= 1 - 1e-6


@pytest.mark.parametrize("fuse,mfma", [(2, False), (4, False), (3, True), (4, True), (5, True)])
def test_compensated_brickwork_matches_oracle(fuse, mfma):
    n = 12
    prog = desv.Program(n)
    want = O.zero_state(n)
    for u, a, b in workloads.brickwork_su4(n, 8, seed=5):
        prog.add_matrix(u, [a, b])
        want = O.apply_matrix(want, u, [a, b])
    prog.compile("complex64", fuse_bits=fuse, use_mfma=mfma, compensated=True)
    if mfma:
        assert prog.stats()["mfma_blocks"] > 0
    h = desv.StateHandle(n, complex128=False)
    prog.run(h)
    assert O.fidelity(h.download(), want) >= 1 - 1e-6
    h.close()


def _layer(kind, n, rng):
    """(op-table rows or (matrix, qubits) gates) of one layer that compiles to
    the named device op kind.  Layers of the fused kinds are a Trotter step,
    the workload whose repetition shows the drift (a brickwork of Haar gates
    fused into 5-qubit blocks drifts too little to tell the modes apart)."""
    if kind == "dense1":
        return [("o", [10, q, 0, rng.uniform(0, 6)]) for q in range(n)]
    if kind == "diag":
        return [("o", [11, q, 0, rng.uniform(0, 6)]) for q in range(n)] + \
               [("o", [22, q, q + 1, rng.uniform(0, 6)]) for q in range(0, n - 1, 2)]
    if kind == "scalar":
        return [("o", [14, 0, 0, rng.uniform(0, 6)])] + [("o", [9, q, 0, rng.uniform(0, 6)]) for q in range(n)]
    if kind != "dense2":
        # fused blocks: one first-order Trotter step of the XXZ chain
        step = workloads.xxz_trotter_circuit(n, 0.05, 1, seed=0)[2 * n:]
        return [("o", list(r)) for r in O.encode_ops(step, n)]
    # dense 2-qubit: a brickwork of Haar SU(4) gates
    gates = []
    for d in range(4):
        for a in range(d % 2, n - 1, 2):
            gates.append(("m", workloads.haar_su4(rng), [a, a + 1]))
    return gates


KINDS = [
    ("dense1", dict(fuse_bits=1)),
    ("diag", dict(fuse_bits=1)),
    ("scalar", dict(fuse_bits=1)),
    ("dense2", dict(fuse_bits=2)),
    ("valu3", dict(fuse_bits=3)),
    ("valu5", dict(fuse_bits=5)),
    ("mfma3", dict(fuse_bits=3, use_mfma=True)),
    ("mfma4", dict(fuse_bits=4, use_mfma=True)),
    ("mfma5", dict(fuse_bits=5, use_mfma=True)),
]


@pytest.mark.parametrize("kind,opts", KINDS, ids=[k for k, _ in KINDS])
def test_compensation_removes_coherent_norm_drift(kind, opts):
    """One fixed layer applied R times from a random state: plain complex64
    drifts in norm linearly in R (the same rounded matrices every time), the
    compensated run stays at random-rounding level."""
    n, reps = 12, 400
    rng = np.random.default_rng(11)
    layer = _layer(kind, n, rng)
    prog = desv.Program(n)
    for _ in range(reps):
        for g in layer:
            if g[0] == "o":
                prog.add_ops(np.array([g[1]], dtype=np.float64))
            else:
                prog.add_matrix(g[1], g[2])
    sv0 = rand_state(n, np.random.default_rng(3))
    want = sv0.copy()
    drift = {}
    for comp in (False, True):
        prog.compile("complex64", tile_bits=11, compensated=comp, **opts)
        h = desv.StateHandle(n, complex128=False)
        h.upload(sv0)
        prog.run(h)
        got = h.download().astype(np.complex128)
        drift[comp] = (np.vdot(got, got).real - 1.0, got)
        h.close()
    # exact reference state: apply the layer reps times in complex128
    for _ in range(reps):
        for g in layer:
            want = O.run_ops_fast(want, np.array([g[1]], dtype=np.float64)) if g[0] == "o" \
                else O.apply_matrix(want, g[1], g[2])
    plain, comp = drift[False][0], drift[True][0]
    f_comp = O.fidelity(drift[True][1], want)
    print(f"{kind}: norm - 1 plain {plain:.3e}, compensated {comp:.3e}; compensated 1 - F = {1 - f_comp:.3e}")
    # random rounding alone leaves a few 1e-7 after these 4,000-9,000 ops
    # (a CPU emulation of both arithmetics gave plain 3e-6 to 4e-5 and
    # compensated 7e-8 to 3e-7 for the first five layer kinds)
    assert abs(comp) < 1e-6
    assert abs(plain) > 5 * abs(comp)
    assert f_comp >= 1 - 1e-6


def test_simulator_default_is_compensated():
    sim = desv.DenseSVSimulator(4, use_float32=True)
    assert sim.compensated
    desv.set_compensation(False)
    try:
        assert not desv.DenseSVSimulator(4, use_float32=True).compensated
        assert desv.DenseSVSimulator(4, use_float32=True, compensated=True).compensated
    finally:
        desv.set_compensation(True)
    assert not desv.DenseSVSimulator(4).compensated  # complex128 never compensates
