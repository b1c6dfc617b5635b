"""The host compiler (fusion + cache-blocked stage scheduling) preserves the
circuit: every compiled plan is replayed by a NumPy interpreter of its device
ops and compared with the oracle.  No kernel runs here; this pins the plan,
not the kernels."""
import numpy as np
import pytest

import desv
from oracle import sv_oracle as O

OPK_DENSE1, OPK_DENSE2, OPK_DENSEK, OPK_DIAG, OPK_SCALAR, OPK_GRAD, OPK_GRADDIAG, OPK_MFMA = range(8)


def interpret(plan, sv):
    """Apply every device op of a plan to the full vector."""
    tile_masks, stage_ops, ops, mats = plan
    sv = sv.astype(np.complex128).copy()
    N = sv.shape[0]
    idx = np.arange(N, dtype=np.int64)
    for tm, so in zip(tile_masks, stage_ops):
        b, e, sperm = int(so[0]), int(so[1]), [int(x) for x in so[2:]]
        tbits = [i for i in range(64) if (int(tm) >> i) & 1]

        def g(pos):
            return tbits[pos] if pos >= 0 else -1 - pos

        for op in ops[b:e]:
            kind, k, moff = int(op[0]), int(op[1]), int(op[2])
            cm = int(op[10])
            for j in range(32):
                if (int(op[9]) >> j) & 1:
                    cm |= 1 << tbits[j]
            bits = [g(int(op[4 + j])) for j in range(k)]
            ctrl_ok = (idx & cm) == cm
            if kind == OPK_SCALAR:
                sv *= mats[moff]
            elif kind == OPK_DIAG:
                d = np.zeros(N, dtype=np.int64)
                for j, bb in enumerate(bits):
                    d |= ((idx >> bb) & 1) << j
                f = mats[moff:moff + (1 << k)][d]
                sv = np.where(ctrl_ok, f * sv, sv)
            elif kind in (OPK_DENSE1, OPK_DENSE2, OPK_DENSEK, OPK_MFMA):
                D = 1 << k
                M = mats[moff:moff + D * D].reshape(D, D)
                tmask = sum(1 << bb for bb in bits)
                base = idx[((idx & tmask) == 0) & ctrl_ok]
                offs = np.array([sum(((s >> j) & 1) << bits[j] for j in range(k)) for s in range(D)])
                A = sv[base[:, None] + offs[None, :]]
                sv[base[:, None] + offs[None, :]] = A @ M.T
            else:
                raise AssertionError(f"unexpected op kind {kind}")
        # write-back permutation: physical tile bit tbits[sperm[j]] moves to tbits[j]
        if any(sperm[j] != j for j in range(len(tbits))):
            dest = np.zeros(N, dtype=np.int64)
            rest = idx & ~np.int64(int(tm))
            dest |= rest
            for j in range(len(tbits)):
                dest |= ((idx >> tbits[sperm[j]]) & 1) << tbits[j]
            out = np.empty_like(sv)
            out[dest] = sv
            sv = out
    return sv


def random_ops(n, count, rng):
    rows = []
    for _ in range(count):
        if rng.random() < 0.4:
            gid = int(rng.choice([20, 21, 22, 23, 24, 25]))
            a, b = rng.choice(n, 2, replace=False)
            rows.append([gid, a, b, rng.uniform(-np.pi, np.pi)])
        else:
            rows.append([int(rng.integers(0, 15)), int(rng.integers(0, n)), 0, rng.uniform(-np.pi, np.pi)])
    return np.array(rows, dtype=np.float64)


OPTS = [
    dict(fuse_bits=1, tile_bits=6, low_bits=2),
    dict(fuse_bits=2, tile_bits=6, low_bits=2),
    dict(fuse_bits=3, tile_bits=7, low_bits=3),
    dict(fuse_bits=5, tile_bits=8, low_bits=2),
    dict(fuse_bits=4, tile_bits=8, low_bits=3, per_gate=True),
    dict(fuse_bits=2, tile_bits=5, low_bits=1, per_gate=True),
    dict(),
]


@pytest.mark.parametrize("opts", OPTS)
@pytest.mark.parametrize("n", [3, 9, 11])
def test_plan_replays_the_circuit(n, opts, rng):
    ops = random_ops(n, 90, rng)
    sv0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    sv0 /= np.linalg.norm(sv0)
    prog = desv.Program(n).add_ops(ops).compile("complex128", **opts)
    got = interpret(prog.plan(), sv0)
    want = O.run_ops(sv0.copy(), ops)
    assert np.allclose(got, want, atol=1e-10)
    st = prog.stats()
    assert st["gates"] == len(ops)
    if opts.get("per_gate"):
        assert st["passes"] == st["blocks"]


def test_remapped_schedules_restore_the_layout(rng):
    n = 14
    gates = desv.workloads.brickwork_su4(n, 10, seed=3)
    prog = desv.Program(n)
    want = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    want /= np.linalg.norm(want)
    sv0 = want.copy()
    for u, a, b in gates:
        prog.add_matrix(u, [a, b])
        want = O.apply_matrix(want, u, [a, b])
    for tb, lb in ((8, 3), (9, 4), (10, 5)):
        prog.compile("complex128", fuse_bits=2, tile_bits=tb, low_bits=lb)
        tm, so, ops, mats = prog.plan()
        assert any(any(so[s, 2 + j] != j for j in range(tb)) for s in range(len(so)))
        assert np.allclose(interpret(prog.plan(), sv0), want, atol=1e-10)


@pytest.mark.parametrize("fuse,tile,low", [(2, 7, 3), (3, 8, 4), (4, 7, 3)])
def test_trotter_staircase_fuses_single_qubit_gates(fuse, tile, low):
    """A Trotter staircase (XX, YY, ZZ on every bond, as h/sdg/cx/rz
    sequences) compiles to blocks only: every 1-qubit gate joins a block of a
    neighbouring 2-qubit gate, including the last gates of a qubit before a
    stage boundary, which wait for the stage of that qubit's next gate.  The
    plan still replays the circuit exactly."""
    n = 12
    circ = desv.workloads.xxz_trotter_circuit(n, 0.05, 6, seed=0)
    ops = desv.simulator.encode_op_table(desv.QuantumTranspiler.transpile(circ), n)
    prog = desv.Program(n).add_ops(ops).compile("complex128", fuse_bits=fuse, tile_bits=tile, low_bits=low)
    tm, so, dops, mats = prog.plan()
    assert prog.stats()["passes"] > 1
    lone = [o for o in dops if int(o[1]) == 1 and int(o[0]) in (OPK_DENSE1, OPK_DIAG)]
    assert lone == []
    # at most one block per bond and step (each bond's 21 h/sdg/s/cx/rz gates fuse into one)
    assert len(dops) <= (n - 1) * 6
    rng = np.random.default_rng(5)
    sv0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    sv0 /= np.linalg.norm(sv0)
    assert np.allclose(interpret((tm, so, dops, mats), sv0), O.run_ops(sv0.copy(), ops), atol=1e-10)


def test_cache_blocking_reduces_passes(rng):
    n = 16
    prog = desv.Program(n).add_circuit(desv.workloads.random_layers_circuit(n, 4, seed=1))
    prog.compile("complex64", fuse_bits=2, tile_bits=12, low_bits=5)
    blocked = prog.stats()["passes"]
    prog.compile("complex64", fuse_bits=1, per_gate=True, tile_bits=12, low_bits=5)
    plain = prog.stats()["passes"]
    assert blocked * 4 < plain


def test_matrix_programs_and_wide_blocks(rng):
    n = 10
    prog = desv.Program(n)
    sv0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    sv0 /= np.linalg.norm(sv0)
    want = sv0.copy()
    for _ in range(25):
        q = int(rng.integers(0, n - 1))
        u = desv.workloads.haar_su4(rng)
        prog.add_matrix(u, [q, q + 1])
        want = O.apply_matrix(want, u, [q, q + 1])
    for fb, mfma in ((5, True), (4, False), (3, False), (2, False)):
        prog.compile("complex64", fuse_bits=fb, tile_bits=8, low_bits=3, use_mfma=mfma)
        got = interpret(prog.plan(), sv0)
        assert np.allclose(got, want, atol=1e-10)
        if mfma:
            assert prog.stats()["mfma_blocks"] > 0


def test_wider_than_tile_is_rejected():
    prog = desv.Program(12).add_ops(np.array([[1, 0, 0, 0.0]]))
    with pytest.raises(ValueError):
        prog.compile("complex64", fuse_bits=5, tile_bits=4, low_bits=2, per_gate=True)
