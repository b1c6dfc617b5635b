"""Register-resident stages on the GPU against the oracle.

Stages of 2^12-amplitude complex64 tiles whose blocks are all uncontrolled
dense or diagonal blocks of at most 3 qubits keep the tile in registers
(`reg_groups=True`); their layouts and exchanges are replayed on the CPU in
test_reg_groups.py; so are the wave-tile stages (`reg_groups=2`: one 2^11-
or 2^12-amplitude tile per wave, loaded from HBM into registers and stored
back, no LDS).  Here the kernels run: brickwork circuits of Haar SU(4)
gates (plain and compensated matrices) and mixed op tables (1-2-qubit dense
and diagonal gates, with and without controls, so register and LDS stages
alternate) against the oracle, and the register path against the LDS path of
the same compiled plan.
"""
import numpy as np
import pytest

import desv
from desv import workloads
from oracle import sv_oracle as O

pytestmark = pytest.mark.gpu


def _brickwork(n, depth, seed):
    gates = workloads.brickwork_su4(n, depth, seed=seed)
    prog = desv.Program(n)
    want = O.zero_state(n)
    for u, a, b in gates:
        prog.add_matrix(u, [a, b])
        want = O.apply_matrix(want, u, [a, b])
    return prog, want


@pytest.mark.parametrize("n,fuse,comp", [(13, 3, False), (14, 3, True), (14, 2, False), (16, 3, False)])
def test_brickwork_register_stages_against_oracle(n, fuse, comp):
    prog, want = _brickwork(n, 8, seed=n)
    prog.compile("complex64", fuse_bits=fuse, tile_bits=12, use_mfma=True, reg_groups=True, compensated=comp)
    sw = prog.reg_program(comp)[0]
    assert (sw >= 0).any()
    h = desv.StateHandle(n, False)
    prog.run(h)
    got = h.download().astype(np.complex128)
    h.close()
    # FP32 arithmetic: amplitudes of a random 2^n state are ~2^(-n/2)
    assert np.abs(got - want).max() < 2e-6 * 2 ** (-n / 2) * 64
    assert abs(np.vdot(want, got)) ** 2 >= 1 - 1e-6


def _mixed_ops(n, count, rng, controlled):
    rows = []
    for _ in range(count):
        u = rng.random()
        if controlled and u < 0.2:
            a, b = rng.choice(n, 2, replace=False)
            rows.append([int(rng.choice([20, 21, 22, 24, 25])), a, b, rng.uniform(-np.pi, np.pi)])
        elif u < 0.45:
            a, b = rng.choice(n, 2, replace=False)
            rows.append([23, a, b, 0.0])
        else:
            rows.append([int(rng.integers(1, 14)), int(rng.integers(0, n)), 0, rng.uniform(-np.pi, np.pi)])
    return np.array(rows, dtype=np.float64)


@pytest.mark.parametrize("fuse,controlled", [(1, False), (2, False), (3, True)])
def test_mixed_register_stages_against_oracle(fuse, controlled, rng):
    n = 14
    ops = _mixed_ops(n, 150, rng, controlled)
    prog = desv.Program(n).add_ops(ops)
    prog.compile("complex64", fuse_bits=fuse, tile_bits=12, use_mfma=True, reg_groups=True)
    sv0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    sv0 /= np.linalg.norm(sv0)
    h = desv.StateHandle(n, False)
    h.upload(sv0)
    prog.run(h)
    got = h.download().astype(np.complex128)
    h.close()
    want = O.run_ops_fast(sv0, ops)
    assert abs(np.vdot(want, got)) ** 2 >= 1 - 1e-6


def test_register_and_lds_paths_agree():
    n = 15
    prog, _ = _brickwork(n, 10, seed=7)
    outs = []
    for rg in (False, True):
        prog.compile("complex64", fuse_bits=3, tile_bits=12, use_mfma=True, reg_groups=rg)
        h = desv.StateHandle(n, False)
        prog.run(h)
        outs.append(h.download().astype(np.complex128))
        h.close()
    assert np.abs(outs[0] - outs[1]).max() < 1e-6


@pytest.mark.parametrize("n,tile,fuse,comp", [(13, 11, 3, False), (14, 12, 3, True), (15, 12, 2, False),
                                              (16, 11, 3, True), (17, 12, 3, False)])
def test_brickwork_wave_stages_against_oracle(n, tile, fuse, comp):
    prog, want = _brickwork(n, 8, seed=n + 1)
    prog.compile("complex64", fuse_bits=fuse, tile_bits=tile, use_mfma=True, reg_groups=2, compensated=comp)
    sw = prog.reg_program(comp)[0]
    assert (sw >= 0).all()
    h = desv.StateHandle(n, False)
    prog.run(h)
    got = h.download().astype(np.complex128)
    h.close()
    assert np.abs(got - want).max() < 2e-6 * 2 ** (-n / 2) * 64
    assert abs(np.vdot(want, got)) ** 2 >= 1 - 1e-6


@pytest.mark.parametrize("fuse,controlled", [(1, False), (2, False), (3, True)])
def test_mixed_wave_stages_against_oracle(fuse, controlled, rng):
    n = 14
    ops = _mixed_ops(n, 150, rng, controlled)
    prog = desv.Program(n).add_ops(ops)
    prog.compile("complex64", fuse_bits=fuse, tile_bits=11, use_mfma=True, reg_groups=2)
    sv0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    sv0 /= np.linalg.norm(sv0)
    h = desv.StateHandle(n, False)
    h.upload(sv0)
    prog.run(h)
    got = h.download().astype(np.complex128)
    h.close()
    want = O.run_ops_fast(sv0, ops)
    assert abs(np.vdot(want, got)) ** 2 >= 1 - 1e-6


def test_wave_and_lds_paths_agree():
    n = 16
    prog, _ = _brickwork(n, 10, seed=9)
    outs = []
    for rg in (0, 2):
        prog.compile("complex64", fuse_bits=3, tile_bits=12, use_mfma=True, reg_groups=rg)
        h = desv.StateHandle(n, False)
        prog.run(h)
        outs.append(h.download().astype(np.complex128))
        h.close()
    assert np.abs(outs[0] - outs[1]).max() < 1e-6


@pytest.mark.parametrize("n,count,tile,fuse,comp", [(14, 120, 12, 3, False), (16, 120, 12, 2, True)])
def test_relabelled_tail_against_oracle(n, count, tile, fuse, comp):
    """A wave-tile program ending with an out-of-place relabelling stage (the
    qubits outside its tile move with the tile addresses), run twice in a row
    (the handle's state buffer swaps each time), and its in-place form on a
    handle without a second buffer, against the oracle."""
    rng = np.random.default_rng(n * 1000 + count)
    ops = _mixed_ops(n, count, rng, False)
    prog = desv.Program(n).add_ops(ops)
    prog.compile("complex64", fuse_bits=fuse, tile_bits=tile, use_mfma=True, reg_groups=2, compensated=comp)
    assert prog.tail(comp)["fb_from"] >= 0
    sv0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    sv0 /= np.linalg.norm(sv0)
    want1 = O.run_ops_fast(sv0, ops)
    want2 = O.run_ops_fast(want1, ops)
    for double_buffer in (True, False):
        h = desv.StateHandle(n, False)
        h.set_double_buffer(double_buffer)
        h.upload(sv0)
        for want in (want1, want2):
            prog.run(h)
            got = h.download().astype(np.complex128)
            assert np.abs(got - want).max() < 2e-6 * 2 ** (-n / 2) * 64
            assert abs(np.vdot(want, got)) ** 2 >= 1 - 1e-6
        h.close()
