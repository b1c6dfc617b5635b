"""Register-group stages, replayed on the CPU.

A complex64 stage of 2^12-amplitude tiles whose blocks are all uncontrolled
dense or diagonal blocks of at most 3 bits can keep the tile in registers
(`desv_compile_opts.reg_groups`, `run_reg_groups` in
dense-evolution_amd/csrc/engine_kernels.hip): 4 waves x 64 lanes x 16
registers hold the tile; a group reads it from the swizzled LDS tile in a
layout (which tile bit sits at which lane, register or wave bit), exchanges
lane and register bits, runs each block as a stacked real 16x16 MFMA product
and writes the tile back in the layout it ends in.  The host
(`plan.cpp::build_reg_program`) chooses the groups, layouts, exchanges and the
row/column order of every block's A fragments.

This module restates the kernel's semantics lane by lane in NumPy -- the LDS
slot addressing, the instruction-level behaviour of v_permlane16/32_swap and
of the DPP row shifts, bank masks and quad permutations (as measured on
MI355X: lane l of a row_shr:n result reads lane l - n of its 16-lane row,
permlane32_swap exchanges the upper half of its first operand with the lower
half of its second, permlane16_swap the odd rows of the first with the even
rows of the second), and the 16x16x4 MFMA operand layout -- and replays the
dumped register programs.  The result must equal the oracle.  With the
compensated fragments (hi + lo) the replay carries the matrices to about
2^-48, so a wrong layout, exchange, fragment order or slot shows up far above
the tolerance.
"""
import numpy as np
import pytest

import desv
from desv import workloads
from oracle import sv_oracle as O

HDR = 47
RG_XLANE, RG_MFMA3, RG_MFMA3T = 1, 2, 3


def slot(i):
    return i ^ ((i >> 5) & 31)


LANE = np.arange(64)


# ---- instruction-level restatements --------------------------------------

def permlane32_swap(x, y):
    return np.concatenate([x[:32], y[:32]]), np.concatenate([x[32:], y[32:]])


def permlane16_swap(x, y):
    nx, ny = x.copy(), y.copy()
    for r in (0, 2):
        # odd rows (1, 3) of x <-> even rows (0, 2) of y
        nx[(r + 1) * 16:(r + 2) * 16] = y[r * 16:(r + 1) * 16]
        ny[r * 16:(r + 1) * 16] = x[(r + 1) * 16:(r + 2) * 16]
    return nx, ny


def dpp_row_shift(old, src, n, bank_mask):
    """row_shr:n (n > 0) or row_shl:-n (n < 0) with a bank mask: enabled lanes
    whose source lane lies in the same 16-lane row read it, others keep old."""
    out = old.copy()
    for l in range(64):
        if not (bank_mask >> ((l % 16) // 4)) & 1:
            continue
        s = l - n
        if s // 16 == l // 16 and 0 <= s < 64:
            out[l] = src[s]
    return out


def quad_perm_xor(src, b):
    return src[LANE ^ (1 << b)]


def xlane(V, r, b):
    """V: (64, NJ) complex register file of one wave."""
    for j in range(V.shape[1]):
        if (j >> r) & 1:
            continue
        a, c = V[:, j].copy(), V[:, j | (1 << r)].copy()
        if b == 5:
            na, nc = permlane32_swap(a, c)
        elif b == 4:
            na, nc = permlane16_swap(a, c)
        elif b >= 2:
            bm1, bm0 = (0xa, 0x5) if b == 2 else (0xc, 0x3)
            na = dpp_row_shift(a, c, 1 << b, bm1)
            nc = dpp_row_shift(c, a, -(1 << b), bm0)
        else:
            up = ((LANE >> b) & 1).astype(bool)
            ta, tc = quad_perm_xor(c, b), quad_perm_xor(a, b)
            na = np.where(up, ta, a)
            nc = np.where(up, c, tc)
        V[:, j], V[:, j | (1 << r)] = na, nc


def mfma3(V, s, frag):
    """frag: (64, 4) A fragments (lane, ks) of the stacked real 16x16 matrix."""
    A = np.zeros((16, 16))
    for l in range(64):
        for ks in range(4):
            A[l % 16, 4 * ks + l // 16] = frag[l, ks]
    for j in range(V.shape[1]):
        if (j >> s) & 1:
            continue
        x0, x1 = V[:, j], V[:, j | (1 << s)]
        Bk = [x0.real, x1.real, x0.imag, x1.imag]
        Bm = np.zeros((16, 16))  # B[k][n]: k = 4 ks + l // 16, n = l % 16
        for ks in range(4):
            for l in range(64):
                Bm[4 * ks + l // 16, l % 16] = Bk[ks][l]
        D = A @ Bm
        y = np.array([[D[4 * (l // 16) + jj, l % 16] for jj in range(4)] for l in range(64)])
        V[:, j] = y[:, 0] + 1j * y[:, 1]
        V[:, j | (1 << s)] = y[:, 2] + 1j * y[:, 3]


def lane_slot(lpos, wpos, wave):
    idx = np.zeros(64, dtype=np.int64)
    for i in range(6):
        idx |= ((LANE >> i) & 1) << lpos[i]
    for i in range(2):
        idx |= ((wave >> i) & 1) << wpos[i]
    return slot(idx) * 8


def run_groups(lds, words, frags, fb):
    """lds: complex array indexed by LDS element (slot); words: the stage's
    register words; frags: this stage's fragment blocks (fb per MFMA op)."""
    ng, nq = int(words[0]), int(words[1])
    w, q = 2, 0
    for _ in range(ng):
        h = words[w:]
        nmu = int(h[0])
        for wave in range(4):
            lb = lane_slot(h[1:7], h[7:9], wave)
            V = np.zeros((64, 16), dtype=np.complex128)
            for j in range(16):
                V[:, j] = lds[(lb ^ int(h[9 + j])) // 8]
            qq = q
            for u in range(nmu):
                code = int(h[HDR + u])
                a1 = (code >> 4) & 15
                if code & 15 == RG_XLANE:
                    xlane(V, a1, (code >> 8) & 15)
                else:
                    assert code & 15 == RG_MFMA3
                    fr = frags[qq * fb].astype(np.float64)
                    if fb == 2:
                        fr = fr + frags[qq * fb + 1].astype(np.float64)
                    mfma3(V, a1, fr)
                    qq += 1
            lbo = lane_slot(h[25:31], h[7:9], wave)
            for j in range(16):
                lds[(lbo ^ int(h[31 + j])) // 8] = V[:, j]
        q = qq
        w += HDR + nmu
    assert q == nq


def goff_fn(tm, n, m):
    tbits = [b for b in range(n) if (tm >> b) & 1]
    L = 0
    while L < n and (tm >> L) & 1:
        L += 1
    hbits = [b for b in tbits if b >= L]

    def goff(e):
        o = e & ((1 << L) - 1)
        r, j = e >> L, 0
        while r:
            if r & 1:
                o |= 1 << hbits[j]
            r >>= 1
            j += 1
        return o
    return goff


def replay(prog, sv, n, compensated):
    """Replay a compiled program: register stages through run_groups, the
    others through the LDS replay of test_tile_layout."""
    from test_tile_layout import run_stage
    tm, so, ops, mats = prog.plan()
    sw, sf, words, frags = prog.reg_program(compensated)
    fb = 2 if compensated else 1
    nreg = 0
    for s in range(len(tm)):
        if sw[s] < 0:
            sv = run_stage(sv, n, int(tm[s]), so[s], ops, mats, 2)
            continue
        nreg += 1
        m = bin(int(tm[s])).count("1")
        assert m == 12
        goff = goff_fn(int(tm[s]), n, m)
        sperm = [int(x) for x in so[s][2:2 + m]]

        def perm(e):
            o, j = 0, 0
            while e:
                if e & 1:
                    o |= 1 << sperm[j]
                e >>= 1
                j += 1
            return o
        free = [b for b in range(n) if not (int(tm[s]) >> b) & 1]
        gidx = np.array([goff(i) for i in range(1 << m)])
        pidx = np.array([perm(e) for e in range(1 << m)])
        sl = slot(np.arange(1 << m))
        out = sv.copy()
        for t in range(1 << (n - m)):
            base = 0
            for i, b in enumerate(free):
                if (t >> i) & 1:
                    base |= 1 << b
            lds = np.zeros(1 << m, dtype=np.complex128)
            lds[sl] = sv[base | gidx]
            run_groups(lds, words[sw[s]:], frags[sf[s]:], fb)
            T = lds[sl]
            out[base | gidx] = T[pidx]
        sv = out
    return sv, nreg


def _rand_state(n, rng):
    sv = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return sv / np.linalg.norm(sv)


@pytest.mark.parametrize("n,depth,fuse", [(13, 6, 3), (14, 5, 3), (13, 6, 2)])
def test_brickwork_register_stages_replay(n, depth, fuse, rng):
    gates = workloads.brickwork_su4(n, depth, seed=3)
    prog = desv.Program(n)
    for u, a, b in gates:
        prog.add_matrix(u, [a, b])
    prog.compile("complex64", fuse_bits=fuse, tile_bits=12, use_mfma=True, reg_groups=True)
    sv0 = _rand_state(n, rng)
    got, nreg = replay(prog, sv0, n, compensated=True)
    assert nreg >= 1
    want = sv0.copy()
    for u, a, b in gates:
        want = O.apply_matrix(want, u, [a, b])
    assert np.abs(got - want).max() < 1e-10
    # the hi fragments alone: FP32 matrix rounding only
    got32, _ = replay(prog, sv0, n, compensated=False)
    assert np.abs(got32 - want).max() < 1e-5


def _mixed_ops(n, count, rng, controlled):
    rows = []
    for _ in range(count):
        u = rng.random()
        if controlled and u < 0.2:
            a, b = rng.choice(n, 2, replace=False)
            rows.append([int(rng.choice([20, 21, 22, 24, 25])), a, b, rng.uniform(-np.pi, np.pi)])
        elif u < 0.45:
            a, b = rng.choice(n, 2, replace=False)
            rows.append([23, a, b, 0.0])  # swap: dense 2-qubit, uncontrolled
        else:
            rows.append([int(rng.integers(1, 14)), int(rng.integers(0, n)), 0, rng.uniform(-np.pi, np.pi)])
    return np.array(rows, dtype=np.float64)


@pytest.mark.parametrize("fuse,controlled", [(1, False), (2, False), (3, False), (3, True)])
def test_mixed_blocks_register_stages_replay(fuse, controlled, rng):
    n = 13
    ops = _mixed_ops(n, 120, rng, controlled)
    prog = desv.Program(n).add_ops(ops)
    prog.compile("complex64", fuse_bits=fuse, tile_bits=12, use_mfma=True, reg_groups=True)
    sv0 = _rand_state(n, rng)
    got, nreg = replay(prog, sv0, n, compensated=True)
    if not controlled:
        assert nreg >= 1
    assert np.abs(got - O.run_ops_fast(sv0, ops)).max() < 1e-10


def test_register_program_words_are_consistent():
    n = 14
    gates = workloads.brickwork_su4(n, 8, seed=5)
    prog = desv.Program(n)
    for u, a, b in gates:
        prog.add_matrix(u, [a, b])
    prog.compile("complex64", fuse_bits=3, tile_bits=12, use_mfma=True, reg_groups=True)
    sw, sf, words, frags = prog.reg_program(False)
    tm, so, ops, _ = prog.plan()
    for s in range(len(tm)):
        if sw[s] < 0:
            continue
        w = words[sw[s]:]
        ng, nq, i, q = int(w[0]), int(w[1]), 2, 0
        for _ in range(ng):
            h = w[i:]
            lin = list(h[1:7]) + list(h[7:9])
            lout = list(h[25:31]) + list(h[7:9])
            # every layout places 8 distinct tile bits on lanes and waves; the
            # register slots enumerate the other 4
            assert len(set(lin)) == 8 and len(set(lout)) == 8
            regs_in = sorted({int(x) // 8 for x in h[9:25]})
            assert len(regs_in) == 16
            for u in range(int(h[0])):
                code = int(h[HDR + u])
                assert code & 15 in (RG_XLANE, RG_MFMA3)
                q += code & 15 == RG_MFMA3
            i += HDR + int(h[0])
        assert q == nq
        assert sf[s] + nq <= len(frags)


def test_reg_groups_option_rules():
    p = desv.Program(12)
    p.add_matrix(np.eye(4), [0, 1])
    with pytest.raises(Exception):
        p.compile("complex64", use_mfma=False, reg_groups=True)
    with pytest.raises(Exception):
        p.compile("complex128", reg_groups=True)


# ---- wave-tile stages (reg_groups=2): one tile per wave, HBM <-> registers --

def run_micro(V, mu, frags, fb, q):
    for code in mu:
        code = int(code)
        a1 = (code >> 4) & 15
        if code & 15 == RG_XLANE:
            xlane(V, a1, (code >> 8) & 15)
        else:
            fr = frags[q * fb].astype(np.float64)
            if fb == 2:
                fr = fr + frags[q * fb + 1].astype(np.float64)
            if code & 15 == RG_MFMA3:
                mfma3(V, a1, fr)
            else:
                assert code & 15 == RG_MFMA3T
                mfma3t(V, a1, (code >> 8) & 15, (code >> 12) & 15, fr.reshape(16, 16))
            q += 1
    return q


def mfma_4b(a, b):
    """v_mfma_f32_16x16x1_4b_f32 (one k-step, as measured on MI355X): four
    16x16 outer products, block bb from lanes 16 bb .. 16 bb + 15 of both
    operands (A_bb[m] = a[16 bb + m], B_bb[n] = b[16 bb + n]); D_bb[m][n] is
    returned at lane 16 (m // 4) + n, accumulator register 4 bb + m % 4."""
    D = np.zeros((64, 16))
    for bb in range(4):
        for m in range(16):
            for n in range(16):
                D[16 * (m // 4) + n, 4 * bb + m % 4] = a[16 * bb + m] * b[16 * bb + n]
    return D


def treg(x, y, z, nr, g, t):
    """Register holding target bits x, y, z = t's bits 0, 1, 2 in group g
    (g's bits fill the other register bits in increasing order)."""
    j, gi = 0, 0
    for b in range(nr):
        if b == x:
            j |= (t & 1) << b
        elif b == y:
            j |= ((t >> 1) & 1) << b
        elif b == z:
            j |= ((t >> 2) & 1) << b
        else:
            j |= ((g >> gi) & 1) << b
            gi += 1
    return j


# mfma_4b as index arrays: output (lane, register) of block bb, row m, column
# n, and the operand lanes it reads
_BB, _M, _N = np.meshgrid(np.arange(4), np.arange(16), np.arange(16), indexing="ij")
_OUT_L, _OUT_R = 16 * (_M // 4) + _N, 4 * _BB + _M % 4
_IN_A, _IN_B = 16 * _BB + _M, 16 * _BB + _N


def mfma_4b_chain(Ak, Bk):
    """sum_k mfma_4b(Ak[:, k], Bk[:, k]) (Ak, Bk: (64, K) operand lanes)."""
    D = np.zeros((64, 16))
    D[_OUT_L, _OUT_R] = np.einsum("...k,...k->...", Ak[_IN_A], Bk[_IN_B])
    return D


def test_mfma_4b_chain_matches_the_instruction(rng):
    Ak, Bk = rng.standard_normal((64, 5)), rng.standard_normal((64, 5))
    want = sum(mfma_4b(Ak[:, k], Bk[:, k]) for k in range(5))
    assert np.allclose(mfma_4b_chain(Ak, Bk), want, rtol=0, atol=1e-12)


def mfma3t(V, x, y, z, frag):
    """3-bit block in the 4-block form: per group of 8 registers, k-step k
    takes float k & 1 of register treg(k >> 1) as every lane's B value and
    frag[lane % 16][k] as its A value; accumulator register r is written to
    float r & 1 of register treg(r >> 1)."""
    nr = int(np.log2(V.shape[1]))
    assert x < y < z < nr
    Ak = frag[LANE % 16, :]  # (64, 16): lane's A value at k-step k
    for g in range(V.shape[1] // 8):
        regs = [treg(x, y, z, nr, g, t) for t in range(8)]
        Bk = np.empty((64, 16))
        for t in range(8):
            Bk[:, 2 * t], Bk[:, 2 * t + 1] = V[:, regs[t]].real, V[:, regs[t]].imag
        D = mfma_4b_chain(Ak, Bk)
        for t in range(8):
            V[:, regs[t]] = D[:, 2 * t] + 1j * D[:, 2 * t + 1]


def wave_stages(prog, compensated, fallback=False):
    """The stages a wave-tile program runs, as (tile mask, stage_ops row,
    word offset, first fragment block, store mask or 0, omap): the
    relabelled tail (the default, with a second state buffer) or its in-place
    form (``fallback``)."""
    tm, so, _, _ = prog.plan()
    sw, sf, _, _ = prog.reg_program(compensated)
    t = prog.tail(compensated)
    out = []
    head = len(tm) if (t["fb_from"] < 0 or not fallback) else t["fb_from"]
    for s in range(head):
        out.append((int(tm[s]), so[s], int(sw[s]), int(sf[s]), int(t["store_masks"][s]), t["omaps"][s]))
    if fallback and t["fb_from"] >= 0:
        for k in range(len(t["fb_tile_masks"])):
            out.append((int(t["fb_tile_masks"][k]), t["fb_stage_ops"][k], int(t["fb_stage_word"][k]),
                        int(t["fb_stage_frag"][k]), 0, None))
    return out


def replay_wave(prog, sv, n, compensated, fallback=False):
    from test_tile_layout import run_stage
    _, _, ops, mats = prog.plan()
    _, _, words, frags = prog.reg_program(compensated)
    fb = 2 if compensated else 1
    nwave = 0
    for tmask, srow, sw, sf, smask, omap in wave_stages(prog, compensated, fallback):
        if sw < 0:
            assert smask == 0  # a relabelling stage always runs as a wave tile
            sv = run_stage(sv, n, tmask, srow, ops, mats, 2)
            continue
        nwave += 1
        m = bin(tmask).count("1")
        nr = m - 6
        w = words[sw:]
        nmu, nq = int(w[0]), int(w[1])
        gl, gs = [int(x) for x in w[2:2 + m]], [int(x) for x in w[2 + m:2 + 2 * m]]
        mu = w[2 + 2 * m:2 + 2 * m + nmu]
        flag = int(w[2 + 2 * m + nmu])
        assert flag == (1 if smask else 0)
        if smask:
            om = [int(x) for x in w[3 + 2 * m + nmu:3 + 2 * m + nmu + n - m]]
            assert om == [int(x) for x in omap[:n - m]] and all(x == -1 for x in omap[n - m:])
            # every bit goes somewhere exactly once: tile bits to the store mask
            assert sorted(om + [b for b in range(n) if (smask >> b) & 1]) == list(range(n))
            assert sorted(gs) == [b for b in range(n) if (smask >> b) & 1]
        # coalescing constraints of the kernel's 16-byte accesses
        assert gl[6] == 0 and gs[6] == 0
        assert sorted(gl[:4]) == [1, 2, 3, 4] and sorted(gs[:4]) == [1, 2, 3, 4]
        free = [b for b in range(n) if not (tmask >> b) & 1]

        def index(gb):
            idx = np.zeros((64, 1 << nr), dtype=np.int64)
            for i in range(6):
                idx |= ((np.arange(64)[:, None] >> i) & 1) << gb[i]
            for r in range(nr):
                idx |= ((np.arange(1 << nr)[None, :] >> r) & 1) << gb[6 + r]
            return idx
        il, is_ = index(gl), index(gs)
        out = sv.copy() if not smask else np.full_like(sv, np.nan)
        for t in range(1 << (n - m)):
            base = sbase = 0
            for i, b in enumerate(free):
                if (t >> i) & 1:
                    base |= 1 << b
                    sbase |= 1 << (om[i] if smask else b)
            V = sv[base | il].astype(np.complex128)
            q = run_micro(V, mu, frags[sf:], fb, 0)
            assert q == nq
            out[sbase | is_] = V
        assert not np.isnan(out.real).any()
        sv = out
    return sv, nwave


@pytest.mark.parametrize("n,depth,tile,fuse", [(13, 6, 11, 3), (14, 5, 12, 3), (13, 6, 12, 2), (15, 4, 11, 3)])
def test_brickwork_wave_stages_replay(n, depth, tile, fuse, rng):
    gates = workloads.brickwork_su4(n, depth, seed=11)
    prog = desv.Program(n)
    for u, a, b in gates:
        prog.add_matrix(u, [a, b])
    prog.compile("complex64", fuse_bits=fuse, tile_bits=tile, use_mfma=True, reg_groups=2)
    sv0 = _rand_state(n, rng)
    want = sv0.copy()
    for u, a, b in gates:
        want = O.apply_matrix(want, u, [a, b])
    for fallback in (False, True):
        got, nw = replay_wave(prog, sv0, n, compensated=True, fallback=fallback)
        assert nw == len(wave_stages(prog, True, fallback))  # every stage, restore passes included
        assert np.abs(got - want).max() < 1e-10


@pytest.mark.parametrize("fuse,controlled", [(2, False), (3, True)])
def test_mixed_blocks_wave_stages_replay(fuse, controlled, rng):
    n = 13
    ops = _mixed_ops(n, 120, rng, controlled)
    prog = desv.Program(n).add_ops(ops)
    prog.compile("complex64", fuse_bits=fuse, tile_bits=11, use_mfma=True, reg_groups=2)
    sv0 = _rand_state(n, rng)
    want = O.run_ops_fast(sv0, ops)
    for fallback in (False, True):
        got, nw = replay_wave(prog, sv0, n, compensated=True, fallback=fallback)
        if not controlled:
            assert nw >= 1
        assert np.abs(got - want).max() < 1e-10


@pytest.mark.parametrize("n,count,tile,fuse", [(14, 120, 12, 3), (16, 120, 12, 2)])
def test_relabelled_tail_replay(n, count, tile, fuse):
    """Programs whose in-place form needs two or more restore passes end with
    one out-of-place stage that writes every qubit to its home bit (the
    qubits outside its tile move with the tile addresses); both that tail and
    the kept in-place form replay to the oracle's state."""
    rng = np.random.default_rng(n * 1000 + count)
    ops = _mixed_ops(n, count, rng, False)
    prog = desv.Program(n).add_ops(ops)
    prog.compile("complex64", fuse_bits=fuse, tile_bits=tile, use_mfma=True, reg_groups=2)
    t = prog.tail(True)
    assert t["fb_from"] >= 0 and len(t["fb_tile_masks"]) >= 2
    assert (t["store_masks"] != 0).sum() == 1 and t["store_masks"][-1] != 0
    assert len(wave_stages(prog, True, False)) < len(wave_stages(prog, True, True))
    sv0 = _rand_state(n, rng)
    want = O.run_ops_fast(sv0, ops)
    for fallback in (False, True):
        got, _ = replay_wave(prog, sv0, n, compensated=True, fallback=fallback)
        assert np.abs(got - want).max() < 1e-10
    # without wave tiles (or with register groups) no stage relabels
    for rg in (0, 1):
        prog.compile("complex64", fuse_bits=fuse, tile_bits=tile, use_mfma=True, reg_groups=rg)
        assert prog.tail()["fb_from"] == -1


def test_wave_programs_take_the_cheapest_planner_pass():
    """Wave-tile programs are compiled by the cheapest of the planner's
    deterministic pass and 64 randomised ones (fixed seeds): the same program
    always gets the same plan, and the 30-qubit depth-24 brickwork needs at
    most the deterministic pass's 13 sweeps (12 compute stages and one
    relabelling pass)."""
    gates = workloads.brickwork_su4(30, 24, seed=0)
    prog = desv.Program(30)
    for u, a, b in gates:
        prog.add_matrix(u, [a, b])
    dumps = []
    for _ in range(2):
        prog.compile("complex64", fuse_bits=3, tile_bits=12, use_mfma=True, reg_groups=2)
        tm, so, ops, _ = prog.plan()
        dumps.append((tm.copy(), so.copy(), ops.copy()))
    assert all(np.array_equal(a, b) for a, b in zip(*dumps))
    st = prog.stats()
    assert st["passes"] <= 13 and st["blocks"] <= 184
    assert st["passes_in_place"] >= st["passes"]
