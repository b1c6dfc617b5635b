"""The stage kernel's tile addressing, replayed on the CPU.

`k_stage` (dense-evolution_amd/csrc/engine_kernels.hip) keeps a tile in LDS
under an XOR swizzle (tile index i at element i ^ ((i >> 5) & 31)), lets each
lane own the groups tid + 256 j of an op (a group = the amplitudes an op mixes,
enumerated over the tile bits that are neither targets nor controls), builds
every address as the XOR of a per-lane slot and wave-uniform slots, and moves
tiles between HBM and LDS in chunks of 256 * (amplitudes per 16-byte access)
through per-chunk offset tables, with the write-back permutation applied on
the way out.  This module restates that addressing lane by lane in Python and
replays compiled plans with it; the result must equal the oracle.  It checks
the addressing design (a wrong slot, group split, chunk table or permutation
shows up as a wrong amplitude); the kernels themselves are checked on the GPU.
"""
import numpy as np
import pytest

import desv
from oracle import sv_oracle as O

KT = 256
OPK_DENSE1, OPK_DENSE2, OPK_DENSEK, OPK_DIAG, OPK_SCALAR, OPK_GRAD, OPK_GRADDIAG, OPK_MFMA = range(8)


def slot(i):
    return i ^ ((i >> 5) & 31)


def insert_zero(i, p):
    return ((i >> p) << (p + 1)) | (i & ((1 << p) - 1))


def spread(x, fixed):
    f = fixed
    while f:
        p = (f & -f).bit_length() - 1
        x = insert_zero(x, p)
        f &= f - 1
    return x


def deposit(x, mask):
    r = 0
    f = mask
    while f:
        if x & 1:
            r |= f & -f
        f &= f - 1
        x >>= 1
    return r


def groups(m, fixed):
    G = 1 << (m - bin(fixed).count('1'))
    J = G // KT if G > KT else 1
    return G, J


def run_stage(sv, n, tm, so, ops, mats, PV):
    tbits = [b for b in range(n) if (tm >> b) & 1]
    m = len(tbits)
    L = 0
    while L < n and (tm >> L) & 1:
        L += 1
    hbits = [b for b in tbits if b >= L]
    b_, e_ = int(so[0]), int(so[1])
    sperm = [int(x) for x in so[2:2 + m]]
    permuted = any(sperm[j] != j for j in range(m))
    size = 1 << m
    CH = KT * PV
    nch = size // CH if size >= CH else 1

    def goff(e):
        o = e & ((1 << L) - 1)
        r, j = e >> L, 0
        while r:
            if r & 1:
                o |= 1 << hbits[j]
            r >>= 1
            j += 1
        return o

    def perm(e):
        o, j = 0, 0
        while e:
            if e & 1:
                o |= 1 << sperm[j]
            e >>= 1
            j += 1
        return o
    jtab = [goff(j * CH) for j in range(nch)]
    qtab = [slot(perm(j * CH) if permuted else j * CH) for j in range(nch)]
    free = [b for b in range(n) if not (tm >> b) & 1]
    out = sv.copy()
    for t in range(1 << (n - m)):
        base = 0
        for i, b in enumerate(free):
            if (t >> i) & 1:
                base |= 1 << b
        lds = np.zeros(size, dtype=np.complex128)
        for tid in range(KT):
            le = tid * PV
            if le >= size:
                continue
            gl = base | goff(le)
            sl = slot(le)
            for j in range(nch):
                off = gl | jtab[j]
                S = slot(j * CH)
                for i in range(PV):
                    lds[(sl ^ S) ^ i] = sv[off + i]
        for o in range(b_, e_):
            op = ops[o]
            kind, k, mat = int(op[0]), int(op[1]), int(op[2])
            pos = [int(x) for x in op[4:9]]
            cl, cg = int(op[9]), int(op[10])
            if (base & cg) != cg:
                continue
            new = lds.copy()
            if kind in (OPK_DENSE1, OPK_DENSE2, OPK_DENSEK, OPK_MFMA):
                D = 1 << k
                M = mats[mat:mat + D * D].reshape(D, D)
                fixed = cl
                for j in range(k):
                    fixed |= 1 << pos[j]
                soff = []
                for s in range(D):
                    oo = cl
                    for j in range(k):
                        if (s >> j) & 1:
                            oo |= 1 << pos[j]
                    soff.append(slot(oo))
                G, J = groups(m, fixed)
                for tid in range(min(G, KT)):
                    P = slot(spread(tid, fixed))
                    for j in range(J):
                        A = P ^ slot(spread(j * KT, fixed))
                        a = np.array([lds[A ^ soff[s]] for s in range(D)])
                        y = M @ a
                        for r in range(D):
                            new[A ^ soff[r]] = y[r]
            elif kind == OPK_DIAG:
                gsub = gmask = lmask = 0
                for j in range(k):
                    pj = pos[j]
                    if pj < 0:
                        gmask |= 1 << j
                        if (base >> (-1 - pj)) & 1:
                            gsub |= 1 << j
                    else:
                        lmask |= 1 << pj
                G, J = groups(m, lmask | cl)
                for tid in range(min(G, KT)):
                    P = slot(spread(tid, lmask | cl))
                    for j in range(J):
                        A = P ^ slot(spread(j * KT, lmask | cl))
                        for d in range(1 << k):
                            if (d & gmask) != gsub:
                                continue
                            oo = cl
                            for q in range(k):
                                if pos[q] >= 0 and (d >> q) & 1:
                                    oo |= 1 << pos[q]
                            oo = slot(oo)
                            new[A ^ oo] = mats[mat + d] * lds[A ^ oo]
            elif kind == OPK_SCALAR:
                new = lds * mats[mat]
            else:
                raise AssertionError(kind)
            lds = new
        for tid in range(KT):
            le = tid * PV
            if le >= size:
                continue
            gl = base | goff(le)
            ql = [slot(perm(le + i) if permuted else le + i) for i in range(PV)]
            for j in range(nch):
                off = gl | jtab[j]
                for i in range(PV):
                    out[off + i] = lds[ql[i] ^ qtab[j]]
    return out


def emulate(plan, sv, n, PV):
    tm, so, ops, mats = plan
    for s in range(len(tm)):
        sv = run_stage(sv, n, int(tm[s]), so[s], ops, mats, PV)
    return sv




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


@pytest.mark.parametrize("n,dtype,opts", [
    (9, "complex128", dict(tile_bits=7, low_bits=3)),
    (10, "complex128", dict(fuse_bits=3, tile_bits=8, low_bits=2)),
    (10, "complex64", dict(tile_bits=8, low_bits=3)),
    (10, "complex64", dict(fuse_bits=5, tile_bits=9, low_bits=2, use_mfma=True)),
    (8, "complex64", dict(tile_bits=8)),
])
def test_swizzled_tile_addressing_replays_plans(n, dtype, opts, rng):
    ops = random_ops(n, 80, rng)
    sv0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    sv0 /= np.linalg.norm(sv0)
    prog = desv.Program(n).add_ops(ops).compile(dtype, **opts)
    tm, so, pops, mats = prog.plan()
    if opts.get("tile_bits", n) < n:
        assert len(tm) > 1
    got = emulate(prog.plan(), sv0, n, 2 if dtype == "complex64" else 1)
    assert np.allclose(got, O.run_ops_fast(sv0, ops), atol=1e-12)


def test_slot_is_linear_and_an_involution():
    i = np.arange(1 << 14)
    s = i ^ ((i >> 5) & 31)
    assert np.array_equal(s ^ ((s >> 5) & 31), i)
    a, b = np.meshgrid(np.arange(0, 1 << 14, 37), np.arange(0, 1 << 14, 91))
    sl = lambda x: x ^ ((x >> 5) & 31)
    assert np.array_equal(sl(a ^ b), sl(a) ^ sl(b))
    # lanes stepping by 2^s, s <= 5, land on 32 distinct 8-byte bank pairs
    for st in range(6):
        lanes = sl(np.arange(32) << st)
        assert len(set((lanes % 32).tolist())) == 32
