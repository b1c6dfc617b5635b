"""The complex64 streaming kernels' addressing, replayed on the CPU.

`k_direct_b0` (a gate whose targets or controls include index bit 0) lets a
thread own float4 pairs (x, x | 1) at x | off(c) for the c-th combination of
the other targets, and `k_direct_lane` (targets in index bits 1..6) lets lane l
of wave w, step s own the float4 f = 256 w + 64 s + l and exchanges group
members between lanes with xor shuffles (dense-evolution_amd/csrc/
engine_kernels.hip).  This module restates both index maps thread by thread
(lane by lane) and applies them to a random state for every placement of one
or two targets and of controls (bit 0 among them), in place and out of place:
every amplitude must be owned exactly once (out of place) or at most once (in
place), and the result must equal a dense application of the gate.  It checks
the addressing design; the kernels themselves are checked on the GPU.
"""
import itertools

import numpy as np
import pytest


def dense_apply(sv, pos, cm, M):
    """Reference: M on index bits `pos` (matrix bit j <-> pos[j]) where all
    bits of the control mask `cm` are set."""
    out = sv.copy()
    K = len(pos)
    for x in range(len(sv)):
        if any(x >> p & 1 for p in pos) or (x & cm) != cm:
            continue
        idx = [x | sum(((s >> j) & 1) << pos[j] for j in range(K)) for s in range(1 << K)]
        out[idx] = M @ sv[idx]
    return out


def pair_kernel(sv, n, pos, cm, M, oop):
    """k_direct_b0: J0 = matrix bit of index bit 0 (-1: bit 0 is a control)."""
    K, D = len(pos), 1 << len(pos)
    J0 = pos.index(0) if 0 in pos else -1
    NH = K if J0 < 0 else K - 1
    mbit = [i + 1 if (J0 >= 0 and i >= J0) else i for i in range(NH)]
    off = [sum(((c >> i) & 1) << pos[mbit[i]] for i in range(NH)) for c in range(1 << NH)]
    outer = sorted(pos[mbit[i]] for i in range(NH))
    dst = np.full_like(sv, np.nan) if oop else sv.copy()
    seen = np.zeros(len(sv), dtype=int)
    for u in range(1 << (n - 1 - NH)):
        y = u
        for p in outer:  # zeros at the outer bits of the pair index
            y = ((y >> (p - 1)) << p) | (y & ((1 << (p - 1)) - 1))
        x = y << 1
        act = (x & cm) == cm if J0 >= 0 else ((x | 1) & cm) == cm
        if not oop and not act:
            continue
        a = [[sv[x | o], sv[x | o | 1]] for o in off]
        r = [list(v) for v in a]
        if act:
            if J0 >= 0:
                vec = np.zeros(D, dtype=complex)
                for c in range(1 << NH):
                    for h in range(2):
                        q = (h << J0) | sum(((c >> i) & 1) << mbit[i] for i in range(NH))
                        vec[q] = a[c][h]
                res = M @ vec
                for c in range(1 << NH):
                    for h in range(2):
                        r[c][h] = res[(h << J0) | sum(((c >> i) & 1) << mbit[i] for i in range(NH))]
            else:
                res = M @ np.array([a[c][1] for c in range(1 << NH)])
                for c in range(1 << NH):
                    r[c][1] = res[c]
        for c, o in enumerate(off):
            dst[x | o], dst[x | o | 1] = r[c]
            seen[x | o] += 1
            seen[x | o | 1] += 1
    return dst, seen


def lane_kernel(sv, n, pos, cm, M, oop, lane_shift=1):
    """k_direct_lane: waves of 4 x 64 float4s, members exchanged by xor;
    index bit p is lane bit p - lane_shift (1 in the kernel)."""
    K, D, S = len(pos), 1 << len(pos), 4
    lanes = np.arange(64)
    m = sum(((lanes >> (pos[j] - lane_shift)) & 1) << j for j in range(K))
    dep = [sum(((d >> j) & 1) << (pos[j] - lane_shift) for j in range(K)) for d in range(D)]
    dst = np.full_like(sv, np.nan) if oop else sv.copy()
    seen = np.zeros(len(sv), dtype=int)
    nvec4 = len(sv) // 2
    for f0 in range(0, nvec4, S * 64):
        for s in range(S):
            f = f0 + s * 64 + lanes
            a = np.stack([sv[2 * f], sv[2 * f + 1]], axis=1)
            e0 = 2 * f
            c0, c1 = (e0 & cm) == cm, ((e0 | 1) & cm) == cm
            acc = a * M[m, m][:, None]
            for d in range(1, D):
                acc = acc + M[m, m ^ d][:, None] * a[lanes ^ dep[d]]
            r = a.copy()
            r[c0, 0] = acc[c0, 0]
            r[c1, 1] = acc[c1, 1]
            st = np.ones(64, dtype=bool) if oop else (c0 | c1)
            dst[2 * f[st]], dst[2 * f[st] + 1] = r[st, 0], r[st, 1]
            seen[2 * f[st]] += 1
            seen[2 * f[st] + 1] += 1
    return dst, seen


def _cases(n, targets, rng):
    for K in (1, 2):
        for pos in itertools.permutations(targets, K):
            free = [b for b in range(n) if b not in pos]
            for cbits in [()] + [(b,) for b in free[:3]] + [tuple(free[:2])]:
                cm = sum(1 << b for b in cbits)
                M = rng.normal(size=(1 << K, 1 << K)) + 1j * rng.normal(size=(1 << K, 1 << K))
                yield list(pos), cm, M


@pytest.mark.parametrize("oop", [False, True])
def test_pair_kernel_addressing(oop):
    rng = np.random.default_rng(3)
    n = 6
    count = 0
    for pos, cm, M in _cases(n, range(n), rng):
        if 0 not in pos and not cm & 1:
            continue  # bit 0 neither target nor control: not this kernel's case
        sv = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        got, seen = pair_kernel(sv, n, pos, cm, M, oop)
        assert (seen == 1).all() if oop else seen.max() <= 1
        assert np.allclose(got, dense_apply(sv, pos, cm, M)), (pos, cm)
        count += 1
    assert count > 40


@pytest.mark.parametrize("oop", [False, True])
def test_lane_kernel_addressing(oop):
    rng = np.random.default_rng(4)
    n = 9  # 256 float4s: whole waves
    count = 0
    for pos, cm, M in _cases(n, range(1, 7), rng):
        sv = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        got, seen = lane_kernel(sv, n, pos, cm, M, oop)
        assert (seen == 1).all() if oop else seen.max() <= 1
        assert np.allclose(got, dense_apply(sv, pos, cm, M)), (pos, cm)
        count += 1
    assert count > 100


def test_a_misplaced_exchange_is_caught():
    """The replay is sensitive: taking index bit p as lane bit p instead of
    p - 1 (the float4 holds the pair along bit 0) gives a wrong state."""
    rng = np.random.default_rng(5)
    n, pos, cm = 9, [2, 4], 0
    M = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    sv = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    got, _ = lane_kernel(sv, n, pos, cm, M, True, lane_shift=0)
    assert not np.allclose(got, dense_apply(sv, pos, cm, M))
    good, _ = lane_kernel(sv, n, pos, cm, M, True)
    assert np.allclose(good, dense_apply(sv, pos, cm, M))
