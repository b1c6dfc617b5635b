"""NumPy restatement of the reference dense-statevector semantics.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): this module is the
checker the HIP engine is compared against.  It is written independently of
the product package (it imports nothing from it) so that the two cannot share
a bug.

Conventions (reference ``backends/statevector.py:22-33``): the state is a flat
``(2**n,)`` complex vector, qubit ``q`` is bit ``n-1-q`` of the basis index
(MSB-first).  A 2-qubit matrix acting on ``(q1, q2)`` is indexed
``2*bit(q1) + bit(q2)`` (``statevector.py:188-203``).
"""
from __future__ import annotations

import numpy as np

INV2 = 1.0 / np.sqrt(2.0)

# ---------------------------------------------------------------------------
# gate tables
# ---------------------------------------------------------------------------

#: name -> id, restating ``circuits/gates.py:53-63``.
GATE_IDS = {
    'id': 0, 'h': 1, 'x': 2, 'y': 3, 'z': 4, 's': 5, 'sdg': 6, 't': 7, 'tdg': 8,
    'rx': 9, 'ry': 10, 'rz': 11, 'p': 12, 'u1': 12, 'phase': 12, 'sx': 13,
    'gphase': 14, 'cx': 20, 'cz': 21, 'cp': 22, 'cphase': 22, 'cy': 24, 'crz': 25,
}


def gate_1q_matrix(g_id: int, param: float) -> np.ndarray:
    """2x2 matrix for 1-qubit gate id 0..14 in complex128.

    Restates ``circuits/compiler.py:52-134`` (``_gate_1q_matrix``): the
    rotation gates use half angles, P/U1 is diag(1, e^{i p}), GPhase is
    e^{i p} * I.
    """
    c, s = np.cos(param / 2.0), np.sin(param / 2.0)
    e = np.exp(1j * param)
    table = {
        0: [[1, 0], [0, 1]],
        1: [[INV2, INV2], [INV2, -INV2]],
        2: [[0, 1], [1, 0]],
        3: [[0, -1j], [1j, 0]],
        4: [[1, 0], [0, -1]],
        5: [[1, 0], [0, 1j]],
        6: [[1, 0], [0, -1j]],
        7: [[1, 0], [0, np.exp(1j * np.pi / 4)]],
        8: [[1, 0], [0, np.exp(-1j * np.pi / 4)]],
        9: [[c, -1j * s], [-1j * s, c]],
        10: [[c, -s], [s, c]],
        11: [[np.exp(-0.5j * param), 0], [0, np.exp(0.5j * param)]],
        12: [[1, 0], [0, e]],
        13: [[0.5 + 0.5j, 0.5 - 0.5j], [0.5 - 0.5j, 0.5 + 0.5j]],
        14: [[e, 0], [0, e]],
    }
    return np.array(table[int(g_id)], dtype=np.complex128)


def gate_2q_matrix(g_id: int, param: float) -> np.ndarray:
    """4x4 matrix (index 2*bit(ctrl)+bit(tgt)) of 2-qubit gate ids 20..25.

    Restates the branch bodies of ``compiler.py:235-283``: CX flips the
    target when the control is set, CZ negates |11>, CP multiplies |11> by
    e^{i p}, SWAP exchanges |01>/|10>, CY applies Y under the control (+-i
    phases, ``:257-261``), CRZ applies Rz(p) to the target under the control
    (``:270-272``).
    """
    m = np.eye(4, dtype=np.complex128)
    g = int(g_id)
    if g == 20:
        m[2:, 2:] = [[0, 1], [1, 0]]
    elif g == 21:
        m[3, 3] = -1
    elif g == 22:
        m[3, 3] = np.exp(1j * param)
    elif g == 23:
        m[1:3, 1:3] = [[0, 1], [1, 0]]
    elif g == 24:
        m[2:, 2:] = [[0, -1j], [1j, 0]]
    elif g == 25:
        m[2, 2] = np.exp(-0.5j * param)
        m[3, 3] = np.exp(0.5j * param)
    else:
        raise ValueError(f"unknown 2-qubit gate id {g}")
    return m


#: fixed matrices restating ``circuits/gates.py:33-51``.
GATES = {
    'h': gate_1q_matrix(1, 0), 'x': gate_1q_matrix(2, 0), 'y': gate_1q_matrix(3, 0),
    'z': gate_1q_matrix(4, 0), 's': gate_1q_matrix(5, 0), 'sdg': gate_1q_matrix(6, 0),
    't': gate_1q_matrix(7, 0), 'tdg': gate_1q_matrix(8, 0), 'sx': gate_1q_matrix(13, 0),
    'id': np.eye(2, dtype=np.complex128),
    'cx': gate_2q_matrix(20, 0), 'cz': gate_2q_matrix(21, 0), 'cy': gate_2q_matrix(24, 0),
    'swap': gate_2q_matrix(23, 0),
    'iswap': np.array([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]], dtype=np.complex128),
    'ecr': INV2 * np.array([[0, 0, 1, 1j], [0, 0, 1j, 1], [1, -1j, 0, 0], [-1j, 1, 0, 0]],
                           dtype=np.complex128),
}


def parametric_matrix(name: str, *args) -> np.ndarray:
    """Parametric matrices restating ``circuits/gates.py:78-91``."""
    if name in ('rx', 'ry', 'rz', 'p', 'u1', 'gphase'):
        return gate_1q_matrix(GATE_IDS[name], args[0])
    if name in ('cp', 'crz'):
        return gate_2q_matrix(GATE_IDS[name], args[0])
    if name == 'u3':
        th, ph, la = args
        return np.array([[np.cos(th / 2), -np.exp(1j * la) * np.sin(th / 2)],
                         [np.exp(1j * ph) * np.sin(th / 2), np.exp(1j * (ph + la)) * np.cos(th / 2)]])
    if name == 'u2':
        ph, la = args
        return INV2 * np.array([[1.0, -np.exp(1j * la)], [np.exp(1j * ph), np.exp(1j * (ph + la))]])
    raise ValueError(name)


# ---------------------------------------------------------------------------
# transpiler (decomposition only) -- compiler.py:362-488
# ---------------------------------------------------------------------------

def transpile(circuit):
    """Decompose ccx/swap/iswap/ecr/u3/u2; restates ``compiler.py:442-488``.

    ccx -> 15 gates (``:376-386``), swap -> 3 cx (``:391``), iswap ->
    s,h,s,cx,cx,h (``:401``), ecr -> s,sx,cx,x,gphase(-pi/4) (``:412``),
    u3 -> rz(lam), ry(theta), rz(phi), gphase((phi+lam)/2) (``:430-435``),
    u2(phi,lam) = u3(pi/2,phi,lam) (``:440``).  u2/u3 tuples without
    parameter values pass through (``:482-485``).
    """
    out = []
    for cmd in circuit:
        name = cmd[0].lower()
        a = cmd[1:]
        if name == 'ccx':
            c1, c2, t = a[:3]
            out += [('h', t), ('cx', c2, t), ('tdg', t), ('cx', c1, t), ('t', t), ('cx', c2, t),
                    ('tdg', t), ('cx', c1, t), ('t', c2), ('t', t), ('cx', c1, c2), ('h', t),
                    ('t', c1), ('tdg', c2), ('cx', c1, c2)]
        elif name == 'swap':
            q1, q2 = a[:2]
            out += [('cx', q1, q2), ('cx', q2, q1), ('cx', q1, q2)]
        elif name == 'iswap':
            q1, q2 = a[:2]
            out += [('s', q1), ('h', q1), ('s', q2), ('cx', q1, q2), ('cx', q2, q1), ('h', q2)]
        elif name == 'ecr':
            q1, q2 = a[:2]
            out += [('s', q1), ('sx', q2), ('cx', q1, q2), ('x', q1), ('gphase', q1, -np.pi / 4.0)]
        elif name == 'u3' and len(cmd) >= 5:
            q, th, ph, la = a[:4]
            out += [('rz', q, la), ('ry', q, th), ('rz', q, ph), ('gphase', q, (ph + la) / 2.0)]
        elif name == 'u2' and len(cmd) >= 4:
            q, ph, la = a[:3]
            th = np.pi / 2.0
            out += [('rz', q, la), ('ry', q, th), ('rz', q, ph), ('gphase', q, (ph + la) / 2.0)]
        else:
            out.append(cmd)
    return out


def encode_ops(circuit, n):
    """Op table ``[g_id, q1, q2, param]`` built by ``run_circuit_jit``.

    Restates ``statevector.py:414-464``: names resolved through GATE_IDS
    (unknown -> ValueError), qubits range-checked, 1-qubit parametric gates
    read their angle from args[1], cp/crz/cphase from args[2].
    """
    rows = []
    for cmd in transpile(circuit):
        name = str(cmd[0]).lower()
        if name not in GATE_IDS:
            raise ValueError(f"unknown gate '{cmd[0]}'")
        gid = GATE_IDS[name]
        a = cmd[1:]
        if name in ('rx', 'ry', 'rz', 'p', 'u1', 'phase', 'gphase'):
            row = [gid, a[0], 0, a[1] if len(a) > 1 else 0.0]
        elif name in ('cp', 'crz', 'cphase'):
            row = [gid, a[0], a[1] if len(a) > 1 else 0, a[2] if len(a) > 2 else 0.0]
        elif name in ('cx', 'cz', 'swap', 'cy'):
            row = [gid, a[0], a[1] if len(a) > 1 else 0, 0.0]
        else:
            row = [gid, a[0] if a else 0, 0, 0.0]
        for q in ([row[1], row[2]] if gid >= 20 else [row[1]]):
            if not 0 <= int(q) < n:
                raise ValueError(f"Qubit index {int(q)} out of range [0, {n})")
        rows.append([float(x) for x in row])
    return np.array(rows, dtype=np.float64).reshape(-1, 4)


# ---------------------------------------------------------------------------
# gate application
# ---------------------------------------------------------------------------

def zero_state(n, dtype=np.complex128):
    """|0...0>; restates ``statevector.py:101-107``."""
    sv = np.zeros(1 << n, dtype=dtype)
    sv[0] = 1.0
    return sv


def apply_op_gather(sv, op):
    """One step of the scanned per-gate kernel, gather formulation.

    Restates ``compiler.py:136-315``: matrices are computed in float64 and
    cast to the state dtype (``:59-67``, ``:172-179``); every amplitude i
    reads its partner i ^ stride through an index gather and the two
    branches are selected with ``where``.
    """
    dt = sv.dtype
    dim = sv.shape[0]
    n = dim.bit_length() - 1
    g, q1, q2, p = int(op[0]), int(op[1]), int(op[2]), float(op[3])
    idx = np.arange(dim, dtype=np.int64)
    if g <= 14:
        m = gate_1q_matrix(g, p).astype(dt)
        stride = np.int64(1) << (n - 1 - q1)
        mask0 = (idx & stride) == 0
        a_self = sv[idx]
        a_pair = sv[idx ^ stride]
        new0 = m[0, 0] * a_self + m[0, 1] * a_pair
        new1 = m[1, 0] * a_pair + m[1, 1] * a_self
        return np.where(mask0, new0, new1).astype(dt)
    cb = np.int64(1) << (n - 1 - q1)
    tb = np.int64(1) << (n - 1 - q2)
    cset = (idx & cb) != 0
    tset = (idx & tb) != 0
    if g == 20:
        return np.where(cset, sv[idx ^ tb], sv)
    if g == 21:
        return np.where(cset & tset, -sv, sv)
    if g == 22:
        e = dt.type(np.exp(1j * p))
        return np.where(cset & tset, e * sv, sv)
    if g == 23:
        sw = cset ^ tset
        return np.where(sw, sv[idx ^ cb ^ tb], sv)
    if g == 24:
        ph = np.where(tset, dt.type(1j), dt.type(-1j))
        return np.where(cset, ph * sv[idx ^ tb], sv)
    if g == 25:
        ph = np.where(tset, dt.type(np.exp(0.5j * p)), dt.type(np.exp(-0.5j * p)))
        return np.where(cset, ph * sv, sv)
    raise ValueError(f"unknown gate id {g}")


def run_ops(sv, ops):
    """``lax.scan`` of the per-gate kernel over an op table (``compiler.py:318-334``)."""
    for op in np.asarray(ops, dtype=np.float64).reshape(-1, 4):
        sv = apply_op_gather(sv, op)
    return sv


def apply_op_fast(sv, op):
    """Same map as :func:`apply_op_gather`, computed on reshaped views.

    The state is viewed as ``(2,)*n`` so a gate is a small contraction over
    its axes instead of a full-index gather; the matrices are the same
    float64 tables cast to the state dtype.  Pinned against the gather
    formulation in ``tests/test_oracle.py``; used to keep 20-qubit parity
    checks fast.
    """
    dt = sv.dtype
    n = sv.shape[0].bit_length() - 1
    g, q1, q2, p = int(op[0]), int(op[1]), int(op[2]), float(op[3])
    t = sv.reshape([2] * n)
    if g <= 14:
        m = gate_1q_matrix(g, p).astype(dt)
        a0 = t.take(0, axis=q1)
        a1 = t.take(1, axis=q1)
        out = np.stack([m[0, 0] * a0 + m[0, 1] * a1, m[1, 0] * a0 + m[1, 1] * a1], axis=q1)
        return out.reshape(-1).astype(dt)
    m = gate_2q_matrix(g, p).astype(dt)
    if q1 == q2:
        raise ValueError("Control and target qubits must differ")
    v = np.moveaxis(t, [q1, q2], [0, 1]).reshape(4, -1)
    r = (m @ v).reshape([2, 2] + [2] * (n - 2))
    return np.moveaxis(r, [0, 1], [q1, q2]).reshape(-1).astype(dt)


def run_ops_fast(sv, ops):
    for op in np.asarray(ops, dtype=np.float64).reshape(-1, 4):
        sv = apply_op_fast(sv, op)
    return sv


def apply_matrix(sv, mat, qubits):
    """Eager k-qubit contraction, the reference's ``apply_gate_1q/2q``.

    Restates ``statevector.py:167-203`` (moveaxis the target axes last,
    ``(N/2^k, 2^k) @ gate.T``, move back), generalised to k qubits with the
    first listed qubit as the most significant matrix-index bit.
    """
    n = sv.shape[0].bit_length() - 1
    k = len(qubits)
    if len(set(qubits)) != k:
        raise ValueError("qubits must differ")
    mat = np.asarray(mat, dtype=sv.dtype).reshape(1 << k, 1 << k)
    t = np.moveaxis(sv.reshape([2] * n), list(qubits), list(range(n - k, n)))
    shp = t.shape
    r = t.reshape(-1, 1 << k) @ mat.T
    return np.moveaxis(r.reshape(shp), list(range(n - k, n)), list(qubits)).reshape(-1)


def kron_operator(mat, qubits, n):
    """Brute-force full 2^n x 2^n operator of a k-qubit matrix (n <= 10).

    Builds the operator column by column from basis states using bit
    arithmetic only (independent of ``apply_matrix``'s axis shuffling).
    """
    k = len(qubits)
    dim = 1 << n
    op = np.zeros((dim, dim), dtype=np.complex128)
    pos = [n - 1 - q for q in qubits]
    for col in range(dim):
        sub = 0
        for j, p in enumerate(pos):
            sub |= ((col >> p) & 1) << (k - 1 - j)
        rest = col
        for p in pos:
            rest &= ~(1 << p)
        for r in range(1 << k):
            row = rest
            for j, p in enumerate(pos):
                row |= ((r >> (k - 1 - j)) & 1) << p
            op[row, col] += mat[r, sub]
    return op


def run_circuit(circuit, n, dtype=np.complex128, sv=None):
    """``DenseSVSimulator(n).run_circuit(circuit)`` through the table path."""
    if sv is None:
        sv = zero_state(n, dtype)
    ops = encode_ops(circuit, n)
    return run_ops(sv.astype(dtype), ops)


def fidelity(a, b):
    """|<a|b>|^2; restates ``utils/measurement.py:71-94``."""
    return float(abs(np.vdot(np.asarray(a, np.complex128), np.asarray(b, np.complex128))) ** 2)


def probabilities(sv):
    """|a|^2 clipped to [0,1] and renormalised; restates ``statevector.py:599-607``."""
    p = np.clip(np.abs(np.asarray(sv)) ** 2, 0.0, 1.0).astype(np.float64)
    s = p.sum()
    return p / s if s > 1e-12 else p


# ---------------------------------------------------------------------------
# Pauli strings -- physics/observables.py
# ---------------------------------------------------------------------------

PAULI = {
    'I': np.eye(2, dtype=np.complex128),
    'X': np.array([[0, 1], [1, 0]], dtype=np.complex128),
    'Y': np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
    'Z': np.array([[1, 0], [0, -1]], dtype=np.complex128),
}


def normalize_terms(pauli_terms, n_qubits=None):
    """str | dict | pairs -> {q: 'X'|'Y'|'Z'}; restates ``observables.py:38-60``."""
    if isinstance(pauli_terms, str):
        if n_qubits is not None and len(pauli_terms) != n_qubits:
            raise ValueError("length mismatch")
        d = {q: c.upper() for q, c in enumerate(pauli_terms) if c.upper() != 'I'}
    elif isinstance(pauli_terms, dict):
        d = {int(q): str(c).upper() for q, c in pauli_terms.items() if str(c).upper() != 'I'}
    else:
        d = {int(q): str(c).upper() for q, c in pauli_terms if str(c).upper() != 'I'}
    for q, c in d.items():
        if c not in 'XYZ' or len(c) != 1:
            raise ValueError(f"unknown Pauli operator {c!r}")
        if q < 0:
            raise ValueError("negative qubit")
    return d


def apply_pauli(sv, terms, n):
    """P|psi> by flip mask + source-bit phase; restates ``observables.py:63-94``."""
    if not terms:
        return np.asarray(sv, dtype=np.complex128).copy()
    flip = 0
    for q, c in terms.items():
        if c in 'XY':
            flip |= 1 << (n - 1 - q)
    idx = np.arange(sv.shape[0])
    src = idx ^ flip
    coeff = np.ones(sv.shape[0], dtype=np.complex128)
    for q, c in terms.items():
        bit = (src >> (n - 1 - q)) & 1
        if c == 'Y':
            coeff = coeff * np.where(bit == 0, 1j, -1j)
        elif c == 'Z':
            coeff = coeff * np.where(bit == 0, 1.0, -1.0)
    return np.asarray(sv)[src] * coeff


def pauli_expectation(sv, pauli_terms, n_qubits=None):
    """<psi|P|psi> accumulated in complex128 (``observables.py:97-149``)."""
    sv = np.asarray(sv)
    n = sv.shape[0].bit_length() - 1
    t = normalize_terms(pauli_terms, n_qubits)
    if t and max(t) >= n:
        raise ValueError("qubit out of range")
    return float(np.real(np.vdot(sv.astype(np.complex128), apply_pauli(sv.astype(np.complex128), t, n))))


def pauli_sum_expectation(sv, terms, n_qubits=None):
    """sum_k c_k <P_k>; restates ``observables.py:152-186``."""
    total = 0.0
    for c, p in terms:
        total += c * pauli_expectation(sv, p, n_qubits)
    return total


def pauli_sum_matvec(v, terms):
    """H v matrix-free; restates ``observables.py:250-308``."""
    v = np.asarray(v, dtype=np.complex128)
    n = v.shape[0].bit_length() - 1
    out = np.zeros_like(v)
    for c, p in terms:
        out += c * apply_pauli(v, normalize_terms(p), n)
    return out


def pauli_matrix(terms, n):
    """Dense H = sum c_k kron(P_k); restates ``observables.py:189-247``."""
    dim = 1 << n
    H = np.zeros((dim, dim), dtype=np.complex128)
    for c, p in terms:
        d = normalize_terms(p, n)
        m = np.array([[1.0 + 0j]])
        for q in range(n):
            m = np.kron(m, PAULI[d.get(q, 'I')])
        H += c * m
    return H


# ---------------------------------------------------------------------------
# energy and gradients -- solvers/autodiff.py, backends/variational_numpy.py
# ---------------------------------------------------------------------------

#: gates that consume a theta entry (``autodiff.py:21``).
AD_PARAMETRIC = ('rx', 'ry', 'rz', 'u1', 'p', 'cp', 'crz')
AD_TWO_QUBIT = ('cx', 'cy', 'cz', 'cp', 'crz', 'swap')


def build_template(op_dicts, n):
    """``[g_id, q1, q2, sentinel]`` rows of ``_build_template`` (``autodiff.py:33-89``).

    Only names and qubits are read; literal angles are ignored; ops touching
    a qubit >= n are dropped (``:54-55``); sentinel -1 marks a theta slot.
    """
    tuples = []
    for op in op_dicts:
        name = str(op['name']).lower().strip()
        qs = [int(q) for q in op.get('qubits', [])]
        if not qs or any(q >= n for q in qs):
            continue
        tuples.append((name, *qs))
    rows = []
    for cmd in transpile(tuples):
        name = cmd[0].lower()
        if name not in GATE_IDS:
            raise ValueError(f"gate '{name}' has no GATE_IDS entry")
        sent = -1.0 if name in AD_PARAMETRIC else 0.0
        qs = cmd[1:]
        if name in AD_TWO_QUBIT and len(qs) >= 2:
            rows.append([GATE_IDS[name], qs[0], qs[1], sent])
        elif qs:
            rows.append([GATE_IDS[name], qs[0], 0, sent])
    return np.array(rows, dtype=np.float64).reshape(-1, 4)


def patch_template(template, theta):
    """Inject theta into sentinel rows in order (``autodiff.py:150-157``)."""
    ops = np.array(template, dtype=np.float64, copy=True)
    k = 0
    for r in ops:
        if r[3] == -1.0:
            r[3] = theta[k]
            k += 1
    return ops


def energy(ops, hterms, n, sv0=None):
    """E = Re <psi|H|psi> with psi = U(ops)|sv0>, complex128 (``autodiff.py:136-167``)."""
    sv = zero_state(n) if sv0 is None else np.asarray(sv0, np.complex128)
    sv = run_ops_fast(sv, ops)
    return float(np.real(np.vdot(sv, pauli_sum_matvec(sv, hterms)))), sv


def _op_unitary_on(sv, op, adjoint=False):
    """Apply op (or its inverse) with explicit matrices (complex128)."""
    g, q1, q2, p = int(op[0]), int(op[1]), int(op[2]), float(op[3])
    if g <= 14:
        m = gate_1q_matrix(g, p)
        return apply_matrix(sv, m.conj().T if adjoint else m, [q1])
    m = gate_2q_matrix(g, p)
    return apply_matrix(sv, m.conj().T if adjoint else m, [q1, q2])


def _generator_apply(sv, op):
    """G psi with dU/dtheta = G U for the theta-carrying gate ids."""
    g, q1, q2 = int(op[0]), int(op[1]), int(op[2])
    if g == 9:
        return apply_matrix(sv, -0.5j * PAULI['X'], [q1])
    if g == 10:
        return apply_matrix(sv, -0.5j * PAULI['Y'], [q1])
    if g == 11:
        return apply_matrix(sv, -0.5j * PAULI['Z'], [q1])
    if g == 12:
        return apply_matrix(sv, np.diag([0, 1j]), [q1])
    if g == 22:
        return apply_matrix(sv, np.diag([0, 0, 0, 1j]), [q1, q2])
    if g == 25:
        return apply_matrix(sv, np.diag([0, 0, -0.5j, 0.5j]), [q1, q2])
    raise ValueError(f"gate id {g} carries no theta")


def adjoint_gradient(template, theta, hterms, n):
    """Energy and full gradient in one reverse sweep (c128 reference).

    Not in the reference (which uses JAX reverse mode through the scan,
    ``autodiff.py:136-167``); mathematically the same derivative:
    dE/dtheta_k = 2 Re <lambda_k| G_k |psi_k>, with psi_k the state after
    gate k and lambda_k = U_{k+1}^dag ... U_G^dag H psi.
    """
    ops = patch_template(template, theta)
    e, psi = energy(ops, hterms, n)
    lam = pauli_sum_matvec(psi, hterms)
    grad = np.zeros(len(theta))
    slots = [i for i, r in enumerate(template) if r[3] == -1.0]
    slot_of = {row: k for k, row in enumerate(slots)}
    for i in range(len(ops) - 1, -1, -1):
        if i in slot_of:
            grad[slot_of[i]] = 2.0 * np.real(np.vdot(lam, _generator_apply(psi, ops[i])))
        psi = _op_unitary_on(psi, ops[i], adjoint=True)
        lam = _op_unitary_on(lam, ops[i], adjoint=True)
    return e, grad


def param_shift_gradient(template, theta, hterms, n):
    """(E(theta+pi/2 e_k) - E(theta-pi/2 e_k)) / 2 per parameter.

    Restates ``backends/variational_numpy.py:149-156``.  Exact for the
    single-Pauli-generator rotations rx/ry/rz; not valid for p/cp/crz.
    """
    theta = np.asarray(theta, dtype=np.float64)
    g = np.zeros_like(theta)
    for k in range(len(theta)):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += np.pi / 2
        tm[k] -= np.pi / 2
        g[k] = 0.5 * (energy(patch_template(template, tp), hterms, n)[0]
                      - energy(patch_template(template, tm), hterms, n)[0])
    return g


def fd_gradient(template, theta, hterms, n, h=1e-4):
    """Central finite differences in complex128 (``test_autodiff.py:59-69``)."""
    theta = np.asarray(theta, dtype=np.float64)
    g = np.zeros_like(theta)
    for k in range(len(theta)):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        g[k] = (energy(patch_template(template, tp), hterms, n)[0]
                - energy(patch_template(template, tm), hterms, n)[0]) / (2 * h)
    return g


# ---------------------------------------------------------------------------
# circuit generators -- circuits/trotter.py
# ---------------------------------------------------------------------------

def pauli_rotation_ops(pauli_dict, angle):
    """exp(-i angle P) gate tuples; restates ``trotter.py:52-99``."""
    qs = sorted(pauli_dict)
    if not qs:
        return []
    ops = []
    for q in qs:
        if pauli_dict[q] == 'X':
            ops.append(('h', q))
        elif pauli_dict[q] == 'Y':
            ops += [('sdg', q), ('h', q)]
    for i in range(len(qs) - 1):
        ops.append(('cx', qs[i], qs[i + 1]))
    ops.append(('rz', qs[-1], 2 * angle))
    for i in reversed(range(len(qs) - 1)):
        ops.append(('cx', qs[i], qs[i + 1]))
    for q in qs:
        if pauli_dict[q] == 'X':
            ops.append(('h', q))
        elif pauli_dict[q] == 'Y':
            ops += [('h', q), ('s', q)]
    return ops


def trotter_evolve_ops(terms, t, n_steps, order=1):
    """First/second-order Trotter product; restates ``trotter.py:102-149``."""
    if order not in (1, 2):
        raise ValueError("order must be 1 or 2")
    dt = t / n_steps
    step = []
    if order == 1:
        for c, d in terms:
            step += pauli_rotation_ops(d, c * dt)
    else:
        for c, d in terms:
            step += pauli_rotation_ops(d, c * dt / 2)
        for c, d in reversed(terms):
            step += pauli_rotation_ops(d, c * dt / 2)
    return step * n_steps


# ---------------------------------------------------------------------------
# stochastic Kraus channels (NumPy path) -- noise/kraus_channels.py, kraus/*.py
# ---------------------------------------------------------------------------

NOISE_MODELS = ['ideal', 'depolarizing', 'bitflip', 'phaseflip', 'amplitude_damping', 'combined']


def _noise_pairs(dim, q):
    """idx_0 (bit q clear) and idx_0 | 1 << q: the reference indexes noise
    qubits LSB-first (``kraus_channels.py:38-58``, shift at ``:54``)."""
    step = 1 << q
    i = np.arange(dim)
    i0 = i[(i & step) == 0]
    return i0, i0 | step


def _depolarizing(sv, i0, i1, p, rng):
    """``kraus/depolarizing.py:18-47`` (NumPy branch): r, ch drawn in that
    order; fire if r < p; X if ch < 1/3, Y if ch < 2/3, else Z."""
    r, ch = rng.random(), rng.random()
    if r < p:
        v0, v1 = sv[i0].copy(), sv[i1].copy()
        if ch < 1.0 / 3.0:
            sv[i0], sv[i1] = v1, v0
        elif ch < 2.0 / 3.0:
            sv[i0], sv[i1] = -1j * v1, 1j * v0
        else:
            sv[i1] = -v1


def _bitflip(sv, i0, i1, p, rng):
    """``kraus/bitflip.py:10-22``."""
    if rng.random() < p:
        v0, v1 = sv[i0].copy(), sv[i1].copy()
        sv[i0], sv[i1] = v1, v0


def _phaseflip(sv, i0, i1, p, rng):
    """``kraus/phaseflip.py:10-21``."""
    if rng.random() < p:
        sv[i1] = -sv[i1]


def _amplitude_damping(sv, i0, i1, p, rng):
    """``kraus/amplitude_damping.py:21-47``: one jump decision per qubit with
    the whole-state probability P(K1) = gamma * sum |v1|^2."""
    gamma = float(np.clip(p, 0.0, 1.0))
    r = rng.random()
    v0, v1 = sv[i0].copy(), sv[i1].copy()
    p1 = float(np.clip(gamma * np.sum(np.abs(v1) ** 2), 0.0, 1.0))
    if r < p1:
        sv[i0] = v1 * np.sqrt(gamma) / np.sqrt(max(p1, 1e-15))
        sv[i1] = 0.0
    else:
        nn = np.sqrt(max(1.0 - p1, 1e-15))
        sv[i0] = v0 / nn
        sv[i1] = v1 * np.sqrt(1.0 - gamma) / nn


def _combined(sv, i0, i1, p, rng):
    """``kraus/combined.py:10-16``: depolarizing(p/2) then damping(p*0.333333)."""
    _depolarizing(sv, i0, i1, p * 0.5, rng)
    _amplitude_damping(sv, i0, i1, p * 0.333333, rng)


def noise_apply_to_sv(sv, n, model, p, rng, qubits=None):
    """``NoiseModel.apply_to_sv`` NumPy path (``kraus_channels.py:104-235``):
    'ideal' or p <= 0 returns the input; otherwise a copy gets one channel
    application per target qubit (all qubits by default, in list order) and
    is divided by (norm + 1e-15)."""
    if model == 'ideal' or p <= 0.0:
        return sv
    channel = {'depolarizing': _depolarizing, 'bitflip': _bitflip, 'phaseflip': _phaseflip,
               'amplitude_damping': _amplitude_damping, 'combined': _combined}[model]
    out = np.array(sv, dtype=np.complex128, copy=True)
    for q in (qubits if qubits is not None else range(n)):
        i0, i1 = _noise_pairs(len(out), q)
        channel(out, i0, i1, p, rng)
    return out / (np.linalg.norm(out) + 1e-15)
