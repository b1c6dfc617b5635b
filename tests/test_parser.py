"""QASMParser behaviour, following the reference's parser test cases
(tests/unit/test_parser.py) -- the expected op lists are the pinned answers."""
import numpy as np
import pytest

from desv import QASMCircuit, QASMParser, QuantumTranspiler

P = QASMParser


def names(c):
    return [op['name'] for op in c.ops]


def test_range_expands_single_qubit_gates():
    c = P().parse('OPENQASM 2.0; include "qelib1.inc"; qreg q[4]; h q[0:3];')
    assert [op['qubits'] for op in c.ops] == [[0], [1], [2]]
    assert names(c) == ['h'] * 3


def test_two_qubit_operands_not_expanded():
    c = P().parse('OPENQASM 2.0; qreg q[4]; cx q[0],q[1];')
    assert len(c.ops) == 1 and c.ops[0]['qubits'] == [0, 1]


def test_for_loop_unrolls_inclusive_and_keeps_following_gate():
    c = P().parse('OPENQASM 3.0; qubit[3] q; for int i in [0:2] { h q[i]; } cx q[0], q[1];')
    assert names(c) == ['h', 'h', 'h', 'cx']
    assert [op['qubits'] for op in c.ops] == [[0], [1], [2], [0, 1]]


def test_for_loop_bound_from_declared_int():
    c = P().parse('OPENQASM 3.0; qubit[3] q; const int n = 3; for int i in [0:n-1] { rx(0.5) q[i]; }')
    assert len(c.ops) == 3 and all(op['name'] == 'rx' and op['params'] == [0.5] for op in c.ops)


def test_for_body_with_several_statements():
    c = P().parse('qreg q[2]; for int i in [0:1] { h q[i]; x q[i]; }')
    assert names(c) == ['h', 'x', 'h', 'x']
    assert [op['qubits'][0] for op in c.ops] == [0, 0, 1, 1]


def test_if_and_while_blocks_are_dropped_cleanly():
    c = P().parse('OPENQASM 3.0; qubit[2] q; bit[2] c; if (c[0]) { x q[0]; } h q[1];')
    assert c.ops == [{'type': 'gate', 'name': 'h', 'qubits': [1], 'params': []}]
    c = P().parse('OPENQASM 3.0; qubit[2] q; while (true) { x q[0]; } h q[1];')
    assert c.ops == [{'type': 'gate', 'name': 'h', 'qubits': [1], 'params': []}]


def test_unresolvable_for_dropped():
    c = P().parse('OPENQASM 3.0; qubit[2] q; for int i in [0:m] { h q[i]; } cx q[0], q[1];')
    assert names(c) == ['cx'] and c.ops[0]['qubits'] == [0, 1]


def test_validate():
    ok, msg = P().validate(P().parse('qreg q[2]; h q[0]; cx q[0],q[1];'))
    assert ok and msg == 'OK'
    ok, msg = P().validate(QASMCircuit(0, 0, []))
    assert not ok and 'n_qubits' in msg
    ok, msg = P().validate(QASMCircuit(2, 0, []))
    assert not ok and 'No gate operations' in msg
    ok, msg = P().validate(QASMCircuit(2, 0, [{'type': 'gate', 'name': 'h', 'qubits': [5], 'params': []}]))
    assert not ok and 'qubit 5' in msg


def test_qasm3_bits_and_unbalanced_paren():
    c = P().parse('OPENQASM 3.0; qubit[2] q; bit[2] c; h q[0]; cx q[0], q[1];')
    assert c.n_cbits == 2 and names(c) == ['h', 'cx']
    c = P().parse('OPENQASM 2.0; include "qelib1.inc"; qreg q[2]; rx(0.5 q[0]; h q[1];')
    assert names(c) == ['h']


def test_unbalanced_brace_does_not_hang():
    c = P().parse('OPENQASM 3.0;\nqreg q[2];\nfor int i in [0:1] { h q[i];\ncx q[0], q[1];\n')
    assert isinstance(c.ops, list)


def test_params_and_fallback_index():
    c = P().parse('OPENQASM 2.0; qreg q[1]; u2(0.1,0.2) q[0];')
    assert c.ops[0]['name'] == 'u2' and c.ops[0]['params'] == pytest.approx([0.1, 0.2])
    c = P().parse('OPENQASM 2.0; qreg q[2]; h q[5];')
    assert c.ops[0]['qubits'] == [5]


def test_aliases_and_bare_register():
    c = P().parse('qreg q[3]; cnot q[0],q[1]; cu1(0.3) q[1],q[2]; toffoli q[0],q[1],q[2]; u1(0.2) q[0]; h q;')
    assert names(c) == ['cx', 'cp', 'ccx', 'p', 'h']
    assert c.ops[-1]['qubits'] == [0]


def test_comments_are_stripped():
    c = P().parse('qreg q[2]; // h q[0];\n /* x q[1]; */ cx q[0],q[1];')
    assert names(c) == ['cx']


def test_iteration_and_transpile_accept_circuits():
    c = P().parse('qreg q[3]; ccx q[0],q[1],q[2];')
    assert list(c) == c.to_tuples()
    assert len(QuantumTranspiler.transpile(c)) == 15


_ESCAPES = [
    '().__class__.__bases__[0].__subclasses__().__len__()',
    '__import__("os").system("echo pwned")',
    'getattr(1, "__class__")',
    '[x for x in range(10)][0]',
    '(lambda: 1)()',
    'exec("1")',
    'globals()',
    '().__class__.__init__.__globals__',
]


@pytest.mark.parametrize('payload', _ESCAPES)
def test_expression_evaluator_rejects_escapes(payload):
    with pytest.raises(ValueError):
        P()._eval_param(payload)
    assert P()._resolve_int_expr(payload, {}) is None


def test_escape_through_parse_raises_and_in_for_bound_yields_nothing():
    with pytest.raises(ValueError):
        P().parse('OPENQASM 3.0; qubit[1] q; rx(().__class__.__bases__[0].__subclasses__().__len__()) q[0];')
    c = P().parse('OPENQASM 3.0; qubit[1] q; for int i in [0:().__class__.__bases__[0].__subclasses__().__len__()] { x q[0]; }')
    assert c.ops == []


@pytest.mark.parametrize('expr,val', [
    ('pi', np.pi), ('pi/2', np.pi / 2), ('-pi/4', -np.pi / 4), ('0.5', 0.5), ('-0.5', -0.5),
    ('sqrt(2)', np.sqrt(2)), ('cos(0.3)', np.cos(0.3)), ('sin(pi/4)*2', np.sin(np.pi / 4) * 2),
    ('2*pi/3', 2 * np.pi / 3), ('1+2*3', 7.0), ('tau', 2 * np.pi), ('2**3 % 5', 3.0),
])
def test_legitimate_expressions(expr, val):
    assert P()._eval_param(expr) == pytest.approx(val)


@pytest.mark.parametrize('bad', ['pi * / 2', 'pi +', '(pi', 'pi 2', '**pi', 'pi // 2'])
def test_malformed_expressions_raise(bad):
    with pytest.raises(ValueError):
        P()._eval_param(bad)


def test_malformed_expression_through_parse_raises():
    with pytest.raises(ValueError):
        P().parse('OPENQASM 2.0; include "qelib1.inc"; qreg q[1]; rx(pi * / 2) q[0];')


def test_user_gate_definitions():
    c = P().parse('OPENQASM 2.0; gate bell a,b { h a; cx a,b; } qreg q[2]; bell q[0],q[1];')
    assert [(o['name'], o['qubits']) for o in c.ops] == [('h', [0]), ('cx', [0, 1])]
    c = P().parse('OPENQASM 2.0;\ngate bell a,b { h a; cx a,b; }\ngate rot(t) a { rz(t/2) a; }\n'
                  'gate two(t) a,b,c { bell a,b; cx b,c; rot(2*t) c; }\nqreg q[3]; two(pi) q[0],q[1],q[2]; h q[2];')
    assert names(c) == ['h', 'cx', 'cx', 'rz', 'h']
    assert c.ops[3]['qubits'] == [2] and c.ops[3]['params'] == pytest.approx([np.pi])
    with pytest.raises(ValueError, match='expects'):
        P().parse('gate bell a,b { h a; cx a,b; } qreg q[3]; bell q[0],q[1],q[2];')


CDKM = """OPENQASM 2.0;
include "qelib1.inc";
gate gate_MAJ q0,q1,q2 { cx q0,q1; cx q0,q2; ccx q2,q1,q0; }
gate gate_UMA q0,q1,q2 { ccx q2,q1,q0; cx q0,q2; cx q2,q1; }
gate circuit_41 q0,q1,q2,q3,q4,q5 { gate_MAJ q1,q3,q0; gate_MAJ q2,q4,q1; cx q2,q5; gate_UMA q2,q4,q1; gate_UMA q1,q3,q0; }
gate gate_CDKMRippleCarryAdder q0,q1,q2,q3,q4,q5 { circuit_41 q0,q1,q2,q3,q4,q5; }
qreg q[6];
gate_CDKMRippleCarryAdder q[0],q[1],q[2],q[3],q[4],q[5];
"""


def test_exported_cuccaro_adder_adds():
    """Reference test_parser.py:451-466: the Qiskit-exported ripple-carry
    adder maps |a, b> -> |a, (a + b) mod 8> with a = qubits (1, 2) and
    b = qubits (3, 4, 5), register values LSB-first in list order
    (reference circuits/arithmetic.py:38-51), on every basis input with
    qubit 0 clear; checked with the oracle."""
    from oracle import sv_oracle as O
    c = P().parse(CDKM)
    assert {o['name'] for o in c.ops} == {'cx', 'ccx'}
    n = 6
    for i in range(2 ** n):
        if (i >> (n - 1)) & 1:
            continue
        bit = lambda x, q: (x >> (n - 1 - q)) & 1  # noqa: E731
        a = bit(i, 1) + 2 * bit(i, 2)
        b = bit(i, 3) + 2 * bit(i, 4) + 4 * bit(i, 5)
        s_ = (a + b) % 8
        j = i
        for k, q in enumerate((3, 4, 5)):
            sh = n - 1 - q
            j = (j & ~(1 << sh)) | (((s_ >> k) & 1) << sh)
        sv = np.zeros(1 << n, dtype=np.complex128)
        sv[i] = 1
        out = O.run_circuit(list(c), n, sv=sv)
        assert abs(abs(out[j]) - 1) < 1e-12
