"""OpenQASM 2/3 front end: ``QASMParser`` and ``QASMCircuit``.

Host-side input layer of the drop-in surface.  The accepted language and the
resulting op lists follow the reference parser's documented behaviour
(``dense_evolution/circuits/parser.py:13-695``, pinned by its
``tests/unit/test_parser.py``):

* ``//`` and ``/* */`` comments are removed;
* ``gate NAME(params) args { body }`` definitions are lifted out and every
  call is expanded (recursively, nesting limit 64, arity checked);
* ``for <type> v in [a:b] { ... }`` is unrolled with an INCLUSIVE end bound,
  where a and b may use ``int``/``const int`` declarations; ``if``, ``while``,
  ``def`` and unresolvable loops are dropped; an unbalanced brace stops the
  block pass;
* statements are split on ``;``; ``OPENQASM``, ``include``, ``barrier``,
  ``measure``, ``reset`` and stray block keywords are skipped;
* ``qreg``/``creg`` (QASM 2) and ``qubit``/``bit`` (QASM 3) declare registers;
  a bare register name means its first qubit; ``q[a:b]`` is an exclusive
  range; an index missing from the map falls back to the literal number;
* gate names are lower-cased and aliased (cu1->cp, u1->p, toffoli->ccx,
  fredkin->cswap, cnot->cx, not->x, id->i); single-qubit gates applied to
  several qubits become one op per qubit;
* parameters are evaluated by an AST whitelist (numbers, + - * / % **, unary
  +/-, the constants pi/tau/euler and a fixed set of math functions); any
  other expression raises ValueError.
"""
from __future__ import annotations

import ast
import math
import operator
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np


@dataclass
class QASMCircuit:
    """Parsed circuit: register sizes plus ``ops`` dicts
    ``{'type': 'gate', 'name', 'qubits': [int], 'params': [float]}``."""
    n_qubits: int = 0
    n_cbits: int = 0
    ops: List[Dict] = field(default_factory=list)

    def to_tuples(self) -> List[Tuple]:
        """``(name, *qubits, *params)`` per op -- the run_circuit tuple format."""
        return [(op['name'],) + tuple(op['qubits']) + tuple(op['params']) for op in self.ops]

    def __iter__(self):
        return iter(self.to_tuples())


_ALIASES = {
    'cu1': 'cp', 'u1': 'p', 'toffoli': 'ccx', 'fredkin': 'cswap', 'cnot': 'cx',
    'not': 'x', 'id': 'i',
}
_ONE_QUBIT = frozenset(('h', 'x', 'y', 'z', 's', 'sdg', 't', 'tdg', 'sx', 'id', 'rx', 'ry', 'rz',
                        'p', 'u1', 'u2', 'u3'))
_SKIPPED = frozenset(('openqasm', 'include', 'barrier', 'measure', 'reset', 'gate', 'def', 'if',
                      'for', 'while'))
_CONSTANTS = {'pi': math.pi, 'tau': 2.0 * math.pi, 'euler': math.e}
_FUNCTIONS = {
    'sin': np.sin, 'cos': np.cos, 'tan': np.tan, 'sqrt': np.sqrt, 'exp': np.exp, 'log': np.log,
    'asin': np.arcsin, 'acos': np.arccos, 'atan': np.arctan,
    'arcsin': np.arcsin, 'arccos': np.arccos, 'arctan': np.arctan,
    'abs': abs, 'round': round,
}
_BIN = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul,
        ast.Div: operator.truediv, ast.Pow: operator.pow, ast.Mod: operator.mod}
_UN = {ast.UAdd: operator.pos, ast.USub: operator.neg}

_IDENT = r'[A-Za-z_]\w*'
_RE_QREG = re.compile(rf'^qreg\s+({_IDENT})\s*\[(\d+)\]')
_RE_CREG = re.compile(rf'^creg\s+({_IDENT})\s*\[(\d+)\]')
_RE_QUBIT = re.compile(rf'^qubit(?:\s*\[(\d+)\])?\s+({_IDENT})')
_RE_BIT = re.compile(rf'^bit(?:\s*\[(\d+)\])?\s+({_IDENT})')
_RE_RANGE = re.compile(rf'^({_IDENT})\[(\d+):(\d+)\]$')
_RE_INDEX = re.compile(r'\[(\d+)\]')
_RE_BLOCK = re.compile(r'\b(for|if|while|def|gate)\b[^{]*\{')
_RE_FOR = re.compile(r'for\s+(?:\w+\s+)?(\w+)\s+in\s*\[\s*([^\]]+?)\s*:\s*([^\]]+?)\s*\]\s*\{')
_RE_INTDECL = re.compile(r'(?:const\s+)?int(?:\s*\[\d+\])?\s+(\w+)\s*=\s*(-?\d+)\s*;')
_RE_GATEDEF = re.compile(rf'\bgate\s+({_IDENT})\s*(?:\(([^)]*)\))?\s*([^{{]*)\{{')


def _closing_brace(s: str, open_at: int) -> Optional[int]:
    depth = 0
    for k in range(open_at, len(s)):
        c = s[k]
        if c == '{':
            depth += 1
        elif c == '}':
            depth -= 1
            if depth == 0:
                return k
    return None


def _split_top_level(s: str) -> List[str]:
    """Split on commas that are not inside parentheses."""
    parts, cur, depth = [], [], 0
    for c in s:
        if c == ',' and depth == 0:
            parts.append(''.join(cur).strip())
            cur = []
            continue
        depth += (c == '(') - (c == ')')
        cur.append(c)
    if cur:
        parts.append(''.join(cur).strip())
    return [p for p in parts if p]


def _evaluate(node, names: Dict):
    """Whitelisted evaluation of a parsed expression (never eval/exec)."""
    if isinstance(node, ast.Expression):
        return _evaluate(node.body, names)
    if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)) \
            and not isinstance(node.value, bool):
        return node.value
    if isinstance(node, ast.BinOp) and type(node.op) in _BIN:
        return _BIN[type(node.op)](_evaluate(node.left, names), _evaluate(node.right, names))
    if isinstance(node, ast.UnaryOp) and type(node.op) in _UN:
        return _UN[type(node.op)](_evaluate(node.operand, names))
    if isinstance(node, ast.Name) and node.id in names and not callable(names[node.id]):
        return names[node.id]
    if (isinstance(node, ast.Call) and isinstance(node.func, ast.Name) and not node.keywords
            and node.func.id in names and callable(names[node.func.id])):
        return names[node.func.id](*[_evaluate(a, names) for a in node.args])
    raise ValueError(f"disallowed expression node: {type(node).__name__}")


class QASMParser:
    """OpenQASM 2.0 / 3.0 parser producing :class:`QASMCircuit`."""

    _MATH_ENV = {**_CONSTANTS, **_FUNCTIONS}

    # -- expressions ---------------------------------------------------------
    def _eval_param(self, tok: str) -> float:
        try:
            return float(_evaluate(ast.parse(tok, mode='eval'), self._MATH_ENV))
        except Exception as exc:  # noqa: BLE001 -- re-raised as ValueError
            raise ValueError(f"Invalid gate parameter expression '{tok}': {exc}") from exc

    def _resolve_int_expr(self, expr: str, decls: Dict[str, int]) -> Optional[int]:
        expr = expr.strip()
        for name, val in decls.items():
            expr = re.sub(r'\b' + re.escape(name) + r'\b', str(val), expr)
        try:
            return int(_evaluate(ast.parse(expr, mode='eval'), {}))
        except Exception:  # noqa: BLE001 -- unresolvable bound
            return None

    # -- source preprocessing -------------------------------------------------
    @staticmethod
    def _lift_gate_definitions(s: str) -> Tuple[str, Dict]:
        defs: Dict[str, Tuple[List[str], List[str], str]] = {}
        while True:
            m = _RE_GATEDEF.search(s)
            if not m:
                return s, defs
            close = _closing_brace(s, m.end() - 1)
            if close is None:
                return s, defs
            params = [p.strip() for p in (m.group(2) or '').split(',') if p.strip()]
            args = [a.strip() for a in m.group(3).split(',') if a.strip()]
            defs[m.group(1).lower()] = (params, args, s[m.end():close])
            s = s[:m.start()] + s[close + 1:]

    def _unroll_blocks(self, s: str) -> str:
        decls = {k: int(v) for k, v in _RE_INTDECL.findall(s)}
        while True:
            m = _RE_BLOCK.search(s)
            if not m:
                return s
            brace = m.end() - 1
            close = _closing_brace(s, brace)
            if close is None:
                return s
            text = ''
            if m.group(1).lower() == 'for':
                fm = _RE_FOR.search(s[m.start():brace] + '{')
                if fm:
                    lo = self._resolve_int_expr(fm.group(2), decls)
                    hi = self._resolve_int_expr(fm.group(3), decls)
                    if lo is not None and hi is not None:
                        var = re.compile(r'\b' + re.escape(fm.group(1)) + r'\b')
                        body = s[brace + 1:close]
                        text = ' '.join(var.sub(str(v), body) for v in range(lo, hi + 1))
            s = s[:m.start()] + text + s[close + 1:]

    def _expand_call(self, stmt: str, defs: Dict, depth: int = 0) -> List[str]:
        stmt = re.sub(r'\s+', ' ', stmt).strip()
        head = re.split(r'[\s(]', stmt, 1)[0].lower()
        if head not in defs:
            return [stmt]
        if depth > 64:
            raise ValueError(f"gate definitions nested too deeply at '{head}'")
        params, args, body = defs[head]
        rest = stmt[len(head):].strip()
        values: List[str] = []
        if rest.startswith('('):
            depth_p, close = 0, -1
            for k, c in enumerate(rest):
                depth_p += (c == '(') - (c == ')')
                if depth_p == 0:
                    close = k
                    break
            values = _split_top_level(rest[1:close])
            rest = rest[close + 1:]
        actual = [a.strip() for a in rest.split(',') if a.strip()]
        if len(actual) != len(args) or len(values) != len(params):
            raise ValueError(f"gate '{head}' expects {len(params)} parameters and "
                             f"{len(args)} qubits, got '{stmt}'")
        subst = dict(zip(args, actual))
        subst.update({p: f'({v})' for p, v in zip(params, values)})
        if subst:
            pat = re.compile(r'\b(' + '|'.join(re.escape(k) for k in subst) + r')\b')
            body = pat.sub(lambda mm: subst[mm.group(1)], body)
        out: List[str] = []
        for part in body.split(';'):
            if part.strip():
                out.extend(self._expand_call(part, defs, depth + 1))
        return out

    # -- operands ---------------------------------------------------------------
    @staticmethod
    def _resolve_qubits(s: str, qmap: Dict[str, int]) -> List[int]:
        out: List[int] = []
        for tok in (t.strip() for t in s.split(',')):
            if not tok:
                continue
            m = _RE_RANGE.match(tok)
            if m:
                out.extend(qmap[f'{m.group(1)}[{i}]'] for i in range(int(m.group(2)), int(m.group(3)))
                           if f'{m.group(1)}[{i}]' in qmap)
                continue
            if tok in qmap:
                out.append(qmap[tok])
                continue
            b = _RE_INDEX.search(tok)
            if b:
                key = f"{tok[:tok.index('[')]}[{b.group(1)}]"
                out.append(qmap[key] if key in qmap else int(b.group(1)))
                continue
            digits = re.findall(r'\d+', tok)
            if digits and not re.search(r'[A-Za-z_]', tok):
                out.append(int(digits[-1]))
        return out

    def _parse_gate(self, instr: str, qmap: Dict[str, int]) -> Optional[Dict]:
        instr = instr.strip()
        lp = instr.find('(')
        pstr = ''
        if lp != -1:
            depth, rp = 0, -1
            for k in range(lp, len(instr)):
                depth += (instr[k] == '(') - (instr[k] == ')')
                if instr[k] == ')' and depth == 0:
                    rp = k
                    break
            if rp == -1:
                return None  # unbalanced parentheses: statement skipped
            pstr = instr[lp + 1:rp].strip()
            head, operands = instr[:lp].strip(), instr[rp + 1:].strip()
        else:
            bits = instr.split(None, 1)
            head, operands = bits[0], (bits[1] if len(bits) > 1 else '')
        raw = head.strip().lower()
        if not raw:
            return None
        params = [self._eval_param(t) for t in _split_top_level(pstr)] if pstr else []
        qubits = self._resolve_qubits(operands.replace(' ', ''), qmap)
        if not qubits:
            return None
        return {'type': 'gate', 'name': _ALIASES.get(raw, raw), 'qubits': qubits, 'params': params}

    # -- public ------------------------------------------------------------------
    def parse(self, qasm_str: str) -> QASMCircuit:
        text = re.sub(r'/\*.*?\*/', ' ', qasm_str, flags=re.DOTALL)
        text = re.sub(r'//[^\n]*', ' ', text)
        text, defs = self._lift_gate_definitions(text)
        text = self._unroll_blocks(text)
        statements = [e for st in text.split(';') if st.strip() for e in self._expand_call(st, defs)]

        qmap: Dict[str, int] = {}
        cmap: Dict[str, int] = {}
        nq = nc = 0
        ops: List[Dict] = []
        for instr in statements:
            instr = re.sub(r'\s+', ' ', instr).strip()
            if not instr or re.split(r'[\s(]', instr)[0].lower() in _SKIPPED:
                continue
            declared = False
            for rx, is_q, size_first in ((_RE_QREG, True, False), (_RE_CREG, False, False),
                                          (_RE_QUBIT, True, True), (_RE_BIT, False, True)):
                m = rx.match(instr)
                if not m:
                    continue
                if size_first:
                    size, name = (int(m.group(1)) if m.group(1) else 1), m.group(2)
                else:
                    name, size = m.group(1), int(m.group(2))
                table, start = (qmap, nq) if is_q else (cmap, nc)
                for i in range(size):
                    table[f'{name}[{i}]'] = start + i
                table[name] = start
                if is_q:
                    nq += size
                else:
                    nc += size
                declared = True
                break
            if declared:
                continue
            op = self._parse_gate(instr, qmap)
            if op is None:
                continue
            if op['name'] in _ONE_QUBIT and len(op['qubits']) > 1:
                for q in op['qubits']:
                    ops.append({'type': op['type'], 'name': op['name'], 'qubits': [q],
                                'params': list(op['params'])})
                    nq = max(nq, q + 1)
            else:
                ops.append(op)
                nq = max(nq, max(op['qubits']) + 1)
        return QASMCircuit(nq, nc, ops)

    def validate(self, circ: QASMCircuit) -> Tuple[bool, str]:
        """Structural checks only (qubit count, non-empty, qubit ranges)."""
        if circ.n_qubits <= 0:
            return False, 'n_qubits must be > 0.'
        if not circ.ops:
            return False, 'No gate operations found in circuit.'
        for i, op in enumerate(circ.ops):
            for q in op.get('qubits', []):
                if not 0 <= q < circ.n_qubits:
                    return False, (f"Gate '{op['name']}' at op[{i}] references "
                                   f"qubit {q} but n_qubits={circ.n_qubits}.")
        return True, 'OK'
