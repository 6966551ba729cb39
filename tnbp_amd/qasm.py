"""OpenQASM 2 front end: program -> lattice edges -> TPO slices.

The reference's estimator reads a QASM file into ``qasm::Program``
(include/qasm/ir.h:10-40; its parser directories include/qasm/v2, v3 are not
part of the reference tree), takes the lattice from the circuit's two-qubit
gates (``EdgesFromQasm``, include/tnbp/parser/qasmtoedge.h:38-57), builds one
TPO slice per barrier-separated block or per ``num_gates`` count
(``QasmToTPO``, qasmtotpo.h:700-767) from the gate matrices of
``InstructionMatrix`` (qasmtotpo.h:19-144), and optionally compresses each
slice with ``OptTPObySVD`` (optsvdtpo.h:87-214).  This module does the same on
the host (the operators are a few hundred numbers per site); the slices feed
``AbsorbTPO``.

Parsed: ``OPENQASM``, ``include``, ``qreg``/``creg``, gate applications with
parameter expressions (numbers, ``pi``, + - * / ^, unary minus, parentheses,
sin cos tan exp ln sqrt) on indexed qubits, ``barrier``.  Gates: u3/u, u2,
u1, rx, ry, rz, h, x, y, z, s, sdg, t, tdg, id, cx, cz, swap, rzz.  Like the
reference, a gate without a matrix (measure, reset, custom gates, three-qubit
gates on this TPO path) is an error.
"""
from __future__ import annotations

import cmath
import math
import re
from dataclasses import dataclass, field

import numpy as np

from . import circuits


@dataclass
class Instruction:
    op: str
    params: list
    qubits: list          # global qubit indices (registers laid out in declaration order)
    line: int = 0


@dataclass
class Program:
    qregs: list = field(default_factory=list)   # [(name, size)]
    instructions: list = field(default_factory=list)


ONE_QUBIT = {"u3", "u2", "u1", "rx", "ry", "rz", "h", "x", "y", "z", "s", "sdg", "t", "tdg", "id"}
TWO_QUBIT = {"cx", "cz", "swap", "rzz"}
NPARAMS = {"u3": 3, "u2": 2, "u1": 1, "rx": 1, "ry": 1, "rz": 1, "rzz": 1}


def gate_matrix(op, params=()):
    """Row-major U[out][in] of ``InstructionMatrix`` (qasmtotpo.h:19-144);
    for two-qubit gates the first listed qubit is the low bit."""
    c = np.complex128
    if op == "u3":
        th, ph, la = params
        return np.array([[math.cos(th / 2), -cmath.exp(1j * la) * math.sin(th / 2)],
                         [cmath.exp(1j * ph) * math.sin(th / 2), cmath.exp(1j * (ph + la)) * math.cos(th / 2)]], dtype=c)
    if op == "u2":
        ph, la = params
        s = 1 / math.sqrt(2.0)
        return np.array([[s, -cmath.exp(1j * la) * s], [cmath.exp(1j * ph) * s, cmath.exp(1j * (ph + la)) * s]], dtype=c)
    if op == "u1":
        return np.array([[1, 0], [0, cmath.exp(1j * params[0])]], dtype=c)
    if op == "rx":
        a = params[0] / 2
        return np.array([[math.cos(a), -1j * math.sin(a)], [-1j * math.sin(a), math.cos(a)]], dtype=c)
    if op == "ry":
        a = params[0] / 2
        return np.array([[math.cos(a), -math.sin(a)], [math.sin(a), math.cos(a)]], dtype=c)
    if op == "rz":
        a = params[0] / 2
        return np.array([[cmath.exp(-1j * a), 0], [0, cmath.exp(1j * a)]], dtype=c)
    fixed = {
        "h": np.array([[1, 1], [1, -1]]) / math.sqrt(2.0),
        "x": [[0, 1], [1, 0]], "y": [[0, -1j], [1j, 0]], "z": [[1, 0], [0, -1]],
        "s": [[1, 0], [0, 1j]], "sdg": [[1, 0], [0, -1j]],
        "t": [[1, 0], [0, cmath.exp(1j * math.pi / 4)]], "tdg": [[1, 0], [0, cmath.exp(-1j * math.pi / 4)]],
        "id": [[1, 0], [0, 1]],
    }
    if op in fixed:
        return np.array(fixed[op], dtype=c)
    m = np.zeros((4, 4), dtype=c)
    if op == "cx":
        m[0, 0] = m[2, 2] = m[1, 3] = m[3, 1] = 1
    elif op == "cz":
        m[0, 0] = m[1, 1] = m[2, 2] = 1
        m[3, 3] = -1
    elif op == "swap":
        m[0, 0] = m[1, 2] = m[2, 1] = m[3, 3] = 1
    elif op == "rzz":
        a = params[0] / 2
        m[0, 0] = m[3, 3] = cmath.exp(-1j * a)
        m[1, 1] = m[2, 2] = cmath.exp(1j * a)
    else:
        raise ValueError(f"no matrix for gate {op!r}")
    return m


# ------------------------------------------------------------ expressions
_TOKEN = re.compile(r"\s*(?:(\d+\.\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?|\d+(?:[eE][-+]?\d+)?)|([A-Za-z_]\w*)|(.))")
_FUNCS = {"sin": math.sin, "cos": math.cos, "tan": math.tan, "exp": math.exp, "ln": math.log,
          "sqrt": math.sqrt}


def eval_expr(text):
    """Value of a QASM 2 parameter expression."""
    toks = []
    for num, name, op in _TOKEN.findall(text):
        if num:
            toks.append(("n", float(num)))
        elif name:
            toks.append(("id", name))
        elif op.strip():
            toks.append(("op", op))
    pos = [0]

    def peek():
        return toks[pos[0]] if pos[0] < len(toks) else (None, None)

    def take(kind=None, val=None):
        t = peek()
        if t[0] is None or (kind and t[0] != kind) or (val and t[1] != val):
            raise ValueError(f"bad expression {text!r}")
        pos[0] += 1
        return t

    def atom():
        k, v = peek()
        if k == "n":
            take()
            return v
        if k == "id":
            take()
            if v == "pi":
                return math.pi
            if v in _FUNCS:
                take("op", "(")
                x = expr()
                take("op", ")")
                return _FUNCS[v](x)
            raise ValueError(f"unknown name {v!r} in {text!r}")
        if (k, v) == ("op", "("):
            take()
            x = expr()
            take("op", ")")
            return x
        if (k, v) == ("op", "-"):
            take()
            return -power()
        if (k, v) == ("op", "+"):
            take()
            return power()
        raise ValueError(f"bad expression {text!r}")

    def power():
        x = atom()
        if peek() == ("op", "^"):
            take()
            return x ** power()
        return x

    def term():
        x = power()
        while peek() in (("op", "*"), ("op", "/")):
            if take()[1] == "*":
                x *= power()
            else:
                x /= power()
        return x

    def expr():
        x = term()
        while peek() in (("op", "+"), ("op", "-")):
            if take()[1] == "+":
                x += term()
            else:
                x -= term()
        return x

    v = expr()
    if pos[0] != len(toks):
        raise ValueError(f"bad expression {text!r}")
    return v


# ------------------------------------------------------------------ parser
_STMT = re.compile(r"^([A-Za-z_]\w*)\s*(?:\((.*)\))?\s*(.*)$", re.S)


def _split_top(s, sep=","):
    out, depth, cur = [], 0, []
    for ch in s:
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        if ch == sep and depth == 0:
            out.append("".join(cur))
            cur = []
        else:
            cur.append(ch)
    out.append("".join(cur))
    return [x.strip() for x in out if x.strip()]


def parse(text):
    """Program of an OpenQASM 2 text (the subset above)."""
    text = re.sub(r"//[^\n]*", "", text)
    prog = Program()
    offset = {}
    size = {}
    line = 1
    for raw in text.split(";"):
        stmt = raw.strip()
        nl = raw.count("\n")
        if not stmt:
            line += nl
            continue
        ln = line + raw[: len(raw) - len(raw.lstrip())].count("\n")
        line += nl
        if stmt.startswith("OPENQASM") or stmt.startswith("include"):
            continue
        m = re.match(r"^(qreg|creg)\s+([A-Za-z_]\w*)\s*\[\s*(\d+)\s*\]$", stmt)
        if m:
            if m.group(1) == "qreg":
                offset[m.group(2)] = sum(n for _, n in prog.qregs)
                size[m.group(2)] = int(m.group(3))
                prog.qregs.append((m.group(2), int(m.group(3))))
            continue
        m = _STMT.match(stmt)
        if not m:
            raise ValueError(f"QASM parse error (line {ln}): {stmt!r}")
        op, ptext, args = m.group(1).lower(), m.group(2), m.group(3)
        if op == "u":
            op = "u3"
        if op in ("measure", "reset", "gate", "opaque", "if"):
            raise ValueError(f"QASM line {ln}: {op!r} has no operator on the TPO path")
        params = [eval_expr(p) for p in _split_top(ptext)] if ptext else []
        qubits = []
        for a in _split_top(args):
            mm = re.match(r"^([A-Za-z_]\w*)\s*\[\s*(\d+)\s*\]$", a)
            if mm:
                reg, idx = mm.group(1), int(mm.group(2))
                if reg not in offset or idx >= size[reg]:
                    raise ValueError(f"QASM line {ln}: unknown qubit {a!r}")
                qubits.append(offset[reg] + idx)
            elif a in offset and op == "barrier":
                qubits += list(range(offset[a], offset[a] + size[a]))
            else:
                raise ValueError(f"QASM line {ln}: bad argument {a!r}")
        if op == "barrier":
            prog.instructions.append(Instruction("barrier", [], qubits, ln))
            continue
        if op not in ONE_QUBIT and op not in TWO_QUBIT:
            raise ValueError(f"QASM line {ln}: unsupported gate {op!r}")
        want = 1 if op in ONE_QUBIT else 2
        if len(qubits) != want or len(params) != NPARAMS.get(op, 0):
            raise ValueError(f"QASM line {ln}: {op} takes {want} qubit(s) and "
                             f"{NPARAMS.get(op, 0)} parameter(s)")
        if want == 2 and qubits[0] == qubits[1]:
            raise ValueError(f"QASM line {ln}: repeated qubit")
        prog.instructions.append(Instruction(op, params, qubits, ln))
    return prog


# ------------------------------------------------------- program -> TPO
def sites_from_qasm(prog):
    """``SitesFromQasm`` (qasmtoedge.h:19-32)."""
    return sorted({q for ins in prog.instructions if ins.op != "barrier" for q in ins.qubits})


def edges_from_qasm(prog):
    """``EdgesFromQasm`` (qasmtoedge.h:38-57): sorted unique (min, max) pairs
    of the two-qubit gates."""
    return sorted({(min(ins.qubits), max(ins.qubits)) for ins in prog.instructions
                   if len(ins.qubits) == 2 and ins.op != "barrier"})


def qasm_to_tpo(prog, edges, num_gates=None):
    """TPO slices (``QasmToTPO``, qasmtotpo.h:700-767): split at barriers (a
    trailing barrier opens no slice), or after num_gates[m] instructions for
    slice m when given."""
    slices, cur = [], []
    if num_gates:
        it = iter(prog.instructions)
        for n in num_gates:
            block = [next(it) for _ in range(n)]
            slices.append([i for i in block if i.op != "barrier"])
    else:
        for k, ins in enumerate(prog.instructions):
            if ins.op == "barrier":
                if k != len(prog.instructions) - 1:
                    slices.append(cur)
                    cur = []
            else:
                cur.append(ins)
        slices.append(cur)
    return [circuits.layer_tpo(edges, [(gate_matrix(i.op, i.params), tuple(i.qubits)) for i in s])
            for s in slices]
