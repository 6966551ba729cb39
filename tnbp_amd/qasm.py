"""OpenQASM 2 / 3 front end: program -> lattice edges -> TPO slices.

The reference's estimator reads a QASM file into ``qasm::Program``
(include/qasm/ir.h:10-40) with its own parsers (include/qasm/any.h picks
v2/parser.h or v3/parser.h from the header), takes the lattice from the
circuit's two-qubit gates (``EdgesFromQasm``, include/tnbp/parser/
qasmtoedge.h:38-57), builds one TPO slice per barrier-separated block or per
``num_gates`` count (``QasmToTPO``, qasmtotpo.h:700-767) from the gate
matrices of ``InstructionMatrix`` (qasmtotpo.h:19-144), and optionally
compresses each slice with ``OptTPObySVD`` (optsvdtpo.h:87-214).  This module
does the same on the host (the operators are a few hundred numbers per site);
the slices feed ``AbsorbTPO``.

Parsed: the subset of the reference's parsers (see ``parse``), with parameter
expressions of numbers, ``pi``, ``tau``, + - * / ^, unary signs, parentheses
and sin cos tan exp ln sqrt.  Gates: u3/u, u2, u1/p, rx, ry, rz, h, x, y, z,
s, sdg, t, tdg, id, cx/cnot, cz, swap, rzz, ccx/toffoli, cswap.  Two-qubit
gates between non-neighbours and three-qubit gates act along the shortest
lattice path (identities in between).  Like the reference, an instruction
without a matrix (measure, reset, custom gates) is an error on the TPO path.
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
    name: str = ""        # the gate's name for "custom" instructions


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
    if op == "ccx":
        # Toffoli, controls on the first two listed qubits (qasmtotpo.h:117-126)
        m = np.eye(8, dtype=c)
        m[3, 3] = m[7, 7] = 0
        m[3, 7] = m[7, 3] = 1
        return m
    if op == "cswap":
        # Fredkin: control on the first listed qubit, which swaps the other two:
        # |011> <-> |101> in the first-qubit-lowest index (qasmtotpo.h:127-136
        # sets the swap, but also m[5][5] = 1 where m[4][4] = 1 belongs, which
        # loses |100>; this is the unitary its comment describes)
        m = np.eye(8, dtype=c)
        m[3, 3] = m[5, 5] = 0
        m[3, 5] = m[5, 3] = 1
        return m
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
            if v.lower() == "pi":
                return math.pi
            if v.lower() == "tau":
                return 2 * math.pi
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

# gate names of the reference's parsers (include/qasm/v2/parser.h:119-143,
# v3/parser.h:324-349), case-insensitive
GATE_NAMES = {"u3": "u3", "u": "u3", "u2": "u2", "u1": "u1", "p": "u1", "rx": "rx", "ry": "ry",
              "rz": "rz", "h": "h", "x": "x", "y": "y", "z": "z", "s": "s", "sdg": "sdg", "t": "t",
              "tdg": "tdg", "id": "id", "cx": "cx", "cnot": "cx", "cz": "cz", "swap": "swap",
              "rzz": "rzz", "ccx": "ccx", "toffoli": "ccx", "cswap": "cswap"}
THREE_QUBIT = {"ccx", "cswap"}


def _split_top(s, sep=","):
    out, depth, cur = [], 0, []
    for ch in s:
        if ch in "([":
            depth += 1
        elif ch in ")]":
            depth -= 1
        if ch == sep and depth == 0:
            out.append("".join(cur))
            cur = []
        else:
            cur.append(ch)
    out.append("".join(cur))
    return [x.strip() for x in out if x.strip()]


def _version(text):
    """3 when a '3' follows the OPENQASM header within 32 characters, else 2
    (the header sniff of include/qasm/any.h:8-20)."""
    pos = text.find("OPENQASM")
    return 3 if pos >= 0 and "3" in text[pos:pos + 32] else 2


def parse(text):
    """Program of an OpenQASM 2 or 3 text (the subset the reference parses,
    include/qasm/v2/parser.h, v3/parser.h): declarations (qreg/creg; qubit and
    bit with or without a size in version 3), gate calls with parameter
    expressions, register broadcast (`h q;` is one gate per qubit, a call's
    qubit list is cut into groups of the gate's arity) and, in version 3,
    slices q[lo:hi]; barrier, reset and measure (arrow form, and the
    assignment forms of version 3).  Unknown gate names become "custom"
    instructions; like the reference, those, measure and reset have no
    operator on the TPO path (qasm_to_tpo rejects them).  Qubits are numbered
    globally, registers in declaration order (with one register, the
    reference's QArg index)."""
    version = _version(text)
    text = re.sub(r"/\*.*?\*/", lambda m: "\n" * m.group(0).count("\n"), text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    prog = Program()
    offset, size = {}, {}
    csize = {}
    line = 1

    def qargs(spec, ln, many=True):
        out = []
        for a in _split_top(spec):
            mm = re.match(r"^([A-Za-z_]\w*)\s*\[\s*(\d+)\s*(?::\s*(\d+)\s*)?\]$", a)
            if mm:
                reg, lo = mm.group(1), int(mm.group(2))
                if reg not in offset:
                    raise ValueError(f"QASM line {ln}: unknown qubit register {reg!r}")
                if mm.group(3) is not None:
                    hi = int(mm.group(3))
                    if version != 3 or hi < lo or hi > size[reg]:
                        raise ValueError(f"QASM line {ln}: bad slice {a!r}")
                    out += [offset[reg] + k for k in range(lo, hi)]
                else:
                    if lo >= size[reg]:
                        raise ValueError(f"QASM line {ln}: qubit {a!r} out of range")
                    out.append(offset[reg] + lo)
            elif a in offset:
                out += list(range(offset[a], offset[a] + size[a]))
            else:
                raise ValueError(f"QASM line {ln}: bad qubit argument {a!r}")
        if not many and len(out) != 1:
            raise ValueError(f"QASM line {ln}: one qubit expected in {spec!r}")
        return out

    def declare_q(name, n):
        offset[name] = sum(k for _, k in prog.qregs)
        size[name] = n
        prog.qregs.append((name, n))

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
                declare_q(m.group(2), int(m.group(3)))
            else:
                csize[m.group(2)] = int(m.group(3))
            continue
        if version == 3:
            m = re.match(r"^(qubit|bit)\s*(?:\[\s*(\d+)\s*\])?\s+([A-Za-z_]\w*)$", stmt)
            if m:
                n = int(m.group(2)) if m.group(2) else 1
                if m.group(1) == "qubit":
                    declare_q(m.group(3), n)
                else:
                    csize[m.group(3)] = n
                continue
            m = re.match(r"^(?:bit\s*(?:\[\s*\d+\s*\])?\s+)?[A-Za-z_]\w*\s*(?:\[\s*\d+\s*\])?\s*=\s*"
                         r"measure\s+(.+)$", stmt, re.S)
            if m:
                for q in qargs(m.group(1), ln):
                    prog.instructions.append(Instruction("measure", [], [q], ln))
                continue
        m = re.match(r"^measure\s+(.+?)\s*->\s*(.+)$", stmt, re.S)
        if m:
            prog.instructions.append(Instruction("measure", [], qargs(m.group(1), ln, False), ln))
            continue
        m = _STMT.match(stmt)
        if not m:
            raise ValueError(f"QASM parse error (line {ln}): {stmt!r}")
        word, ptext, args = m.group(1), m.group(2), m.group(3)
        low = word.lower()
        if low in ("ctrl", "inv", "pow", "gate", "opaque", "if"):
            raise ValueError(f"QASM line {ln}: {word!r} is outside the parsed subset")
        if low == "barrier":
            prog.instructions.append(Instruction("barrier", [], qargs(args, ln), ln))
            continue
        if low == "reset":
            for q in qargs(args, ln):
                prog.instructions.append(Instruction("reset", [], [q], ln))
            continue
        params = [eval_expr(x) for x in _split_top(ptext)] if ptext else []
        qs = qargs(args, ln)
        op = GATE_NAMES.get(low)
        if op is None:
            prog.instructions.append(Instruction("custom", params, qs, ln, name=word))
            continue
        ar = 1 if op in ONE_QUBIT else 2 if op in TWO_QUBIT else 3
        if len(qs) % ar or not qs:
            raise ValueError(f"QASM line {ln}: {op} expects a multiple of {ar} qubits")
        if len(params) != NPARAMS.get(op, 0):
            raise ValueError(f"QASM line {ln}: {op} takes {NPARAMS.get(op, 0)} parameter(s)")
        for k in range(0, len(qs), ar):
            group = qs[k:k + ar]
            if len(set(group)) != ar:
                raise ValueError(f"QASM line {ln}: repeated qubit in {op}")
            prog.instructions.append(Instruction(op, list(params), group, ln))
    return prog


# ------------------------------------------------------- program -> TPO
def sites_from_qasm(prog):
    """``SitesFromQasm`` (qasmtoedge.h:19-32)."""
    return sorted({q for ins in prog.instructions if ins.op not in ("barrier", "custom") for q in ins.qubits})


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
    for sl in slices:
        for i in sl:
            if i.op in ("measure", "reset", "custom"):
                what = i.name or i.op
                raise ValueError(f"QASM line {i.line}: {what!r} has no operator on the TPO path")
    return [circuits.layer_tpo(edges, [(gate_matrix(i.op, i.params), tuple(i.qubits)) for i in s])
            for s in slices]
