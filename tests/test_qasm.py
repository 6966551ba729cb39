"""QASM front end and TPO compression (host logic, no GPU).

Pins: gate matrices against the oracle's restatement of InstructionMatrix and
closed forms (S = T^2, U3 special cases, CX action on basis states); TPO
slices contracted to the full 2^n operator against a brute-force statevector
(which fixes the qubit order, gate order and barrier slicing); OptTPObySVD
against the oracle's label-for-label restatement and against the exact
operator it must leave unchanged.
"""
import math

import numpy as np
import pytest

from oracle import circuits as oc
from statevector import StateVector
from tnbp_amd import circuits, qasm


def test_expressions():
    assert qasm.eval_expr("9*pi/10") == pytest.approx(0.9 * math.pi)
    assert qasm.eval_expr("-pi/2") == pytest.approx(-math.pi / 2)
    assert qasm.eval_expr("2^3^2") == 512
    assert qasm.eval_expr("sin(pi/2) + (1+2)*3 - 1e-1") == pytest.approx(9.9)
    assert qasm.eval_expr("-(-2)") == 2
    with pytest.raises(ValueError):
        qasm.eval_expr("2 +")
    with pytest.raises(ValueError):
        qasm.eval_expr("foo")


def test_gate_matrices():
    rng = np.random.default_rng(1)
    for name in ("u3", "rx", "ry", "rz", "h", "x", "y", "z", "rzz", "cx", "cz"):
        p = list(rng.uniform(-3, 3, qasm.NPARAMS.get(name, 0)))
        np.testing.assert_allclose(qasm.gate_matrix(name, p), oc.gate_matrix(name, p), atol=1e-15)
    t, s = qasm.gate_matrix("t"), qasm.gate_matrix("s")
    np.testing.assert_allclose(t @ t, s, atol=1e-15)
    np.testing.assert_allclose(s @ qasm.gate_matrix("sdg"), np.eye(2), atol=1e-15)
    np.testing.assert_allclose(qasm.gate_matrix("u2", [0.3, 0.4]), qasm.gate_matrix("u3", [math.pi / 2, 0.3, 0.4]), atol=1e-15)
    np.testing.assert_allclose(qasm.gate_matrix("u1", [0.7]), qasm.gate_matrix("u3", [0, 0, 0.7]), atol=1e-15)
    sw = qasm.gate_matrix("swap")
    assert sw[1, 2] == 1 and sw[2, 1] == 1 and sw[0, 0] == 1 and sw[3, 3] == 1
    for m in (qasm.gate_matrix(g, list(rng.uniform(0, 1, qasm.NPARAMS.get(g, 0))))
              for g in qasm.ONE_QUBIT | qasm.TWO_QUBIT):
        np.testing.assert_allclose(m @ m.conj().T, np.eye(m.shape[0]), atol=1e-14)


PROG = """OPENQASM 2.0;
include "qelib1.inc";
qreg q[4];
creg c[4];
// a comment
u(9*pi/10,-pi/2,pi/2) q[0];
cx q[0],q[1];
u1(pi/2) q[1];
cx q[0],q[1];
rzz(0.3) q[2], q[1];
barrier q;
h q[3];
cz q[2],q[3];
ry(-0.4) q[0];
barrier q[0],q[1];
"""


def test_parse_edges_and_slices():
    p = qasm.parse(PROG)
    assert p.qregs == [("q", 4)]
    ops = [i.op for i in p.instructions]
    assert ops == ["u3", "cx", "u1", "cx", "rzz", "barrier", "h", "cz", "ry", "barrier"]
    assert p.instructions[4].qubits == [2, 1]
    assert p.instructions[0].params == pytest.approx([0.9 * math.pi, -math.pi / 2, math.pi / 2])
    assert qasm.edges_from_qasm(p) == [(0, 1), (1, 2), (2, 3)]
    assert qasm.sites_from_qasm(p) == [0, 1, 2, 3]
    # a trailing barrier opens no slice
    assert len(qasm.qasm_to_tpo(p, qasm.edges_from_qasm(p))) == 2
    assert len(qasm.qasm_to_tpo(p, qasm.edges_from_qasm(p), num_gates=[3, 7])) == 2
    for bad in ("qreg q[2]; cx q[0],q[0];", "qreg q[2]; rx q[0];", "qreg q[2]; x q[5];",
                "qreg q[2]; cx q[0];"):
        with pytest.raises(ValueError):
            qasm.parse(bad)
    # parsed like the reference, but no operator on the TPO path
    for bad in ("qreg q[2]; creg c[2]; cx q[0],q[1]; measure q[0] -> c[0];",
                "qreg q[2]; cx q[0],q[1]; foo q[0];", "qreg q[2]; cx q[0],q[1]; reset q;"):
        p = qasm.parse(bad)
        with pytest.raises(ValueError, match="no operator"):
            qasm.qasm_to_tpo(p, qasm.edges_from_qasm(p))


def _full_operator(edges, O):
    """Contract a TPO slice into G[in_0..in_{n-1}, out_0..out_{n-1}]."""
    sites = sorted({s for e in edges for s in e})
    letters = iter("abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ")
    bond = {m: next(letters) for m in range(len(edges))}
    ins = {s: next(letters) for s in sites}
    outs = {s: next(letters) for s in sites}
    terms = []
    for s, o in zip(sites, O):
        lab = "".join(bond[m] for m in circuits.incident(edges, s)) + ins[s] + outs[s]
        terms.append(lab)
    res = "".join(ins[s] for s in sites) + "".join(outs[s] for s in sites)
    return np.einsum(",".join(terms) + "->" + res, *O)


def _apply(G, psi):
    n = psi.ndim
    return np.tensordot(psi, G, axes=(list(range(n)), list(range(n))))


def _check_slices_against_statevector(prog, edges, slices):
    n = 4
    blocks, cur = [], []
    for ins in prog.instructions:
        if ins.op == "barrier":
            blocks.append(cur)
            cur = []
        else:
            cur.append(ins)
    if cur:
        blocks.append(cur)
    rng = np.random.default_rng(3)
    for block, O in zip(blocks, slices):
        G = _full_operator(edges, O)
        psi = rng.standard_normal((2,) * n) + 1j * rng.standard_normal((2,) * n)
        sv = StateVector(n)
        sv.psi = psi.copy()
        for ins in block:
            u = qasm.gate_matrix(ins.op, ins.params)
            sv.apply1(u, ins.qubits[0]) if len(ins.qubits) == 1 else sv.apply2(u, *ins.qubits)
        np.testing.assert_allclose(_apply(G, psi), sv.psi, atol=1e-12)


def test_tpo_slices_equal_statevector():
    p = qasm.parse(PROG)
    edges = qasm.edges_from_qasm(p)
    _check_slices_against_statevector(p, edges, qasm.qasm_to_tpo(p, edges))


def test_opt_tpo_by_svd():
    p = qasm.parse(PROG)
    edges = qasm.edges_from_qasm(p)
    slices = qasm.qasm_to_tpo(p, edges)
    first = slices[0]
    # cx . u1 . cx stacks t = 4 on edge (0, 1); the product has operator rank 2
    assert first[0].shape[0] == 4
    opt = circuits.opt_tpo_by_svd(edges, first, 1e-12)
    assert opt[0].shape[0] == 2 and opt[1].shape[0] == 2
    np.testing.assert_allclose(_full_operator(edges, opt), _full_operator(edges, first), atol=1e-12)
    ref = oc.opt_tpo_by_svd(edges, first, 1e-12)
    assert [o.shape for o in ref] == [o.shape for o in opt]
    np.testing.assert_allclose(_full_operator(edges, ref), _full_operator(edges, first), atol=1e-12)
    _check_slices_against_statevector(p, edges, [circuits.opt_tpo_by_svd(edges, s) for s in slices])


def _embed(u, qs, n):
    """The n-qubit matrix of gate u on qubits qs (first listed = low bit)."""
    k = len(qs)
    t = np.asarray(u).reshape([2] * (2 * k))        # [out_{k-1}..out_0, in_{k-1}..in_0]
    full = np.eye(2 ** n, dtype=complex).reshape([2] * (2 * n))
    res = np.tensordot(t, full, axes=([2 * k - 1 - i for i in range(k)], [n - 1 - q for q in qs]))
    res = np.moveaxis(res, list(range(k)), [n - 1 - qs[k - 1 - j] for j in range(k)])
    return res.reshape(2 ** n, 2 ** n)


def test_three_qubit_and_path_gates_build_the_exact_operator():
    """CCX / CSWAP on qubits that are not neighbours, routed along the
    shortest lattice path through the middle qubit (qasmtotpo.h:439-698), and
    a two-qubit gate between non-neighbours (qasmtotpo.h:333-436): the slice's
    operator equals the dense product of the gates, in order, on a chain and
    on a 2x3 lattice."""
    rng = np.random.default_rng(8)
    u3 = lambda: qasm.gate_matrix("u3", list(rng.uniform(0, 6, 3)))
    for edges, gates in (
        (circuits.chain(5), [(u3(), (0,)), (qasm.gate_matrix("ccx"), (0, 2, 4)), (u3(), (2,)),
                             (qasm.gate_matrix("cswap"), (3, 1, 2)), (qasm.gate_matrix("cx"), (4, 1))]),
        (circuits.square_lattice(3, 2), [(u3(), (4,)), (qasm.gate_matrix("cswap"), (0, 5, 2)),
                                         (qasm.gate_matrix("rzz", [0.7]), (1, 5)),
                                         (qasm.gate_matrix("ccx"), (3, 4, 1))]),
    ):
        n = len({s for e in edges for s in e})
        want = np.eye(2 ** n, dtype=complex)
        for u, q in gates:
            want = _embed(u, q, n) @ want
        O = circuits.layer_tpo(edges, gates)
        G = _full_operator(edges, O)                      # (in_0..in_{n-1}, out_0..out_{n-1})
        got = np.transpose(G, list(range(2 * n - 1, n - 1, -1)) + list(range(n - 1, -1, -1)))
        np.testing.assert_allclose(got.reshape(2 ** n, 2 ** n), want, atol=1e-13)


def test_three_qubit_gate_matrices():
    """CCX is the reference's matrix (qasmtotpo.h:117-126); CSWAP is the
    unitary that swaps the last two listed qubits when the first is 1 (the
    reference's table, qasmtotpo.h:127-136, puts its m[5][5] where m[4][4]
    belongs)."""
    ccx = qasm.gate_matrix("ccx")
    ref = np.eye(8)
    ref[3, 3] = ref[7, 7] = 0
    ref[3, 7] = ref[7, 3] = 1
    np.testing.assert_array_equal(ccx, ref)
    cs = qasm.gate_matrix("cswap")
    np.testing.assert_allclose(cs @ cs.conj().T, np.eye(8), atol=0)
    for i in range(8):
        c, a, b = i & 1, (i >> 1) & 1, (i >> 2) & 1
        j = (c | (b << 1) | (a << 2)) if c else i
        assert cs[j, i] == 1
    p = qasm.parse("OPENQASM 3.0;\nqubit[3] q;\ntoffoli q[0], q[1], q[2];\ncswap q[2], q[0], q[1];\n")
    assert [(i.op, i.qubits) for i in p.instructions] == [("ccx", [0, 1, 2]), ("cswap", [2, 0, 1])]
