"""The front end against the reference's own code (no GPU).

tests/golden/ref_frontend.json holds the outputs of oracle/_ref/tnbp_ref, a
build of the reference's gqten-free headers (include/tnbp/lattice/square.h,
oned.h; include/tnbp/framework/graph.h GreedyEdgeLayering and
FindShortestPath; include/qasm v2/v3 parsers; include/pauli sparse-Pauli
reader) on seeded inputs (tests/golden/make_ref_frontend.py).  This build's
restatements must reproduce them exactly: the lattices and layerings of the
oracle and of the product (bit-exact, edge order included), shortest paths,
every parsed QASM instruction (gate, qubits; parameters to 1 ulp-level
relative 1e-15) and every Pauli term.  Where the binary is built, the fixture
itself is checked against a fresh run of it.
"""
import json
import os
import subprocess

import numpy as np
import pytest

import sys
sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import make_ref_frontend as mk  # noqa: E402
from oracle import tnbp as ot
from tnbp_amd import circuits, pauli, qasm

GOLD = json.load(open(mk.OUT))
CASES = mk.cases()
# qasm::Op (include/qasm/ir.h:13-16) in this build's names
OPS = ["u3", "u2", "u1", "rx", "ry", "rz", "h", "x", "y", "z", "s", "sdg", "t", "tdg", "id",
       "cx", "cz", "swap", "rzz", "ccx", "cswap", "measure", "barrier", "reset", "custom"]


def _pairs(x):
    return [tuple(e) for e in x]


def test_lattices_match_reference():
    for (lx, ly), want in zip(CASES["square"], GOLD["square"]):
        assert ot.bond_square_lattice(lx, ly) == _pairs(want)
        assert [tuple(e) for e in circuits.square_lattice(lx, ly)] == _pairs(want)
    for (lx, px), want in zip(CASES["oned"], GOLD["oned"]):
        assert ot.bond_oned_lattice(lx, px) == _pairs(want)


def _graphs():
    lat = [GOLD["square"][CASES["square"].index([16, 16])], GOLD["square"][CASES["square"].index([7, 3])]]
    return [_pairs(g) for g in CASES["graphs"]] + [_pairs(g) for g in lat]


def test_layering_matches_reference():
    for g, want in zip(_graphs(), GOLD["layers"]):
        got = ot.greedy_edge_layering(g)
        assert [[tuple(e) for e in layer] for layer in got] == [_pairs(layer) for layer in want]


def test_shortest_paths_match_reference():
    graphs = _graphs()
    for k, (a, b, want) in enumerate(GOLD["paths"]):
        g = graphs[k // 6]
        assert ot.find_shortest_path(g, a, b) == want, (g, a, b)
        if want:
            assert circuits._shortest_path(g, a, b) == want, (g, a, b)
        else:  # not connected: the product refuses a gate it cannot route
            with pytest.raises(ValueError, match="not connected"):
                circuits._shortest_path(g, a, b)
    assert any(not w for _, _, w in GOLD["paths"]) and any(len(w) > 3 for _, _, w in GOLD["paths"])


def test_qasm_parser_matches_reference():
    for text, want in zip(CASES["qasm"], GOLD["qasm"]):
        prog = qasm.parse(text)
        assert [(n, k) for n, k in prog.qregs] == [(n, k) for n, k in want["qregs"]]
        got = prog.instructions
        assert len(got) == len(want["instructions"])
        for g, w in zip(got, want["instructions"]):
            assert g.op == OPS[w["op"]], (g, w)
            assert g.qubits == [i for _, i in w["qubits"]], (g, w)
            np.testing.assert_allclose(g.params, w["params"], rtol=1e-15, atol=0)


@pytest.mark.skipif(not os.path.isdir("/root/reference/apps"), reason="reference sample inputs absent")
@pytest.mark.parametrize("name", ["apps/estimator/kicked_ising.qasm", "apps/qasm_tpo/ising_circuit.qasm"])
def test_qasm_samples_match_reference(name):
    if not os.path.exists(mk.REF):
        pytest.skip("oracle/_ref not built")
    path = os.path.join("/root/reference", name)
    want = mk.run("qasm", path)
    prog = qasm.parse(open(path).read())
    assert len(prog.instructions) == len(want["instructions"])
    for g, w in zip(prog.instructions, want["instructions"]):
        assert g.op == OPS[w["op"]] and g.qubits == [i for _, i in w["qubits"]]
        np.testing.assert_allclose(g.params, w["params"], rtol=1e-15, atol=0)


def test_sparse_pauli_matches_reference():
    for text, want in zip(CASES["pauli"], GOLD["pauli"]):
        terms = pauli.parse_sparse_pauli(text)
        assert len(terms) == len(want)
        for (c, label), w in zip(terms, want):
            assert complex(c) == complex(*w["coeff"])
            assert label.upper() == w["label"].upper()
            assert [(q, op) for q, op in pauli.local_ops(label)] == [(q, o) for q, o in w["ops"]]


@pytest.mark.skipif(not os.path.exists(mk.REF), reason="oracle/_ref not built")
def test_fixture_is_the_reference_output():
    """The committed fixture equals a fresh run of the reference build."""
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        for k, text in enumerate(CASES["qasm"][:3]):
            p = os.path.join(d, f"p{k}.qasm")
            open(p, "w").write(text)
            assert mk.run("qasm", p) == GOLD["qasm"][k]
    for (lx, ly), want in zip(CASES["square"][:5], GOLD["square"][:5]):
        assert mk.run("square", str(lx), str(ly)) == want
