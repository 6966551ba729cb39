"""Sparse Pauli front end (host logic, no GPU): the reference's text format
(include/pauli/sparse_pauli.h:47-105), qubit order (pauli_string.h:27-51) and
the path a term is measured on."""
import numpy as np
import pytest

from tnbp_amd import circuits, pauli


def test_parse_formats_and_comments():
    text = """# comment
// another
1.0 0.0 IIXZ
  -0.5 ZIII

2.0 -1.5 yiiy
"""
    terms = pauli.parse_sparse_pauli(text)
    assert terms == [(1 + 0j, "IIXZ"), (-0.5 + 0j, "ZIII"), (2 - 1.5j, "YIIY")]
    with pytest.raises(ValueError):
        pauli.parse_sparse_pauli("1.0 0.0 XQ")
    with pytest.raises(ValueError):
        pauli.parse_sparse_pauli("abc")


def test_qubit_zero_is_rightmost():
    assert pauli.local_ops("XIZZY") == [(0, "Y"), (1, "Z"), (2, "Z"), (4, "X")]
    assert pauli.local_ops("IIII") == []


def test_pauli_matrices_and_tensor_convention():
    y = pauli.op_tensor(pauli.PAULI["Y"])
    # G(in, out) = U[out][in]: paulitotpo.h:29-30 stores {0, -i, i, 0} row-major
    assert y[0, 1] == 1j and y[1, 0] == -1j
    np.testing.assert_array_equal(pauli.op_tensor(pauli.PAULI["Y"]), circuits.op_tensor(circuits.PAULI["Y"]))


def test_support_paths():
    edges = circuits.square_lattice(4, 4)
    assert pauli.support_path(edges, [2, 0, 1]) == [0, 1, 2]
    p = pauli.support_path(edges, [0, 3])
    assert p[0] == 0 and p[-1] == 3 and len(p) == 4
    # a column run and a corner turn are simple paths
    assert pauli.support_path(edges, [1, 5, 9]) == [1, 5, 9]
    path, ops = pauli.term_path_ops(edges, "X" + "I" * 13 + "ZI")  # Z on 1, X on 15
    assert path[0] == 1 and path[-1] == 15 and len(ops) == len(path)
    np.testing.assert_array_equal(ops[1], np.eye(2))
    with pytest.raises(ValueError):
        pauli.support_path(edges, [1, 4, 6, 9])  # 1-0-4-5-6, then 6-5-9 revisits 5


def test_checkpoint_rejects_foreign_files(tmp_path):
    from tnbp_amd.peps import load_peps
    bad = tmp_path / "x.bin"
    bad.write_bytes(b"NOPE" + bytes(16))
    with pytest.raises(ValueError):
        load_peps(None, bad)
    old = tmp_path / "v0.bin"
    old.write_bytes(b"HBPP" + (0).to_bytes(4, "little") + bytes(12))
    with pytest.raises(ValueError):
        load_peps(None, old)


def test_classify_routes_each_term():
    """classify (the host part of estimate, done by a server at submit):
    weight-1 terms go to the site measurement, weight-2 terms on a lattice
    edge to the edge measurement with U_b (x) U_a (qubit a the low bit,
    paulitotpo.h:62-92), every other term to its support path padded with
    identities; the identity string to none (its value stays 1)."""
    edges = circuits.square_lattice(3, 3)
    n = 9

    def label(ops):
        return "".join(ops.get(n - 1 - i, "I") for i in range(n))

    terms = [(1.0, label({4: "Z"})), (0.5, label({0: "X", 1: "Y"})), (2.0, label({0: "X", 4: "Z"})),
             (1j, label({0: "Z", 1: "Z", 2: "X"})), (3.0, "I" * n)]
    one, two, paths = pauli.classify(edges, terms)
    assert one[0] == [0] and one[1] == [4]
    np.testing.assert_array_equal(one[2][0], pauli.op_tensor(pauli.PAULI["Z"]))
    assert two[0] == [1] and two[1] == [(0, 1)]
    np.testing.assert_array_equal(
        two[2][0], circuits.op_tensor(np.kron(pauli.PAULI["Y"], pauli.PAULI["X"])))
    assert paths[0] == [2, 3]
    p = paths[1][0]  # 0 and 4 are diagonal neighbours: a path through 1 or 3
    assert p[0] == 0 and p[-1] == 4 and len(p) == 3
    np.testing.assert_array_equal(paths[2][0][1], np.eye(2))
    assert paths[1][1] == [0, 1, 2]
    # the same plan, whoever builds it
    one2, two2, paths2 = pauli.classify(edges, terms)
    assert one2[:2] == one[:2] and two2[:2] == two[:2] and paths2[:2] == paths[:2]
