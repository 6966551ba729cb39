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
