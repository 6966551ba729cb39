"""The reference's own drivers, compiled unchanged against this backend's TCI
header (tnbp_amd/dropin), run on the GPU.

* kicked_ising_exvalue: the drivers' tnbp:: algorithm calls resolve to the
  fused overloads of include/tci/hbp/fused.h (the PEPS layer);
  kicked_ising_exvalue_plain: the same source built with HBP_TCI_NO_FUSED
  runs the reference's op sequence one TCI call at a time.  Both run config 1
  (4x4 kicked Ising, chi = 4) and must print the golden fixture's iteration
  counts and <X>, <Z>, <ZZ> (tolerance 1e-10 relative, floor 1).
* kicked_ising_exvalue_single: the drivers' own USE_SINGLE build
  (complex<float> tensors) on config 1 with BP to 1e-5, within 1e-5 of the
  FP64 fixture.
* estimator (fused and plain): a QASM circuit on a tree (u3 / rx / h / cx /
  rzz layers separated by barriers) and a sparse-Pauli file of weight-1 and
  nearest-neighbour weight-2 terms; on a tree BP is exact, so every printed
  expectation value equals the statevector's (1e-10).

The binaries are built by `make -C tnbp_amd/dropin` wherever the reference
sources are present (build() does it); the tests are skipped when they have
not been built.
"""
import json
import math
import os
import re
import subprocess

import numpy as np
import pytest

from gpu_util import announce
from oracle import circuits as oc
from statevector import StateVector

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tnbp_amd", "dropin", "_bin")


def _bin(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.skip(f"drop-in driver {name} not built (needs the reference sources at build time)")
    return path


def _kicked_values(stdout):
    vals = {}
    for m in re.finditer(r"measurement of (X|Z) at site (\d+) time step (\d+) = (\S+)", stdout):
        vals[(m.group(1), int(m.group(2)), int(m.group(3)))] = float(m.group(4))
    for m in re.finditer(r"measurement of ZZ at sites \((\d+),(\d+)\) time step (\d+) = (\S+)", stdout):
        vals[("ZZ", (int(m.group(1)), int(m.group(2))), int(m.group(3)))] = float(m.group(4))
    return vals


def test_reference_kicked_driver_single_precision():
    """The drivers' own USE_SINGLE build (complex<float> tensors through the
    TCI calls; contractions and decompositions computed in complex128 and
    rounded to single precision) on config 1 with BP to 1e-5 (single
    precision does not reach the fixture's 1e-10): every <X>, <Z>, <ZZ> of
    every step within 1e-5 of the FP64 golden fixture (relative, floor 1)."""
    exe = _bin("kicked_ising_exvalue_single")
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "config1_kicked_4x4_chi4.json")))
    a = g["args"]
    cmd = [exe, "--lattice", "square", "--L", str(a["L"]), "--num_steps", str(a["num_steps"]),
           "--Jz", repr(a["Jz"]), "--hz", repr(a["hz"]), "--hx", repr(a["hx"]), "--dt", repr(a["dt"]),
           "--max_bp_iterations", "200", "--bp_tolerance", "1e-5",
           "--max_bond_dim", str(a["max_bond_dim"]), "--sv_min", repr(a["sv_min"])]
    announce("drop-in single precision " + " ".join(cmd[1:]))
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    vals = _kicked_values(r.stdout)
    worst, n = 0.0, 0
    for s in g["steps"]:
        for op in ("Z", "X"):
            for i, want in enumerate(s[op.lower()]):
                worst = max(worst, abs(vals[(op, i, s["step"])] - want) / max(abs(want), 1.0))
                n += 1
        for (u, v), want in zip(g["edges"], s["zz"]):
            worst = max(worst, abs(vals[("ZZ", (u, v), s["step"])] - want) / max(abs(want), 1.0))
            n += 1
    announce(f"drop-in single precision: {n} values, worst relative error vs FP64 {worst:.3e}")
    assert worst <= 1e-5


@pytest.mark.parametrize("name", ["kicked_ising_exvalue", "kicked_ising_exvalue_plain"])
def test_reference_kicked_driver_config1(name):
    exe = _bin(name)
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "config1_kicked_4x4_chi4.json")))
    a = g["args"]
    cmd = [exe, "--lattice", "square", "--L", str(a["L"]), "--num_steps", str(a["num_steps"]),
           "--Jz", repr(a["Jz"]), "--hz", repr(a["hz"]), "--hx", repr(a["hx"]), "--dt", repr(a["dt"]),
           "--max_bp_iterations", str(a["max_bp_iterations"]), "--bp_tolerance", repr(a["bp_tolerance"]),
           "--max_bond_dim", str(a["max_bond_dim"]), "--sv_min", repr(a["sv_min"])]
    announce(f"drop-in {name} " + " ".join(cmd[1:]))
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    vals = {}
    for m in re.finditer(r"measurement of (X|Z) at site (\d+) time step (\d+) = (\S+)", r.stdout):
        vals[(m.group(1), int(m.group(2)), int(m.group(3)))] = float(m.group(4))
    zz = {}
    for m in re.finditer(r"measurement of ZZ at sites \((\d+),(\d+)\) time step (\d+) = (\S+)", r.stdout):
        zz[(int(m.group(1)), int(m.group(2)), int(m.group(3)))] = float(m.group(4))
    iters = [len(re.findall(rf"error from belief propagation condition at time step {s} =", r.stdout))
             for s in range(a["num_steps"])]
    assert iters == [s["bp_iterations"] for s in g["steps"]]
    worst = 0.0
    for s in g["steps"]:
        for i, want in enumerate(s["z"]):
            got = vals[("Z", i, s["step"])]
            worst = max(worst, abs(got - want) / max(abs(want), 1.0))
        for i, want in enumerate(s["x"]):
            got = vals[("X", i, s["step"])]
            worst = max(worst, abs(got - want) / max(abs(want), 1.0))
        for (u, v), want in zip(g["edges"], s["zz"]):
            got = zz[(u, v, s["step"])]
            worst = max(worst, abs(got - want) / max(abs(want), 1.0))
    announce(f"drop-in {name} worst relative error {worst:.3e}; "
             f"<Z_5> = {vals[('Z', 5, a['num_steps'] - 1)]!r}")
    assert worst <= 1e-10


# a tree: 0-1, 1-2, 1-3, 3-4, 4-5 (site 1 has degree 3)
TREE_GATES = [
    [("u3", (0.7, 0.2, -0.4), (0,)), ("rx", (1.1,), (1,)), ("h", (), (2,)), ("u3", (1.9, -0.6, 0.3), (3,)),
     ("rx", (0.4,), (4,)), ("u3", (2.3, 0.5, 1.2), (5,))],
    [("cx", (), (0, 1)), ("rzz", (0.9,), (3, 4))],
    [("rzz", (1.3,), (1, 2)), ("cx", (5, 4), None), ("u3", (0.8, 1.4, -0.2), (0,))],
    [("cx", (), (1, 3)), ("rx", (0.6,), (2,)), ("u3", (1.2, -1.0, 0.7), (5,))],
]
PAULI_TERMS = ["IIIIIZ", "IIIIXI", "IIIYII", "IIZIII", "IXIIII", "ZIIIII", "IIIIZZ",
               "IIIXXI", "IIZIYI", "ZYIIII", "YXIIII", "IIIIIX"]


def _tree_circuit():
    lines = ["OPENQASM 2.0;", 'include "qelib1.inc";', "qreg q[6];"]
    sv = StateVector(6)
    for li, layer in enumerate(TREE_GATES):
        if li:
            lines.append("barrier q;")
        for name, params, qubits in layer:
            if qubits is None:  # ("cx", (control, target), None)
                qubits, params = params, ()
            args = ",".join(f"q[{q}]" for q in qubits)
            par = "(" + ",".join(repr(float(x)) for x in params) + ")" if params else ""
            lines.append(f"{'u' if name == 'u3' else name}{par} {args};")
            u = oc.gate_matrix(name, params)
            sv.apply1(u, qubits[0]) if len(qubits) == 1 else sv.apply2(u, qubits[0], qubits[1])
    return "\n".join(lines) + "\n", sv


def _expect(sv, label):
    """<psi| prod_q P_q |psi> for a Pauli label (qubit 0 = rightmost character)."""
    mats = {"X": np.array([[0, 1], [1, 0]], complex), "Y": np.array([[0, -1j], [1j, 0]]),
            "Z": np.diag([1.0 + 0j, -1.0])}
    phi = StateVector(sv.n)
    phi.psi = sv.psi.copy()
    for q, c in enumerate(reversed(label)):
        if c != "I":
            phi.apply1(mats[c], q)
    return complex(np.vdot(sv.psi, phi.psi)).real


@pytest.mark.parametrize("name", ["estimator", "estimator_plain"])
def test_reference_estimator_on_tree_matches_statevector(name, tmp_path):
    exe = _bin(name)
    qasm, sv = _tree_circuit()
    (tmp_path / "tree.qasm").write_text(qasm)
    (tmp_path / "tree.txt").write_text("".join(f"1.0 0.0 {t}\n" for t in PAULI_TERMS))
    cmd = [exe, "--circuit", str(tmp_path / "tree.qasm"), "--sparse_pauli", str(tmp_path / "tree.txt"),
           "--max_bond_dim", "64", "--bp_tolerance", "1e-14", "--max_bp_iterations", "200",
           "--sv_min", "1e-14"]
    announce(f"drop-in {name} on a 6-qubit tree")
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300, cwd=tmp_path)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    one = [(int(m.group(1)), float(m.group(2))) for m in re.finditer(
        r"Expectation value of single site operator at site (\d+) (\S+)", r.stdout)]
    two = [((int(m.group(1)), int(m.group(2))), float(m.group(3))) for m in re.finditer(
        r"Expectation value of two site operator at sites \((\d+),(\d+)\) (\S+)", r.stdout)]
    w1 = [t for t in PAULI_TERMS if sum(c != "I" for c in t) == 1]
    w2 = [t for t in PAULI_TERMS if sum(c != "I" for c in t) == 2]
    assert len(one) == len(w1) and len(two) == len(w2), r.stdout[-3000:]
    worst = 0.0
    for (site, got), t in zip(one, w1):
        assert site == [q for q, c in enumerate(reversed(t)) if c != "I"][0]
        worst = max(worst, abs(got - _expect(sv, t)))
    for (pair, got), t in zip(two, w2):
        assert set(pair) == {q for q, c in enumerate(reversed(t)) if c != "I"}
        worst = max(worst, abs(got - _expect(sv, t)))
    announce(f"drop-in {name}: worst |<P> - statevector| = {worst:.3e}")
    assert worst <= 1e-10
