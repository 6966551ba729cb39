"""The estimator driver end to end on the GPU: QASM text -> lattice edges ->
TPO slices (with and without OptTPObySVD) -> BP PEPS -> sparse-Pauli
observable.  The circuit's lattice is a tree (a chain with a branch), where
BP is exact, so every term must equal the brute-force statevector to 1e-10."""
import math

import numpy as np
import pytest

from gpu_util import announce
from statevector import StateVector

pytestmark = pytest.mark.gpu

QASM = """OPENQASM 2.0;
include "qelib1.inc";
qreg q[6];
u(9*pi/10,-pi/2,pi/2) q[0];
u(0.4,0.1,-0.3) q[1];
cx q[0],q[1];
u1(pi/2) q[1];
cx q[0],q[1];
h q[2];
rzz(0.7) q[1],q[2];
barrier q;
ry(0.3) q[3];
cx q[2],q[3];
rzz(-1.1) q[3],q[4];
u(1.3,0.2,0.5) q[4];
cz q[2],q[5];
barrier q;
rx(0.9) q[5];
swap q[3],q[4];
t q[0];
rzz(0.25) q[0],q[1];
"""

TERMS = """1.0 IIIIIZ
0.5 0.25 IIIIXY
-0.7 IIIZIZ
1.0 IIYXZI
0.3 IZXZYI
2.0 ZIIIIX
0.1 IIIIII
"""


def _statevector():
    from tnbp_amd import qasm
    prog = qasm.parse(QASM)
    sv = StateVector(6)
    for ins in prog.instructions:
        if ins.op == "barrier":
            continue
        u = qasm.gate_matrix(ins.op, ins.params)
        sv.apply1(u, ins.qubits[0]) if len(ins.qubits) == 1 else sv.apply2(u, *ins.qubits)
    return sv


def _expect(sv, label):
    from tnbp_amd import pauli
    phi = StateVector(sv.n)
    phi.psi = sv.psi.copy()
    for q, o in pauli.local_ops(label):
        phi.apply1(pauli.PAULI[o], q)
    return complex(np.vdot(sv.psi, phi.psi))


@pytest.mark.parametrize("opt", [False, True])
def test_estimator_exact_on_tree(opt):
    from tnbp_amd import pauli, serving
    from tnbp_amd.estimator import job_from_qasm
    from tnbp_amd.tci import Context
    announce(f"estimator qasm tree opt_tpo={opt}")
    terms = pauli.parse_sparse_pauli(TERMS)
    job = job_from_qasm(QASM, terms, max_bond_dim=64, bp_tolerance=1e-13, sv_min=1e-13,
                        max_bp_iterations=200, do_opt_tpo=opt, eps_opt_tpo=1e-13)
    assert job.edges == [(0, 1), (1, 2), (2, 3), (2, 5), (3, 4)]
    ctx = Context(0)
    try:
        res = serving.run_job(ctx, job)
    finally:
        ctx.close()
    sv = _statevector()
    want = [_expect(sv, label) for _, label in terms]
    for (c, label), g, w in zip(terms, res["values"], want):
        assert abs(g - w) < 1e-10, (label, g, w)
    total = sum(c * w for (c, _), w in zip(terms, want))
    assert abs(res["value"] - total) < 1e-9
