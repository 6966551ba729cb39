"""Product operators on lattice paths (hbp_peps_measure_paths) and the sparse
Pauli front end on the GPU, against the oracle and, on trees, the statevector.

Tolerances: FP64 1e-10 relative (floor 1) against the oracle and 1e-10
absolute against the statevector; mixed precision 1e-5.
"""
import math

import numpy as np
import pytest

from gpu_util import announce, rel_close
from oracle import circuits as oc
from oracle import tnbp as ot
from statevector import StateVector

pytestmark = pytest.mark.gpu

PATHS_4x4 = [[5], [0, 1], [4, 8], [0, 1, 2], [1, 5, 9, 13], [0, 1, 2, 3, 7, 6],
             [12, 8, 9, 10, 6, 2], [5, 6, 10, 9]]


@pytest.fixture(scope="module")
def ctx():
    from tnbp_amd.tci import Context
    c = Context(0)
    yield c
    c.close()


def _kicked_state(ctx, precision):
    from tnbp_amd import circuits as pc
    from tnbp_amd.peps import AbsorbTPO, InitTensorProductState, Truncation, bp_loop
    edges = pc.square_lattice(4, 4)
    args = (0.125, 0.0, 0.3 / (2 * math.pi), 1.0)
    tpo = pc.kicked_ising_tpo(edges, args[0], args[1], args[2], args[3] * math.pi)
    p = InitTensorProductState(ctx, edges, [2] * 16, precision)
    for _ in range(3):
        AbsorbTPO(ctx, tpo, p)
        bp_loop(ctx, p, 10000, 1e-10)
        Truncation(ctx, p, 4, 1e-8)
    r = oc.run_kicked(4, 3, *args, 10000, 1e-10, 4, 1e-8)
    return edges, p, r["state"]


def _ops(rng, path):
    from tnbp_amd import circuits as pc
    return [pc.op_tensor(pc.PAULI["XYZ"[rng.integers(3)]]) for _ in path]


@pytest.mark.parametrize("precision,tol", [(0, 1e-10), (1, 1e-5)])
def test_paths_match_oracle(ctx, precision, tol):
    from tnbp_amd.peps import MeasurePaths
    announce(f"paths 4x4 precision {precision}")
    rng = np.random.default_rng(3)
    edges, p, st = _kicked_state(ctx, precision)
    ops = [_ops(rng, path) for path in PATHS_4x4]
    got = MeasurePaths(ctx, p, PATHS_4x4, ops)
    want = ot.measure_paths(st, PATHS_4x4, ops)
    assert rel_close(got, want, tol) <= 1.0, (got, want)
    p.close()


def test_paths_reduce_to_measure(ctx):
    from tnbp_amd import circuits as pc
    from tnbp_amd.peps import Measure, MeasureEdges, MeasurePaths
    announce("paths vs Measure")
    edges, p, _ = _kicked_state(ctx, 0)
    x, y = pc.op_tensor(pc.PAULI["X"]), pc.op_tensor(pc.PAULI["Y"])
    got = MeasurePaths(ctx, p, [[s] for s in range(16)], [[y]] * 16)
    want = Measure(ctx, p, list(range(16)), [y] * 16)
    assert rel_close(got, want, 1e-12) <= 1.0
    got = MeasurePaths(ctx, p, [list(e) for e in edges], [[y, x]] * len(edges))
    want = MeasureEdges(ctx, p, edges, [pc.op_tensor(np.kron(pc.PAULI["X"], pc.PAULI["Y"]))] * len(edges))
    assert rel_close(got, want, 1e-12) <= 1.0
    with pytest.raises(ValueError):
        MeasurePaths(ctx, p, [[0, 5]], [[x, x]])  # not neighbours
    p.close()


def test_paths_exact_on_chain(ctx):
    from tnbp_amd import circuits as pc
    from tnbp_amd import driver
    from tnbp_amd.peps import MeasurePaths
    announce("paths chain exact")
    rng = np.random.default_rng(8)
    n = 9
    edges = pc.chain(n)
    layers = [pc.random_layer(rng, edges, n) for _ in range(4)]
    peps, _ = driver.run_layers(ctx, edges, layers, 200, 1e-14, 64, 1e-12, keep=True)
    sv = StateVector(n)
    for gates in layers:
        for u, q in gates:
            sv.apply1(u, q[0]) if len(q) == 1 else sv.apply2(u, q[0], q[1])
    paths = [[0, 1, 2], [2, 3, 4, 5], [3, 4, 5, 6, 7, 8], [8, 7, 6, 5, 4]]
    us = [[pc.PAULI["XYZ"[rng.integers(3)]] for _ in path] for path in paths]
    got = MeasurePaths(ctx, peps, paths, [[pc.op_tensor(u) for u in row] for row in us])
    for g, path, row in zip(got, paths, us):
        phi = StateVector(n)
        phi.psi = sv.psi.copy()
        for q, u in zip(path, row):
            phi.apply1(u, q)
        want = complex(np.vdot(sv.psi, phi.psi))
        assert abs(g - want) < 1e-10, (path, g, want)
    peps.close()


def test_sparse_pauli_estimate(ctx):
    from tnbp_amd import pauli
    announce("sparse pauli estimate")
    edges, p, st = _kicked_state(ctx, 0)
    text = "0.5 0.0 IIIIIIIIIIIIIIIZ\n-1.25 IIIIIIIIIIIIIIZZ\n2.0 0.5 IIIIIIIIIIIIXYZX\n" \
           "0.75 IIIIIIIIIIIIIIII\n1.0 IIZIIIZIIIZIIIZI\n"
    terms = pauli.parse_sparse_pauli(text)
    total, vals = pauli.estimate(ctx, p, edges, terms)
    want = []
    for c, label in terms:
        path, ops = pauli.term_path_ops(edges, label)
        want.append(ot.measure_paths(st, [path], [ops])[0] if path else 1.0)
    assert rel_close(vals, want, 1e-10) <= 1.0
    assert abs(total - sum(c * w for (c, _), w in zip(terms, want))) <= 1e-10 * max(1.0, abs(total))
    p.close()
