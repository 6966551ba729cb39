"""The checkerboard BP schedule (hbp_peps_set_schedule) reaches the reference's
fixed point in fewer sweeps.

Run to a tight tolerance, both schedules converge to the same messages, so the
expectation values agree with the oracle's (Jacobi, bp.h:65-261) to within
the BP tolerance: 1e-10 relative (floor 1) here with tol 1e-13; on a tree
(where BP is exact) the checkerboard run equals the oracle to 1e-10.  The
iteration count is checked to fall (it roughly halves on a bipartite lattice).
"""
import math

import numpy as np
import pytest

from gpu_util import announce, rel_close
from oracle import circuits as oc
from oracle import tnbp as ot

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from tnbp_amd.tci import Context
    c = Context(0)
    yield c
    c.close()


def _kicked_gpu(ctx, schedule, precision, L, steps, chi, tol, hx):
    from tnbp_amd import circuits as pc
    from tnbp_amd.peps import (AbsorbTPO, InitTensorProductState, Measure, MeasureEdges,
                               Truncation, bp_loop)
    edges = pc.square_lattice(L, L)
    tpo = pc.kicked_ising_tpo(edges, 0.125, 0.0, hx, math.pi)
    p = InitTensorProductState(ctx, edges, [2] * (L * L), precision)
    p.set_schedule(schedule)
    its = []
    for _ in range(steps):
        AbsorbTPO(ctx, tpo, p)
        it, _ = bp_loop(ctx, p, 2000, tol)
        its.append(it)
        Truncation(ctx, p, chi, 1e-10)
    z = [v.real for v in Measure(ctx, p, list(range(L * L)), [pc.op_tensor(pc.PAULI["Z"])] * (L * L))]
    zz = [v.real for v in MeasureEdges(ctx, p, edges, [pc.op_tensor(np.diag([1.0, -1, -1, 1]))] * len(edges))]
    p.close()
    return its, z, zz


def test_checkerboard_matches_oracle_at_tight_tolerance(ctx):
    from tnbp_amd.peps import HBP_SCHEDULE_CHECKERBOARD, HBP_SCHEDULE_JACOBI
    hx = 0.3 / (2 * math.pi)
    announce("checkerboard vs oracle, 4x4 kicked Ising chi=4, tol 1e-13")
    r = oc.run_kicked(4, 5, 0.125, 0.0, hx, 1.0, 2000, 1e-13, 4, 1e-10)
    want = r["steps"][-1]
    its_j, zj, zzj = _kicked_gpu(ctx, HBP_SCHEDULE_JACOBI, 0, 4, 5, 4, 1e-13, hx)
    its_c, zc, zzc = _kicked_gpu(ctx, HBP_SCHEDULE_CHECKERBOARD, 0, 4, 5, 4, 1e-13, hx)
    announce(f"  iterations jacobi {its_j} (oracle {[s['bp_iterations'] for s in r['steps']]}), "
             f"checkerboard {its_c}")
    assert its_j == [s["bp_iterations"] for s in r["steps"]]
    assert sum(its_c) < 0.75 * sum(its_j)
    assert rel_close(zc, want["z"], 1e-10) <= 1.0
    assert rel_close(zzc, want["zz"], 1e-10) <= 1.0


@pytest.mark.parametrize("precision,tol", [(0, 1e-10), (1, 1e-5)])
def test_checkerboard_exact_on_tree(ctx, precision, tol):
    """A random circuit on an open chain: BP is exact there; the checkerboard
    run equals the oracle's Jacobi run (FP64 to 1e-10, mixed to 1e-5)."""
    from tnbp_amd import circuits as pc
    from tnbp_amd.peps import (HBP_SCHEDULE_CHECKERBOARD, AbsorbTPO, InitTensorProductState,
                               Measure, MeasureEdges, Truncation, bp_loop)
    announce(f"checkerboard on a chain, precision {precision}")
    rng = np.random.default_rng(21)
    edges = pc.chain(7)
    layers = [pc.random_layer(rng, edges, 7) for _ in range(4)]
    p = InitTensorProductState(ctx, edges, [2] * 7, precision)
    p.set_schedule(HBP_SCHEDULE_CHECKERBOARD)
    for g in layers:
        AbsorbTPO(ctx, pc.layer_tpo(edges, g), p)
        bp_loop(ctx, p, 200, 1e-13)
        Truncation(ctx, p, 64, 1e-12)
    st, _ = oc.run_tpo_layers(edges, [oc.layer_tpo(edges, g) for g in layers], 200, 1e-13, 64, 1e-12)
    for op in ("X", "Y", "Z"):
        o = pc.op_tensor(pc.PAULI[op])
        got = Measure(ctx, p, list(range(7)), [o] * 7)
        assert rel_close(got, ot.measure_sites(st, list(range(7)), [o] * 7), tol) <= 1.0
    o2 = pc.op_tensor(np.kron(pc.PAULI["Y"], pc.PAULI["X"]))
    got = MeasureEdges(ctx, p, edges, [o2] * len(edges))
    assert rel_close(got, ot.measure_edges(st, edges, [o2] * len(edges)), tol) <= 1.0
    p.close()


def test_checkerboard_on_a_loop_falls_back_to_jacobi(ctx):
    """A triangle is not bipartite: the schedule setting is accepted and the
    sweeps stay Jacobi (same iterations and values as an explicit Jacobi run)."""
    from tnbp_amd import circuits as pc
    from tnbp_amd.peps import (HBP_SCHEDULE_CHECKERBOARD, HBP_SCHEDULE_JACOBI, AbsorbTPO,
                               InitTensorProductState, Measure, bp_loop)
    announce("checkerboard on a triangle")
    rng = np.random.default_rng(3)
    edges = [(0, 1), (1, 2), (0, 2)]
    tpo = pc.layer_tpo(edges, pc.random_layer(rng, edges, 3))
    out = []
    for sch in (HBP_SCHEDULE_JACOBI, HBP_SCHEDULE_CHECKERBOARD):
        p = InitTensorProductState(ctx, edges, [2] * 3, 0)
        p.set_schedule(sch)
        AbsorbTPO(ctx, tpo, p)
        it, res = bp_loop(ctx, p, 100, 1e-12)
        out.append((it, res, Measure(ctx, p, [0, 1, 2], [pc.op_tensor(pc.PAULI["Z"])] * 3)))
        p.close()
    assert out[0][0] == out[1][0]
    np.testing.assert_array_equal(out[0][1], out[1][1])
    np.testing.assert_array_equal(out[0][2], out[1][2])
