"""Anderson mixing of the fused BP loop (hbp_peps_set_accel) reaches the
reference's fixed point in fewer sweeps.

The mixed loop computes G = BP(E) and the residual of E exactly as the plain
loop does and stops on the same test, so at a tight tolerance both end on the
same messages: expectation values agree with the oracle (Jacobi, bp.h:65-261,
main.cc:121-159) to 1e-10 relative (floor 1) with tol 1e-13, on a tree (where
BP is exact) they equal the statevector-pinned oracle, and at the 16x16 chi=32
shapes the mixed-precision path with the checkerboard schedule and mixing
stays within 1e-5 of the FP64 path.  Depth 0 is the plain loop, bit for bit.
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


def _kicked(ctx, L, steps, chi, tol, hx, precision=0, schedule=0, accel=0, max_iter=2000,
            J=0.125, seed_h=None):
    from tnbp_amd import circuits as pc
    from tnbp_amd.peps import (AbsorbTPO, InitTensorProductState, Measure, MeasureEdges,
                               Truncation, bp_loop)
    edges = pc.square_lattice(L, L)
    tpo = pc.kicked_ising_tpo(edges, J, 0.0, hx, math.pi)
    p = InitTensorProductState(ctx, edges, [2] * (L * L), precision)
    p.set_schedule(schedule)
    p.set_accel(accel)
    its = []
    for _ in range(steps):
        AbsorbTPO(ctx, tpo, p)
        it, _ = bp_loop(ctx, p, max_iter, tol)
        its.append(it)
        Truncation(ctx, p, chi, 1e-10)
    z = [v.real for v in Measure(ctx, p, list(range(L * L)), [pc.op_tensor(pc.PAULI["Z"])] * (L * L))]
    zz = [v.real for v in MeasureEdges(ctx, p, edges, [pc.op_tensor(np.diag([1.0, -1, -1, 1]))] * len(edges))]
    p.close()
    return its, np.array(z), np.array(zz)


@pytest.mark.parametrize("schedule", [0, 1])
def test_accel_matches_oracle_at_tight_tolerance(ctx, schedule):
    hx = 0.3 / (2 * math.pi)
    announce(f"mixing (depth 4) vs oracle, 4x4 kicked Ising chi=4, tol 1e-13, schedule {schedule}")
    r = oc.run_kicked(4, 5, 0.125, 0.0, hx, 1.0, 2000, 1e-13, 4, 1e-10)
    want = r["steps"][-1]
    its_p, _, _ = _kicked(ctx, 4, 5, 4, 1e-13, hx, schedule=schedule, accel=0)
    its_a, z, zz = _kicked(ctx, 4, 5, 4, 1e-13, hx, schedule=schedule, accel=4)
    announce(f"  iterations plain {its_p}, mixed {its_a} (oracle {[s['bp_iterations'] for s in r['steps']]})")
    assert sum(its_a) <= sum(its_p)
    assert rel_close(z, want["z"], 1e-10) <= 1.0
    assert rel_close(zz, want["zz"], 1e-10) <= 1.0


def test_accel_depth_zero_is_the_plain_loop(ctx):
    hx = 0.3 / (2 * math.pi)
    announce("mixing depth 0 equals the plain loop bitwise")
    a = _kicked(ctx, 4, 4, 4, 1e-12, hx, accel=0)
    b = _kicked(ctx, 4, 4, 4, 1e-12, hx, accel=0)
    from tnbp_amd import circuits as pc
    from tnbp_amd.peps import AbsorbTPO, InitTensorProductState, Measure, Truncation, bp_loop
    edges = pc.square_lattice(4, 4)
    tpo = pc.kicked_ising_tpo(edges, 0.125, 0.0, hx, math.pi)
    p = InitTensorProductState(ctx, edges, [2] * 16, 0)  # never touched by set_accel
    its = []
    for _ in range(4):
        AbsorbTPO(ctx, tpo, p)
        its.append(bp_loop(ctx, p, 2000, 1e-12)[0])
        Truncation(ctx, p, 4, 1e-10)
    z = np.array([v.real for v in Measure(ctx, p, list(range(16)), [pc.op_tensor(pc.PAULI["Z"])] * 16)])
    p.close()
    assert a[0] == b[0] == its
    np.testing.assert_array_equal(a[1], z)
    np.testing.assert_array_equal(a[1], b[1])


@pytest.mark.parametrize("precision,tol", [(0, 1e-10), (1, 1e-5)])
def test_accel_exact_on_tree(ctx, precision, tol):
    """A random circuit on an open chain: BP is exact there; the mixed loop
    (checkerboard schedule and depth-4 mixing) equals the oracle."""
    from tnbp_amd import circuits as pc
    from tnbp_amd.peps import (AbsorbTPO, InitTensorProductState, Measure, MeasureEdges,
                               Truncation, bp_loop)
    announce(f"mixing on a chain, precision {precision}")
    rng = np.random.default_rng(22)
    edges = pc.chain(7)
    layers = [pc.random_layer(rng, edges, 7) for _ in range(4)]
    p = InitTensorProductState(ctx, edges, [2] * 7, precision)
    p.set_schedule(1)
    p.set_accel(4)
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


def test_accel_same_fixed_point_at_every_depth(ctx):
    """Kicked Ising on a loopy 4x4 lattice past its bond-dimension ramp
    (chi = 4, h = 0.9, nine steps, bonds of 8 during BP), FP64, tol 1e-12:
    the mixed loop at every depth ends within 1e-9 of the plain loop's
    values (both within the tolerance of one fixed point), with no more
    sweeps in total (fewer at depth 4)."""
    hx = 0.9 / (2 * math.pi)
    out = {}
    for depth in (0, 2, 4, 8):
        its, z, zz = _kicked(ctx, 4, 9, 4, 1e-12, hx, accel=depth)
        out[depth] = (its, np.concatenate([z, zz]))
        announce(f"  depth {depth}: iterations {its}")
    for depth in (2, 4, 8):
        assert sum(out[depth][0]) <= sum(out[0][0])
        assert rel_close(out[depth][1], out[0][1], 1e-9) <= 1.0
    assert sum(out[4][0]) < sum(out[0][0])


def test_accel_16x16_chi32_mixed_within_1e5_of_fp64(ctx):
    """The headline path's settings (mixed precision, checkerboard, depth-4
    mixing) on the 16x16 chi=32 job's first ten Floquet steps (the last four
    with BP at D = 64) against the FP64 path with the reference's schedule."""
    rng = np.random.default_rng(2024)
    h = float(rng.uniform(0.1, 1.0))
    hx = h / (2 * math.pi)
    announce("16x16 chi=32, 10 steps: fp64 (Jacobi, no mixing)")
    it64, z64, zz64 = _kicked(ctx, 16, 10, 32, 1e-10, hx, precision=0, max_iter=50)
    announce(f"  fp64 BP iterations {it64}; mixed + checkerboard + mixing")
    itmx, zmx, zzmx = _kicked(ctx, 16, 10, 32, 1e-10, hx, precision=1, schedule=1, accel=4,
                              max_iter=50)
    v64, vmx = np.concatenate([z64, zz64]), np.concatenate([zmx, zzmx])
    d = np.abs(vmx - v64)
    announce(f"  mixed BP iterations {itmx}; 736 values: max |diff| {d.max():.2e}, "
             f"max relative {np.max(d / np.abs(v64)):.2e}")
    assert len(v64) == 736
    assert rel_close(vmx, v64, 1e-5) <= 1.0


def test_accel_rejects_bad_arguments(ctx):
    from tnbp_amd import circuits as pc
    from tnbp_amd.peps import InitTensorProductState
    p = InitTensorProductState(ctx, pc.square_lattice(2, 2), [2] * 4, 0)
    for depth, start in ((-1, 1e-2), (9, 1e-2), (2, 0.0)):
        with pytest.raises(Exception):
            p.set_accel(depth, start)
    p.set_accel(8, 0.5)
    p.close()
