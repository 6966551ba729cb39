"""BP loops of several PEPS in one batch (hbp_peps_bp_run_batched): every
site of every PEPS with bonds <= 16 in one persistent launch per iteration,
replayed from a HIP graph.

* Config 1 (4x4 kicked Ising, chi = 4) with its BP loops run through the
  batch: iteration counts and <X>, <Z>, <ZZ> equal the golden fixture (1e-10
  relative, floor 1).
* A batch of differently shaped lattices (3x3, 4x4, a chain, a tree) with
  random circuits: each PEPS gets the iteration count of bp_loop alone and
  residuals, messages and <Z> within 1e-12 (FP64; different kernels, so not
  bitwise); a PEPS run in the batch alone equals itself run in the mix bit
  for bit; in mixed precision within 1e-5 of FP64.
* PEPS above the small-bond limit (bonds of 32 and 64, both precisions) run
  their own fused sweeps recorded into the same per-iteration graph: every
  one gets bp_loop's iteration count, residuals and messages bit for bit,
  also while others leave the batch and the graph is recorded again.
* PEPS the batch does not take (the checkerboard schedule, mixing) run on
  their own and still get bp_loop's results.
"""
import json
import math
import os

import numpy as np
import pytest

from gpu_util import announce, rel_close

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def ctx():
    from tnbp_amd.tci import Context
    c = Context(0)
    yield c
    c.close()


def test_config1_through_the_batch_matches_golden(ctx):
    from tnbp_amd import circuits as pc
    from tnbp_amd.peps import (AbsorbTPO, InitTensorProductState, Measure, MeasureEdges, Truncation,
                               bp_loop_batched)
    g = json.load(open(os.path.join(GOLDEN, "config1_kicked_4x4_chi4.json")))
    a = g["args"]
    announce("config 1 through the batched BP loop")
    edges = pc.square_lattice(4, 4)
    tpo = pc.kicked_ising_tpo(edges, a["Jz"], a["hz"], a["hx"], a["dt"] * math.pi)
    p = InitTensorProductState(ctx, edges)
    for step, want in enumerate(g["steps"]):
        AbsorbTPO(ctx, tpo, p)
        (it, res, conv, bat), = bp_loop_batched(ctx, [p], a["max_bp_iterations"], a["bp_tolerance"])
        assert bat and it == want["bp_iterations"], (step, it, res, want["residuals"])
        dims, _ = Truncation(ctx, p, a["max_bond_dim"], a["sv_min"])
        assert dims == want["bond_dims"]
        z = [v.real for v in Measure(ctx, p, list(range(16)), [pc.op_tensor(pc.PAULI["Z"])] * 16)]
        x = [v.real for v in Measure(ctx, p, list(range(16)), [pc.op_tensor(pc.PAULI["X"])] * 16)]
        zz = [v.real for v in MeasureEdges(ctx, p, edges, [pc.op_tensor(np.diag([1.0, -1, -1, 1]))] * len(edges))]
        assert rel_close(z, want["z"], 1e-10) <= 1.0
        assert rel_close(x, want["x"], 1e-10) <= 1.0
        assert rel_close(zz, want["zz"], 1e-10) <= 1.0
    p.close()


def _lattices():
    from tnbp_amd import circuits as pc
    return [pc.square_lattice(3, 3), pc.square_lattice(4, 4), pc.chain(6),
            [(0, 1), (1, 2), (1, 3), (3, 4), (5, 4), (4, 6)]]


def _layer(rng, edges, n):
    """RZZ(theta ~ U[0.3, 1.3]) on every edge, then random U3 on every site:
    loops of the lattice carry correlations, so BP needs several iterations."""
    from tnbp_amd import circuits as pc
    gates = [(pc.rzz(rng.uniform(0.3, 1.3)), e) for e in edges]
    gates += [(pc.u3(*rng.uniform(0, 2 * math.pi, 3)), (s,)) for s in range(n)]
    return gates


def _prepared(ctx, precision, seed=17, layers=6, chi=8):
    """PEPS with random circuits absorbed (bonds <= 16), one BP-ready state each.
    The first layer leaves |0...0> a product state (RZZ only adds a phase), the
    next ones double the bonds: 2, 4, 8, then 16 cut to 8 (a lossy truncation,
    after which BP on the loopy lattices needs several iterations), and the
    last layer absorbs to bonds of 16."""
    from tnbp_amd import circuits as pc
    from tnbp_amd.peps import AbsorbTPO, InitTensorProductState, Truncation, bp_loop
    rng = np.random.default_rng(seed)
    out = []
    for edges in _lattices():
        n = len({s for e in edges for s in e})
        p = InitTensorProductState(ctx, edges, [2] * n, precision)
        tpos = [pc.layer_tpo(edges, _layer(rng, edges, n)) for _ in range(layers)]
        for t in tpos[:-1]:
            AbsorbTPO(ctx, t, p)
            bp_loop(ctx, p, 100, 1e-12)
            Truncation(ctx, p, chi, 1e-12)
        AbsorbTPO(ctx, tpos[-1], p)
        out.append((edges, p))
    return out


def _state(p):
    return [p.get_message(m) for m in range(2 * p.n_edges)]


def test_batch_equals_each_alone(ctx):
    from tnbp_amd import circuits as pc
    from tnbp_amd.peps import Measure, bp_loop, bp_loop_batched
    announce("batch of four lattices vs bp_loop alone, FP64")
    tol = 1e-11
    ref = []
    for edges, p in _prepared(ctx, 0):
        it, res = bp_loop(ctx, p, 200, tol)
        n = len({s for e in edges for s in e})
        assert max(max(p.site_shape(x)[:-1]) for x in range(n)) <= 16
        ref.append((it, res, _state(p), Measure(ctx, p, list(range(n)), [pc.op_tensor(pc.PAULI["Z"])] * n)))
        p.close()
    got = []
    peps = _prepared(ctx, 0)
    out = bp_loop_batched(ctx, [p for _, p in peps], 200, tol)
    for (edges, p), (it, res, conv, bat) in zip(peps, out):
        n = len({s for e in edges for s in e})
        got.append((it, res, _state(p), Measure(ctx, p, list(range(n)), [pc.op_tensor(pc.PAULI["Z"])] * n), bat))
    assert max(r[0] for r in ref) >= 5  # loopy lattices: real BP loops
    assert max(max(p.site_shape(x)[:-1]) for x in range(9) for p in [peps[0][1]]) == 16
    for (it0, r0, m0, z0), (it1, r1, m1, z1, bat) in zip(ref, got):
        announce(f"  iterations alone {it0}, in the batch {it1}")
        assert bat and it0 == it1
        np.testing.assert_allclose(r1, r0, rtol=1e-8, atol=1e-14)
        for a, b in zip(m1, m0):
            np.testing.assert_allclose(a, b, atol=1e-12)
        assert rel_close(z1, z0, 1e-12) <= 1.0
    # the first PEPS alone in a batch equals itself in the mix, bit for bit
    solo = _prepared(ctx, 0)
    (it, res, conv, bat), = bp_loop_batched(ctx, [solo[0][1]], 200, tol)
    assert it == got[0][0] and res == got[0][1]
    for a, b in zip(_state(solo[0][1]), got[0][2]):
        np.testing.assert_array_equal(a, b)
    for _, p in peps + solo:
        p.close()


def test_mixed_batch_within_1e5_of_fp64(ctx):
    from tnbp_amd import circuits as pc
    from tnbp_amd.peps import Measure, bp_loop_batched
    announce("mixed-precision batch vs FP64 batch")
    vals = []
    for prec in (0, 1):
        peps = _prepared(ctx, prec)
        out = bp_loop_batched(ctx, [p for _, p in peps], 200, 1e-12)
        assert all(b for *_, b in out)
        v = []
        for edges, p in peps:
            n = len({s for e in edges for s in e})
            v += [x.real for x in Measure(ctx, p, list(range(n)), [pc.op_tensor(pc.PAULI["Z"])] * n)]
            p.close()
        vals.append(v)
    assert rel_close(vals[1], vals[0], 1e-5) <= 1.0


def _kicked_big(ctx):
    """A 3x3 kicked-Ising PEPS absorbed to bonds of 32 (above the batch's 16)."""
    from tnbp_amd import circuits as pc
    from tnbp_amd.peps import AbsorbTPO, InitTensorProductState, Truncation, bp_loop
    edges = pc.square_lattice(3, 3)
    tpo = pc.kicked_ising_tpo(edges, 0.125, 0.0, 0.11, math.pi)
    p = InitTensorProductState(ctx, edges)
    for _ in range(6):
        AbsorbTPO(ctx, tpo, p)
        bp_loop(ctx, p, 100, 1e-12)
        Truncation(ctx, p, 16, 1e-12)
    AbsorbTPO(ctx, tpo, p)
    assert max(max(p.site_shape(s)[:-1]) for s in range(9)) > 16
    return p


def test_ineligible_peps_run_on_their_own(ctx):
    from tnbp_amd.peps import (HBP_BATCH_ALONE, HBP_BATCH_GRAPH, HBP_BATCH_SMALL,
                               HBP_SCHEDULE_CHECKERBOARD, bp_loop, bp_loop_batched)
    announce("batch with PEPS it does not take")
    tol = 1e-11
    big, big_ref = _kicked_big(ctx), _kicked_big(ctx)
    c = _prepared(ctx, 0, seed=5)
    c[0][1].set_schedule(HBP_SCHEDULE_CHECKERBOARD)
    c[2][1].set_accel(4)
    d = _prepared(ctx, 0, seed=5)
    d[0][1].set_schedule(HBP_SCHEDULE_CHECKERBOARD)
    d[2][1].set_accel(4)
    out = bp_loop_batched(ctx, [big, c[0][1], c[1][1], c[2][1]], 200, tol)
    assert [bat for *_, bat in out] == [HBP_BATCH_GRAPH, HBP_BATCH_ALONE, HBP_BATCH_SMALL,
                                        HBP_BATCH_ALONE]
    for (it, res, conv, bat), p, q in zip([out[0], out[1], out[3]], [big, c[0][1], c[2][1]],
                                         [big_ref, d[0][1], d[2][1]]):
        it0, res0 = bp_loop(ctx, q, 200, tol)
        assert it == it0 and res == res0
        for a, b in zip(_state(p), _state(q)):
            np.testing.assert_array_equal(a, b)
    for p in [big, big_ref] + [p for _, p in c + d]:
        p.close()


def _graph_members(ctx):
    """PEPS above the small-bond limit: the 3x3 kicked-Ising state at bonds
    of 32 (FP64, converges within a few iterations) and seeded random 3x3
    states with every bond at 64 in both precisions (the merged
    product-message and buffer-form mode-product kernels of the 16x16 chi=32
    steady state; random messages, so they iterate to max_iter)."""
    from tnbp_amd.kernel_probe import build_state
    return [_kicked_big(ctx), build_state(ctx, 3, 64, 0, seed=3), build_state(ctx, 3, 64, 1, seed=4)]


def test_graph_members_equal_bp_loop_bitwise(ctx):
    from tnbp_amd.peps import HBP_BATCH_GRAPH, HBP_BATCH_SMALL, bp_loop, bp_loop_batched
    announce("graph-recorded sweeps of bond-32/64 PEPS with small ones vs bp_loop alone")
    tol, max_iter = 1e-11, 30
    mine = _graph_members(ctx)
    small = [p for _, p in _prepared(ctx, 0, seed=9)] + [p for _, p in _prepared(ctx, 1, seed=9)]
    rec0, rep0, drop0 = ctx.graph_stats()
    out = bp_loop_batched(ctx, mine + small, max_iter, tol)
    rec, rep, drop = (a - b for a, b in zip(ctx.graph_stats(), (rec0, rep0, drop0)))
    its = [o[0] for o in out]
    announce(f"  iterations {its}; graphs recorded {rec}, replays {rep}, dropped {drop}")
    assert [b for *_, b in out] == [HBP_BATCH_GRAPH] * 3 + [HBP_BATCH_SMALL] * 8
    # every iteration after the first is a replay; members leaving at different
    # iterations make the loop record again for the rest
    assert drop == 0 and rep == max(its) - 1 and rec >= 2
    assert len(set(its)) > 1
    refs = _graph_members(ctx)
    for (it, res, conv, _), p, q in zip(out[:3], mine, refs):
        it0, res0 = bp_loop(ctx, q, max_iter, tol)
        announce(f"  iterations {it} (alone {it0}), last residual {res[-1]:.3e}")
        assert it == it0 and res == res0
        for a, b in zip(_state(p), _state(q)):
            np.testing.assert_array_equal(a, b)
    for p in mine + refs + small:
        p.close()


def test_truncate_batched_equals_each_alone(ctx):
    """hbp_peps_truncate_batched over PEPS of both precisions and of small
    (<= 16), medium (32) and large (64, 96: SVDs outside LDS) bonds, each with
    its own chi and sv_min: bond dims, truncation errors, site tensors and
    messages bitwise those of Truncation on each PEPS alone."""
    from tnbp_amd.kernel_probe import build_state
    from tnbp_amd.peps import Truncation, TruncationBatched, bp_loop
    announce("batched truncation vs each PEPS alone")

    def states():
        out = [p for _, p in _prepared(ctx, 0, seed=23)] + [p for _, p in _prepared(ctx, 1, seed=23)]
        out.append(_kicked_big(ctx))
        for D, prec, seed in ((64, 0, 5), (64, 1, 6), (96, 0, 7)):
            q = build_state(ctx, 3, D, prec, seed=seed)
            bp_loop(ctx, q, 2, 1e-30)
            out.append(q)
        return out

    chis = [4, 8, 3, 6, 4, 8, 3, 6, 16, 32, 24, 48]
    svs = [1e-10, 1e-12, 1e-8, 0.0] * 3
    a, b = states(), states()
    assert len(a) == len(chis)
    got = TruncationBatched(ctx, a, chis, svs)
    for p, q, c, e, (dims, errs) in zip(a, b, chis, svs, got):
        d0, e0 = Truncation(ctx, q, c, e)
        assert dims == d0 and errs == e0
        for s in range(len(p.sites)):
            np.testing.assert_array_equal(p.get_site(s), q.get_site(s))
        for x, y in zip(_state(p), _state(q)):
            np.testing.assert_array_equal(x, y)
    announce(f"  {len(a)} PEPS, bond dims after: {[max(g[0]) for g in got]}")
    for p in a + b:
        p.close()


def test_absorb_batched_equals_each_alone(ctx):
    """hbp_peps_absorb_tpo_batched over PEPS of both precisions, eager and lazy,
    small (one batched launch) and large sites (a site of 33.5 M elements at
    bonds 64 absorbs in its own launch): every site tensor and message
    bitwise as AbsorbTPO leaves each PEPS alone."""
    from tnbp_amd import circuits as pc
    from tnbp_amd.kernel_probe import build_state
    from tnbp_amd.peps import (HBP_ABSORB_EAGER, HBP_ABSORB_LAZY, AbsorbTPO, AbsorbTPOBatched,
                               InitTensorProductState, Truncation, bp_loop)
    announce("batched AbsorbTPO vs each PEPS alone")
    rng = np.random.default_rng(31)

    def states():
        out, tpos = [], []
        for k, (edges, prec, mode) in enumerate([(pc.square_lattice(3, 3), 0, HBP_ABSORB_EAGER),
                                                 (pc.square_lattice(4, 4), 1, HBP_ABSORB_EAGER),
                                                 (pc.chain(6), 0, HBP_ABSORB_LAZY),
                                                 (pc.square_lattice(3, 4), 1, HBP_ABSORB_LAZY)]):
            n = len({s for e in edges for s in e})
            r = np.random.default_rng(100 + k)
            layers = [pc.layer_tpo(edges, pc.random_layer(r, edges, n)) for _ in range(3)]
            p = InitTensorProductState(ctx, edges, [2] * n, prec)
            p.set_absorb_mode(mode)
            for t in layers[:-1]:
                AbsorbTPO(ctx, t, p)
                bp_loop(ctx, p, 50, 1e-12)
                Truncation(ctx, p, 8, 1e-12)
            out.append(p)
            tpos.append(layers[-1])
        # bonds 32 -> 64: the interior site leaves the batched launch
        q = build_state(ctx, 3, 32, 1, seed=9)
        edges = pc.square_lattice(3, 3)
        out.append(q)
        tpos.append(pc.kicked_ising_tpo(edges, 0.125, 0.0, 0.11, math.pi))
        return out, tpos

    a, ta = states()
    b, tb = states()
    AbsorbTPOBatched(ctx, ta, a)
    for q, t in zip(b, tb):
        AbsorbTPO(ctx, t, q)
    for p, q in zip(a, b):
        for s in range(len(p.sites)):
            np.testing.assert_array_equal(p.get_site(s), q.get_site(s))
        for x, y in zip(_state(p), _state(q)):
            np.testing.assert_array_equal(x, y)
    assert max(a[-1].site_shape(4)[:-1]) == 64
    for p in a + b:
        p.close()
