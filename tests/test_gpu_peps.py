"""Fused PEPS layer on the GPU against the oracle and its golden fixtures.

Tolerances (stated per test): the plain FP64 path must match the oracle to
1e-10 relative per expectation value (floor 1 for values near zero); the mixed
path to 1e-5.
"""
import ctypes
import json
import math
import os

import numpy as np
import pytest

from gpu_util import announce, rel_close
from oracle import circuits as oc
from oracle import tnbp as ot

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def ctx():
    from tnbp_amd.tci import Context
    c = Context(0)
    yield c
    c.close()


def _golden(name):
    return json.load(open(os.path.join(GOLDEN, name)))


def _run_gpu(ctx, args, precision=0):
    from tnbp_amd import driver
    a = dict(args)
    keep_zero = a.pop("two_site_svd", "trunc") == "full"
    return driver.run_kicked(ctx, two_site_keep_zero=keep_zero, precision=precision, **a)


@pytest.mark.parametrize("name", ["config1_kicked_4x4_chi4.json", "config1_fixed3_kicked_4x4_chi4.json",
                                  "config1_t4_kicked_4x4_chi4.json"])
def test_config1_fp64_matches_golden(ctx, name):
    announce(f"config1 fp64 {name}")
    g = _golden(name)
    r = _run_gpu(ctx, g["args"])
    for a, b in zip(r["steps"], g["steps"]):
        assert a["bp_iterations"] == b["bp_iterations"], (a["step"], a["residuals"], b["residuals"])
        assert a["bond_dims"] == b["bond_dims"]
        for key in ("x", "z", "zz"):
            assert rel_close(a[key], b[key], 1e-10) <= 1.0, (key, a["step"])
        np.testing.assert_allclose(a["trunc_errs"], b["trunc_errs"], rtol=1e-6, atol=1e-14)
        # residuals are gauge invariant; compare where they are above rounding
        for ra, rb in zip(a["residuals"], b["residuals"]):
            if rb > 1e-12:
                assert abs(ra - rb) <= 1e-6 * rb


@pytest.mark.parametrize("name", ["config1_kicked_4x4_chi4.json", "config1_fixed3_kicked_4x4_chi4.json"])
def test_config1_plain_mode_matches_golden(ctx, name):
    """PLAIN mode (the reference's op sequence one call at a time: a copy of V
    per message, a norm and a scale after every bond, a conjugated copy of V,
    a separate condition sweep, bp.h:128-175, :290-350) is the plain FP64 GPU
    path the mixed path's speed is compared with: iteration counts, bond
    dims, <X>, <Z>, <ZZ> equal the golden fixture (1e-10 relative, floor 1),
    and so do its residuals, which keep the reference's phase-sensitive form."""
    from tnbp_amd.peps import HBP_MODE_PLAIN
    announce(f"config1 plain mode {name}")
    g = _golden(name)
    a = dict(g["args"])
    a.pop("two_site_svd", None)
    from tnbp_amd import driver
    r = driver.run_kicked(ctx, mode=HBP_MODE_PLAIN, **a)
    for x, y in zip(r["steps"], g["steps"]):
        assert x["bp_iterations"] == y["bp_iterations"], (x["step"], x["residuals"], y["residuals"])
        assert x["bond_dims"] == y["bond_dims"]
        for key in ("x", "z", "zz"):
            assert rel_close(x[key], y[key], 1e-10) <= 1.0, (key, x["step"])
        for ra, rb in zip(x["residuals"], y["residuals"]):
            if rb > 1e-12:
                assert abs(ra - rb) <= 1e-6 * rb


def test_config1_site11(ctx):
    g = _golden("config1_kicked_4x4_chi4.json")
    r = _run_gpu(ctx, g["args"])
    want = g["steps"][-1]["z"][5]
    got = r["steps"][-1]["z"][5]
    announce(f"<Z_(1,1)> gpu {got!r} oracle {want!r}")
    assert abs(got - want) <= 1e-10 * max(abs(want), 1.0)


def test_config1_mixed_within_1e5(ctx):
    announce("config1 mixed")
    g = _golden("config1_fixed3_kicked_4x4_chi4.json")
    r = _run_gpu(ctx, g["args"], precision=1)
    for a, b in zip(r["steps"], g["steps"]):
        for key in ("x", "z", "zz"):
            assert rel_close(a[key], b[key], 1e-5) <= 1.0, (key, a["step"])


def test_chain_random_gates_exact(ctx):
    """Open chain (a tree): BP is exact; random U3 breaks every symmetry."""
    from tnbp_amd import circuits as pc
    from tnbp_amd import driver
    from tnbp_amd.peps import Measure, MeasureEdges
    announce("chain random gates")
    rng = np.random.default_rng(21)
    edges = pc.chain(7)
    layers = [pc.random_layer(rng, edges, 7) for _ in range(4)]
    peps, recs = driver.run_layers(ctx, edges, layers, 100, 1e-13, 64, 1e-12, keep=True)
    st, orecs = oc.run_tpo_layers(edges, [oc.layer_tpo(edges, g) for g in layers], 100, 1e-13, 64, 1e-12)
    for op in ("X", "Y", "Z"):
        o = pc.op_tensor(pc.PAULI[op])
        got = Measure(ctx, peps, list(range(7)), [o] * 7)
        want = ot.measure_sites(st, list(range(7)), [o] * 7)
        np.testing.assert_allclose(got, want, atol=1e-10)
    o2 = pc.op_tensor(np.kron(pc.PAULI["Y"], pc.PAULI["X"]))
    got = MeasureEdges(ctx, peps, edges, [o2] * len(edges))
    want = ot.measure_edges(st, edges, [o2] * len(edges))
    np.testing.assert_allclose(got, want, atol=1e-10)
    peps.close()


def test_square_random_gates_fixed_iterations(ctx):
    """3x3 loopy lattice with random gates and truncation, fixed BP iterations."""
    from tnbp_amd import circuits as pc
    from tnbp_amd import driver
    from tnbp_amd.peps import Measure, MeasureEdges
    announce("3x3 random gates")
    rng = np.random.default_rng(5)
    edges = pc.square_lattice(3, 3)
    layers = [pc.random_layer(rng, edges, 9) for _ in range(4)]
    peps, recs = driver.run_layers(ctx, edges, layers, 4, 0.0, 4, 1e-8, keep=True)
    st, orecs = oc.run_tpo_layers(edges, [oc.layer_tpo(edges, g) for g in layers], 4, 0.0, 4, 1e-8)
    for a, b in zip(recs, orecs):
        assert a["bond_dims"] == b["bond_dims"]
        np.testing.assert_allclose(a["residuals"], b["residuals"], rtol=1e-6, atol=1e-13)
    zop = pc.op_tensor(pc.PAULI["Z"])
    got = Measure(ctx, peps, list(range(9)), [zop] * 9)
    want = ot.measure_sites(st, list(range(9)), [zop] * 9)
    assert rel_close(got, want, 1e-10) <= 1.0
    o2 = pc.op_tensor(np.kron(pc.PAULI["Z"], pc.PAULI["X"]))
    got = MeasureEdges(ctx, peps, edges, [o2] * len(edges))
    want = ot.measure_edges(st, edges, [o2] * len(edges))
    assert rel_close(got, want, 1e-10) <= 1.0
    peps.close()


def test_site_and_message_round_trip(ctx):
    from tnbp_amd.peps import InitTensorProductState
    announce("round trip")
    p = InitTensorProductState(ctx, [(0, 1), (1, 2)])
    v = p.get_site(1)
    assert v.shape == (1, 1, 2) and v[0, 0, 0] == 1 and v[0, 0, 1] == 0
    e = p.get_message(3)
    assert e.shape == (1, 1) and e[0, 0] == 1
    with pytest.raises(ValueError):
        p.set_site(1, np.zeros((2, 1, 2)))
    p.close()


def test_absorb_odd_bond_dims_matches_oracle(ctx):
    """AbsorbTPO's kernel (lazy and eager) against the oracle's contraction
    (apply.h:52-66) on a 3x3 lattice with bond dims and TPO bond dims that are
    not powers of two (1..7 and 1..3), so that the index split by
    multiply-high division meets odd divisors: sites bitwise-close (1e-13)."""
    from tnbp_amd import circuits as pc
    from tnbp_amd._lib import check, i32
    from tnbp_amd.peps import HBP_ABSORB_EAGER, HBP_ABSORB_LAZY, AbsorbTPO, Peps
    announce("absorb with odd bond dims vs oracle")
    rng = np.random.default_rng(23)
    edges = pc.square_lattice(3, 3)
    dims = [int(x) for x in rng.choice([1, 2, 3, 5, 6, 7], size=len(edges))]
    ts = [int(x) for x in rng.choice([1, 2, 3], size=len(edges))]
    bonds = [sorted(m for m, e in enumerate(edges) if s in e) for s in range(9)]
    sites = [rng.standard_normal([dims[m] for m in bonds[s]] + [2])
             + 1j * rng.standard_normal([dims[m] for m in bonds[s]] + [2]) for s in range(9)]
    ops = [rng.standard_normal([ts[m] for m in bonds[s]] + [2, 2])
           + 1j * rng.standard_normal([ts[m] for m in bonds[s]] + [2, 2]) for s in range(9)]
    st = ot.init_tensor_product_state(edges, [2] * 9)
    st.V = [v.copy() for v in sites]
    st.E = [np.zeros((dims[m % len(edges)],) * 2, dtype=complex) for m in range(2 * len(edges))]
    ot.absorb_tpo(st, ops)
    for mode in (HBP_ABSORB_LAZY, HBP_ABSORB_EAGER):
        p = Peps(ctx, edges, [2] * 9, 0)
        check(ctx.lib.hbp_peps_set_bond_dims(ctx.h, p.h, i32(dims), i32([2] * 9)))
        for s in range(9):
            p.set_site(s, sites[s])
        p.set_absorb_mode(mode)
        AbsorbTPO(ctx, ops, p)
        for s in range(9):
            got = p.get_site(s)
            assert got.shape == st.V[s].shape, (mode, s)
            np.testing.assert_allclose(got, st.V[s], rtol=0, atol=1e-13 * np.abs(st.V[s]).max())
        p.close()


def test_residual_matches_condition(ctx):
    """BeliefPropagationCondition and the fused sweep from random Hermitian PSD
    messages (far from the fixed point), against the oracle."""
    from tnbp_amd import circuits as pc
    from tnbp_amd.peps import (AbsorbTPO, BeliefPropagation, BeliefPropagationCondition,
                               InitTensorProductState)
    announce("residual vs condition")
    rng = np.random.default_rng(17)
    edges = pc.square_lattice(3, 3)
    tpo = pc.kicked_ising_tpo(edges, 0.125, 0.0, 0.05, math.pi)
    p = InitTensorProductState(ctx, edges)
    AbsorbTPO(ctx, tpo, p)
    AbsorbTPO(ctx, tpo, p)
    st = ot.init_tensor_product_state(edges, [2] * 9)
    otpo = oc.circuit_tpo(edges, 0.125, 0.0, 0.05, math.pi, "trunc")
    ot.absorb_tpo(st, otpo)
    ot.absorb_tpo(st, otpo)
    for m in range(2 * len(edges)):
        d = st.E[m].shape[0]
        a = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
        e = a @ a.conj().T
        st.E[m] = e
        p.set_message(m, e)
    layers = ot.greedy_edge_layering(edges)
    r0_gpu = BeliefPropagationCondition(ctx, p)
    r0_ref = ot.belief_propagation_condition(st)
    assert r0_ref > 1e-3 and abs(r0_gpu - r0_ref) <= 1e-10 * r0_ref
    for _ in range(3):
        BeliefPropagation(ctx, p)
        F = [None] * (2 * len(edges))
        for J in layers:
            ot.belief_propagation(st, J, F)
        st.E = F
        r_gpu = BeliefPropagationCondition(ctx, p)
        r_ref = ot.belief_propagation_condition(st)
        assert abs(r_gpu - r_ref) <= 1e-9 * r_ref + 1e-15
        for m in (0, 5, len(edges) + 3):
            np.testing.assert_allclose(p.get_message(m), st.E[m], atol=1e-12)
    p.close()


def test_bp_iterate_pieces_equal_one_call(ctx):
    """bp_loop one iteration per call gives the same iterations, residuals and
    state as hbp_peps_bp_run (also when stopping at max_bp_iterations)."""
    from tnbp_amd import circuits as pc
    from tnbp_amd.peps import AbsorbTPO, InitTensorProductState, Measure, bp_loop
    announce("bp pieces")
    rng = np.random.default_rng(9)
    edges = pc.square_lattice(3, 3)
    tpo = pc.layer_tpo(edges, pc.random_layer(rng, edges, 9))
    for max_iter, tol in ((50, 1e-10), (3, 0.0)):
        runs = []
        for pieces in (False, True):
            p = InitTensorProductState(ctx, edges)
            for _ in range(3):
                AbsorbTPO(ctx, tpo, p)
            prog = (lambda k, r: None) if pieces else None
            it, res = bp_loop(ctx, p, max_iter, tol, progress=prog)
            z = Measure(ctx, p, list(range(9)), [pc.op_tensor(pc.PAULI["Z"])] * 9)
            runs.append((it, res, z, p.get_message(3)))
            p.close()
        assert runs[0][0] == runs[1][0]
        np.testing.assert_array_equal(runs[0][1], runs[1][1])
        np.testing.assert_array_equal(runs[0][2], runs[1][2])
        np.testing.assert_array_equal(runs[0][3], runs[1][3])


@pytest.mark.parametrize("kernel", ["per-site", "auto"])
def test_lazy_and_eager_absorption_agree(ctx, kernel):
    """Lazy absorption materialises the same absorbed tensors the eager path
    stores, so with the same launches every result is bitwise the same
    (per-site launches); with the default sweep the eager path's sites
    (bonds <= 16) go through the small-site kernel while the lazily absorbed
    ones keep their per-site launches, and the results agree to 1e-12."""
    from tnbp_amd import circuits as pc
    from tnbp_amd.peps import (HBP_ABSORB_EAGER, HBP_ABSORB_LAZY, HBP_SWEEP_AUTO, HBP_SWEEP_PER_SITE,
                               AbsorbTPO, InitTensorProductState, Measure, MeasureEdges, Truncation,
                               bp_loop)
    announce(f"lazy vs eager absorption, {kernel} sweep")
    rng = np.random.default_rng(4)
    edges = pc.square_lattice(3, 4)
    tpos = [pc.layer_tpo(edges, pc.random_layer(rng, edges, 12)) for _ in range(3)]
    out = []
    for mode in (HBP_ABSORB_LAZY, HBP_ABSORB_EAGER):
        p = InitTensorProductState(ctx, edges)
        p.set_absorb_mode(mode)
        p.set_sweep_kernel(HBP_SWEEP_PER_SITE if kernel == "per-site" else HBP_SWEEP_AUTO)
        res = []
        for t in tpos:
            AbsorbTPO(ctx, t, p)
            it, r = bp_loop(ctx, p, 30, 1e-12)
            res.append((it, r[-1]))
            Truncation(ctx, p, 4, 1e-10)
        z = Measure(ctx, p, list(range(12)), [pc.op_tensor(pc.PAULI["Z"])] * 12)
        zz = MeasureEdges(ctx, p, edges, [pc.op_tensor(np.kron(pc.PAULI["X"], pc.PAULI["Z"]))] * len(edges))
        out.append((res, z, zz))
        p.close()
    if kernel == "per-site":
        assert out[0][0] == out[1][0]
        np.testing.assert_array_equal(out[0][1], out[1][1])
        np.testing.assert_array_equal(out[0][2], out[1][2])
    else:
        assert [r[0] for r in out[0][0]] == [r[0] for r in out[1][0]]
        assert rel_close(out[0][1], out[1][1], 1e-12) <= 1.0
        assert rel_close(out[0][2], out[1][2], 1e-12) <= 1.0


@pytest.mark.parametrize("precision", [0, 1])
def test_lazy_and_eager_agree_on_merged_sweeps(ctx, precision):
    """The merged sweep on lazily absorbed sites (each sweep materialises a site
    and moves its target pair to the front in a work buffer) and on eagerly
    stored ones (pair-front copies kept for the BP loop): bitwise the same
    messages, residuals and expectation values on a 4x4 kicked-Ising PEPS
    with bonds of 40."""
    from tnbp_amd import circuits as pc
    from tnbp_amd.peps import (HBP_ABSORB_EAGER, HBP_ABSORB_LAZY, AbsorbTPO, InitTensorProductState,
                               Measure, Truncation, bp_loop)
    announce(f"lazy vs eager absorption, merged sweeps, precision {precision}")
    edges = pc.square_lattice(4, 4)
    tpo = pc.kicked_ising_tpo(edges, 0.125, 0.0, 0.11, math.pi)
    out = []
    for mode in (HBP_ABSORB_LAZY, HBP_ABSORB_EAGER):
        p = InitTensorProductState(ctx, edges, [2] * 16, precision)
        p.set_absorb_mode(mode)
        res = []
        for step in range(8):
            AbsorbTPO(ctx, tpo, p)
            it, r = bp_loop(ctx, p, 5, 0.0)
            res.append(list(r))
            Truncation(ctx, p, 20, 1e-12)
        z = Measure(ctx, p, list(range(16)), [pc.op_tensor(pc.PAULI["Z"])] * 16)
        out.append((res, z, [p.get_message(m) for m in range(2 * len(edges))]))
        p.close()
    assert out[0][0] == out[1][0]
    np.testing.assert_array_equal(out[0][1], out[1][1])
    for a, b in zip(out[0][2], out[1][2]):
        np.testing.assert_array_equal(a, b)


def test_grouped_sweep_is_rejected(ctx):
    """The grouped sweep is switched off at the API and its code removed (its
    first GPU run left the device unusable and the cause was not found):
    asking for it is an error and the PEPS keeps its sweep kernel."""
    from tnbp_amd import circuits as pc
    from tnbp_amd.peps import HBP_SWEEP_GROUPED, InitTensorProductState
    announce("grouped sweep rejected")
    p = InitTensorProductState(ctx, pc.square_lattice(2, 2), [2] * 4, 0)
    with pytest.raises(NotImplementedError, match="grouped sweep is disabled"):
        p.set_sweep_kernel(HBP_SWEEP_GROUPED)
    p.close()


@pytest.mark.skip(reason="grouped sweep disabled and its code removed: its first GPU run left "
                         "the device unusable and the cause was not found")
@pytest.mark.parametrize("precision,tol", [(0, 1e-12), (1, 1e-6)])
def test_grouped_sweep_matches_per_site(ctx, precision, tol):
    """The grouped sweep (one launch per level for all sites) against the
    per-site launches on a 4x4 lattice with random gates (degrees 2-4)."""
    from tnbp_amd import circuits as pc
    from tnbp_amd.peps import (HBP_SWEEP_GROUPED, HBP_SWEEP_PER_SITE, AbsorbTPO,
                               InitTensorProductState, Measure, MeasureEdges, Truncation, bp_loop)
    announce(f"grouped vs per-site sweep precision {precision}")
    rng = np.random.default_rng(13)
    edges = pc.square_lattice(4, 4)
    tpos = [pc.layer_tpo(edges, pc.random_layer(rng, edges, 16)) for _ in range(4)]
    out = []
    for kind in (HBP_SWEEP_PER_SITE, HBP_SWEEP_GROUPED):
        p = InitTensorProductState(ctx, edges, [2] * 16, precision)
        p.set_sweep_kernel(kind)
        res = []
        for t in tpos:
            AbsorbTPO(ctx, t, p)
            it, r = bp_loop(ctx, p, 6, 0.0)
            res.append(r)
            Truncation(ctx, p, 8, 1e-10)
        z = Measure(ctx, p, list(range(16)), [pc.op_tensor(pc.PAULI["Y"])] * 16)
        zz = MeasureEdges(ctx, p, edges, [pc.op_tensor(np.kron(pc.PAULI["Z"], pc.PAULI["X"]))] * len(edges))
        out.append((res, z, zz, p.get_message(5)))
        p.close()
    for ra, rb in zip(out[0][0], out[1][0]):
        np.testing.assert_allclose(ra, rb, rtol=1e-6 if precision else 1e-9, atol=1e-14)
    assert rel_close(out[1][1], out[0][1], tol) <= 1.0
    assert rel_close(out[1][2], out[0][2], tol) <= 1.0
    np.testing.assert_allclose(out[1][3], out[0][3], atol=tol)


@pytest.mark.parametrize("precision,tol,res_rtol,res_atol", [(0, 1e-10, 1e-9, 1e-13),
                                                            (1, 2e-5, 0.0, 5e-8)])
def test_merged_sweep_matches_per_site(ctx, precision, tol, res_rtol, res_atol):
    """Merged product-message kernels (the last mode product of each message
    inside the message kernel, and both messages of a bond pair from one pass
    in the first sweep after an absorption) against separate launches: two
    4x4 kicked-Ising PEPS grown identically with separate launches to chi = 20
    and absorbed once more (bonds of 40, the merged kernels' zero-padded
    case); then the same BP iterations with either sweep.  Equal states keep
    the bond gauge common, so the messages themselves compare."""
    from tnbp_amd import circuits as pc
    from tnbp_amd._lib import check
    from tnbp_amd.peps import (HBP_SWEEP_MERGED, HBP_SWEEP_PER_SITE, AbsorbTPO,
                               InitTensorProductState, Measure, MeasureEdges, Truncation, bp_loop)
    announce(f"merged vs per-site sweep precision {precision}")
    edges = pc.square_lattice(4, 4)
    tpo = pc.kicked_ising_tpo(edges, 0.125, 0.0, 0.11, math.pi)
    zop = pc.op_tensor(pc.PAULI["Z"])
    zzop = pc.op_tensor(np.kron(pc.PAULI["Z"], pc.PAULI["Z"]))
    states = []
    for _ in range(2):
        p = InitTensorProductState(ctx, edges, [2] * 16, precision)
        p.set_sweep_kernel(HBP_SWEEP_PER_SITE)
        for _ in range(6):
            AbsorbTPO(ctx, tpo, p)
            bp_loop(ctx, p, 30, 1e-9)
            Truncation(ctx, p, 20, 1e-12)
        AbsorbTPO(ctx, tpo, p)
        states.append(p)
    assert max(max(states[0].site_shape(s)[:-1]) for s in range(16)) == 40
    np.testing.assert_array_equal(states[0].get_site(5), states[1].get_site(5))
    out = []
    for st, kind in zip(states, (HBP_SWEEP_PER_SITE, HBP_SWEEP_MERGED)):
        st.set_sweep_kernel(kind)
        check(ctx.lib.hbp_ctx_profile(ctx.h, 1))
        it, r = bp_loop(ctx, st, 6, 0.0)
        counts = []
        for k in (0, 1, 2):  # reading a kind also releases its records
            n = ctypes.c_int64()
            check(ctx.lib.hbp_ctx_profile_read(ctx.h, k, ctypes.byref(n), None, None, None))
            counts.append(n.value)
        check(ctx.lib.hbp_ctx_profile(ctx.h, 0))
        msgs = [st.get_message(m) for m in range(2 * len(edges))]
        Truncation(ctx, st, 20, 1e-12)
        z = Measure(ctx, st, list(range(16)), [zop] * 16)
        zz = MeasureEdges(ctx, st, edges, [zzop] * len(edges))
        out.append((r, msgs, z, zz, counts))
        st.close()
    # only the merged sweep runs merged launches (at least one per site and sweep),
    # and its first sweep takes fewer message launches (pairs in one pass)
    assert out[0][4][2] == 0 and out[1][4][2] >= 7 * 4
    assert out[1][4][1] < out[0][4][1]
    # mixed precision may stop one iteration apart at its 1e-6 residual floor;
    # its residuals near the floor carry FP32 rounding (absolute tolerance)
    n = min(len(out[0][0]), len(out[1][0]))
    assert n >= 2
    np.testing.assert_allclose(out[1][0][:n], out[0][0][:n], rtol=res_rtol, atol=res_atol)
    for ma, mb in zip(out[0][1], out[1][1]):
        np.testing.assert_allclose(mb, ma, atol=tol)
    assert rel_close(out[1][2], out[0][2], tol) <= 1.0
    assert rel_close(out[1][3], out[0][3], tol) <= 1.0


def test_pair_cache_follows_site_changes(ctx, tmp_path):
    """The merged sweep keeps each site with its target pair moved to the front
    for the whole BP loop; changing a site must drop that copy.  A PEPS whose
    site 5 is replaced after a BP loop continues exactly (bitwise) like a fresh
    PEPS restored from a checkpoint of the changed state."""
    from tnbp_amd import circuits as pc
    from tnbp_amd.peps import AbsorbTPO, InitTensorProductState, Truncation, bp_loop, load_peps, save_peps
    announce("pair cache invalidation")
    edges = pc.square_lattice(4, 4)
    tpo = pc.kicked_ising_tpo(edges, 0.125, 0.0, 0.11, math.pi)
    q = InitTensorProductState(ctx, edges, [2] * 16, 1)
    for _ in range(6):
        AbsorbTPO(ctx, tpo, q)
        bp_loop(ctx, q, 20, 1e-9)
        Truncation(ctx, q, 20, 1e-12)
    AbsorbTPO(ctx, tpo, q)
    assert max(q.site_shape(5)[:-1]) == 40
    bp_loop(ctx, q, 3, 0.0)                      # builds the pair-front copies
    v = q.get_site(5)
    rng = np.random.default_rng(3)
    q.set_site(5, v + 0.05 * np.abs(v).max() * rng.standard_normal(v.shape))
    path = tmp_path / "changed.hbpp"
    save_peps(q, path)
    r = load_peps(ctx, path)
    out = []
    for st in (q, r):
        it, res = bp_loop(ctx, st, 3, 0.0)
        out.append((res, [st.get_message(m) for m in range(2 * len(edges))]))
        st.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    for a, b in zip(out[0][1], out[1][1]):
        np.testing.assert_array_equal(a, b)


def test_checkpoint_round_trip(ctx, tmp_path):
    """save_peps / load_peps: a restored state continues exactly like the
    original (bitwise equal messages and expectation values)."""
    from tnbp_amd import circuits as pc
    from tnbp_amd.peps import AbsorbTPO, InitTensorProductState, Measure, Truncation, bp_loop, load_peps, save_peps
    announce("checkpoint round trip")
    rng = np.random.default_rng(6)
    edges = pc.square_lattice(3, 3)
    tpos = [pc.layer_tpo(edges, pc.random_layer(rng, edges, 9)) for _ in range(3)]
    p = InitTensorProductState(ctx, edges)
    for t in tpos[:2]:
        AbsorbTPO(ctx, t, p)
        bp_loop(ctx, p, 20, 1e-12)
        Truncation(ctx, p, 4, 1e-10)
    path = tmp_path / "state.hbpp"
    save_peps(p, path)
    q = load_peps(ctx, path)
    for s in range(9):
        np.testing.assert_array_equal(q.get_site(s), p.get_site(s))
    zop = pc.op_tensor(pc.PAULI["Z"])
    outs = []
    for st in (p, q):
        AbsorbTPO(ctx, tpos[2], st)
        it, r = bp_loop(ctx, st, 20, 1e-12)
        Truncation(ctx, st, 4, 1e-10)
        outs.append((it, r, Measure(ctx, st, list(range(9)), [zop] * 9), st.get_message(7)))
    assert outs[0][0] == outs[1][0]
    np.testing.assert_array_equal(outs[0][1], outs[1][1])
    np.testing.assert_array_equal(outs[0][2], outs[1][2])
    np.testing.assert_array_equal(outs[0][3], outs[1][3])
    p.close()
    q.close()
